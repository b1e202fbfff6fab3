// Stable LSD radix sort of 64-bit keys for gfx950 (8-bit digits).
//
// Per pass: (1) each block histograms its tile of TILE keys in LDS;
// (2) one block turns the digit-major histogram into exclusive offsets;
// (3) each block re-reads its tile in order and ranks every key among the
// keys of its digit: within a wavefront the peers of a digit are found with
// eight ballots (one per digit bit), across the block's wavefronts with a
// per-digit prefix in LDS, so the scatter is stable. Passes whose digit is
// zero for every key are skipped by the caller.
#include "rigc_internal.hpp"

namespace rigc {
namespace {

constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 8;
constexpr int TILE = SORT_THREADS * SORT_ITEMS;
constexpr int WAVES = SORT_THREADS / 64;

__global__ __launch_bounds__(SORT_THREADS) void hist_kernel(const unsigned long long* keys, long long n, int shift,
                                                           unsigned int* hist, int nblocks) {
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * TILE;
    for (int r = 0; r < SORT_ITEMS; r++) {
        const long long i = base + (long long)r * SORT_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of m entries with one block of 1024 threads: each thread
// sums a contiguous chunk, the chunk sums are scanned with wavefront
// shuffles plus one pass over the 16 wavefront totals, and each thread
// rewrites its chunk from its offset.
__global__ __launch_bounds__(1024) void scan_kernel(unsigned int* a, long long m) {
    __shared__ unsigned int wsum[16];
    const long long chunk = (m + 1023) / 1024;
    const long long lo = (long long)threadIdx.x * chunk, hi = lo + chunk < m ? lo + chunk : m;
    unsigned int s = 0;
    for (long long i = lo; i < hi; i++) s += a[i];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned int x = s;  // inclusive prefix within the wavefront
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    unsigned int run = x - s;
    for (int w = 0; w < wave; w++) run += wsum[w];
    for (long long i = lo; i < hi; i++) {
        const unsigned int v = a[i];
        a[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(SORT_THREADS) void scatter_kernel(const unsigned long long* in, unsigned long long* out,
                                                              long long n, int shift, const unsigned int* offs,
                                                              int nblocks) {
    __shared__ unsigned int wcount[WAVES][256];
    __shared__ unsigned int wpref[WAVES][256];
    __shared__ unsigned int dbase[256];
    __shared__ unsigned int gofs[256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    dbase[tid] = 0;
    gofs[tid] = offs[(size_t)tid * nblocks + blockIdx.x];
    for (int w = 0; w < WAVES; w++) wcount[w][tid] = 0;
    __syncthreads();
    const unsigned long long lt_mask = (1ULL << lane) - 1ULL;
    const long long base = (long long)blockIdx.x * TILE;
    for (int r = 0; r < SORT_ITEMS; r++) {
        const long long i = base + (long long)r * SORT_THREADS + tid;
        const bool valid = i < n;
        const unsigned long long key = valid ? in[i] : 0ULL;
        const unsigned int d = (unsigned int)((key >> shift) & 255u);
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long m = __ballot(valid && bit);
            peers &= bit ? m : ~m;
        }
        const unsigned int rank = __popcll(peers & lt_mask);
        if (valid && rank == 0) wcount[wave][d] = __popcll(peers);
        __syncthreads();
        {
            unsigned int run = dbase[tid];
            for (int w = 0; w < WAVES; w++) {
                wpref[w][tid] = run;
                run += wcount[w][tid];
                wcount[w][tid] = 0;
            }
            dbase[tid] = run;
        }
        __syncthreads();
        if (valid) out[gofs[d] + wpref[wave][d] + rank] = key;
        __syncthreads();
    }
}

// Small inputs: one block ranks every key against all others in LDS
// (stable: equal keys keep their order) and scatters in place.
constexpr int RANK_SORT_MAX = 1024;

__global__ __launch_bounds__(RANK_SORT_MAX) void rank_sort_kernel(unsigned long long* keys, int n) {
    __shared__ unsigned long long k[RANK_SORT_MAX];
    const int i = threadIdx.x;
    if (i < n) k[i] = keys[i];
    __syncthreads();
    if (i >= n) return;
    const unsigned long long mine = k[i];
    int rank = 0;
    for (int j = 0; j < n; j++) {
        const unsigned long long o = k[j];
        rank += (o < mine || (o == mine && j < i)) ? 1 : 0;
    }
    keys[rank] = mine;
}

}  // namespace

// Sorts keys[0..n) ascending; tmp must hold n keys. Only the digits in
// [0, bits) are sorted (keys must be < 2^bits). The result ends in keys.
hipError_t radix_sort_u64(unsigned long long* keys, unsigned long long* tmp, unsigned int* hist, long long n,
                          int bits, hipStream_t stream) {
    if (n <= 1) return hipSuccess;
    if (n <= RANK_SORT_MAX) {
        hipLaunchKernelGGL(rank_sort_kernel, dim3(1), dim3(RANK_SORT_MAX), 0, stream, keys, (int)n);
        return hipGetLastError();
    }
    const int nblocks = (int)((n + TILE - 1) / TILE);
    unsigned long long* src = keys;
    unsigned long long* dst = tmp;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(hist_kernel, dim3(nblocks), dim3(SORT_THREADS), 0, stream, src, n, shift, hist, nblocks);
        hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, stream, hist, (long long)256 * nblocks);
        hipLaunchKernelGGL(scatter_kernel, dim3(nblocks), dim3(SORT_THREADS), 0, stream, src, dst, n, shift, hist,
                           nblocks);
        unsigned long long* t = src;
        src = dst;
        dst = t;
    }
    if (src != keys) return hipMemcpyAsync(keys, src, sizeof(unsigned long long) * n, hipMemcpyDeviceToDevice, stream);
    return hipGetLastError();
}

size_t radix_sort_hist_entries(long long n) { return (size_t)256 * (size_t)((n + TILE - 1) / TILE + 1); }

}  // namespace rigc
