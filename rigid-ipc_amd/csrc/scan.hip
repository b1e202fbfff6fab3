// Exclusive prefix sum of 32-bit counts (three passes: tile scans in LDS,
// one block over the tile totals, tile offsets added back; one launch when
// the input fits one tile). Used for stable compaction of flagged entries.
#include "rigc_internal.hpp"

namespace rigc {
namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ unsigned int block_exclusive_scan(unsigned int v, unsigned int* out_prefix) {
    __shared__ unsigned int wsum[SCAN_THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    unsigned int base = 0, total = 0;
    for (int w = 0; w < SCAN_THREADS / 64; w++) {
        if (w < wave) base += wsum[w];
        total += wsum[w];
    }
    *out_prefix = base + x - v;
    __syncthreads();
    return total;
}

__global__ __launch_bounds__(SCAN_THREADS) void scan_tiles_kernel(const unsigned int* in, unsigned int* out,
                                                                 long long n, unsigned int* tile_sums) {
    const long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
    unsigned int v[SCAN_ITEMS];
    unsigned int s = 0;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    unsigned int pre;
    const unsigned int total = block_exclusive_scan(s, &pre);
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = pre;
        pre += v[k];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// One block: exclusive scan of the tile totals in place; total[0] = sum.
__global__ __launch_bounds__(SCAN_THREADS) void scan_sums_kernel(unsigned int* sums, long long m,
                                                                unsigned int* total) {
    __shared__ unsigned int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (long long b = 0; b < m; b += SCAN_THREADS) {
        const long long i = b + threadIdx.x;
        const unsigned int v = i < m ? sums[i] : 0u;
        unsigned int pre;
        const unsigned int t = block_exclusive_scan(v, &pre);
        if (i < m) sums[i] = carry + pre;
        __syncthreads();
        if (threadIdx.x == 0) carry += t;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

// n <= SCAN_TILE: the whole scan and its total in one launch.
__global__ __launch_bounds__(SCAN_THREADS) void scan_single_tile_kernel(const unsigned int* in, unsigned int* out,
                                                                       long long n, unsigned int* total) {
    const long long base = (long long)threadIdx.x * SCAN_ITEMS;
    unsigned int v[SCAN_ITEMS];
    unsigned int s = 0;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    unsigned int pre;
    const unsigned int t = block_exclusive_scan(s, &pre);
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = pre;
        pre += v[k];
    }
    if (threadIdx.x == 0) *total = t;
}

__global__ __launch_bounds__(SCAN_THREADS) void scan_add_kernel(unsigned int* out, long long n,
                                                               const unsigned int* tile_sums) {
    const long long base = (long long)blockIdx.x * SCAN_TILE;
    const unsigned int add = tile_sums[blockIdx.x];
    for (int k = 0; k < SCAN_ITEMS; k++) {
        const long long i = base + (long long)k * SCAN_THREADS + threadIdx.x;
        if (i < n) out[i] += add;
    }
}

}  // namespace

size_t scan_workspace_entries(long long n) { return (size_t)((n + SCAN_TILE - 1) / SCAN_TILE) + 1; }

// out[i] = sum of in[0..i); *total (device) = sum of all. work needs
// scan_workspace_entries(n) entries. in and out may alias.
void exclusive_scan_u32(const unsigned int* in, unsigned int* out, long long n, unsigned int* work,
                        unsigned int* total, hipStream_t st) {
    const long long tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (tiles <= 1) {
        hipLaunchKernelGGL(scan_single_tile_kernel, dim3(1), dim3(SCAN_THREADS), 0, st, in, out, n, total);
        return;
    }
    if (n > 0)
        hipLaunchKernelGGL(scan_tiles_kernel, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, st, in, out, n, work);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(SCAN_THREADS), 0, st, work, tiles, total);
    if (n > 0) hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, st, out, n, work);
}

}  // namespace rigc
