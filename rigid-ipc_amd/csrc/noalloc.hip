// The narrow phase without allocation (rigc_narrow_phase_noalloc): every
// per-call buffer lives in one static device array per GPU, laid out anew by
// each call. It has no counterpart in the reference; it serves callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops). The array is set up once per device (noalloc_scratch, called by
// rigc_create with the device current); host code reaches it only through
// the address that hipGetSymbolAddress returns for that device.
//
// Calls that overlap on one GPU (two contexts or threads at once) share the
// array and race; each call initialises what it reads before its kernels
// run, so calls one after another need no clearing between them.
#include <mutex>

#include "rigc_internal.hpp"

namespace rigc {
namespace {

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
void* g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

__global__ void fill_u64_kernel(unsigned long long* p, long long n, unsigned long long v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// out[0] = bits of the least TOI (per query: over queries with hit 1; else
// over the per-scene running minima), out[1] = 1 if a query is unresolved.
__global__ __launch_bounds__(256) void noalloc_result_kernel(const double* toi, const signed char* hit, long long nq,
                                                             const unsigned long long* scene_best, int n_scenes,
                                                             int per_query, unsigned long long* out) {
    __shared__ unsigned long long s_best[256];
    __shared__ int s_bad[256];
    unsigned long long best = 0x7ff0000000000000ULL;
    int bad = 0;
    if (per_query) {
        for (long long i = threadIdx.x; i < nq; i += blockDim.x) {
            if (hit[i] < 0) bad = 1;
            if (hit[i] == 1) {
                const unsigned long long b = (unsigned long long)__double_as_longlong(toi[i]);
                best = b < best ? b : best;
            }
        }
    } else {
        for (long long i = threadIdx.x; i < nq; i += blockDim.x)
            if (hit[i] < 0) bad = 1;
        for (int k = threadIdx.x; k < n_scenes; k += blockDim.x) best = scene_best[k] < best ? scene_best[k] : best;
    }
    s_best[threadIdx.x] = best;
    s_bad[threadIdx.x] = bad;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const unsigned long long o = s_best[threadIdx.x + w];
            if (o < s_best[threadIdx.x]) s_best[threadIdx.x] = o;
            s_bad[threadIdx.x] |= s_bad[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s_best[0];
        out[1] = (unsigned long long)s_bad[0];
    }
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// The minimum over the scenes' running minima (bit patterns of t >= 0,
// +inf when none) into out[0].
__global__ void enqueue_result_kernel(const unsigned long long* sb, int n_scenes, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned long long m = 0x7ff0000000000000ULL;
    for (int s = 0; s < n_scenes; s++) m = sb[s] < m ? sb[s] : m;
    out[0] = __longlong_as_double((long long)m);
}

}  // namespace

// rigc_narrow_phase_enqueue: kernels only (fill, the one-pass narrow phase,
// the result), so that the sequence can be captured in a graph.
hipError_t enqueue_narrow_phase(CcdArgs a, int grid, unsigned long long* scene_best, int n_scenes, double* out,
                                hipStream_t stream) {
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)((n_scenes + 255) / 256)), dim3(256), 0, stream, scene_best,
                       (long long)n_scenes, 0x7ff0000000000000ULL);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (a.n_ee + a.n_fv > 0 && (e = launch_ccd_single(a, grid, 256, stream)) != hipSuccess) return e;
    hipLaunchKernelGGL(enqueue_result_kernel, dim3(1), dim3(64), 0, stream, scene_best, n_scenes, out);
    return hipGetLastError();
}

hipError_t noalloc_scratch(int device, void** ptr) {
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = p;
    }
    *ptr = g_scratch[device];
    return hipSuccess;
}

NoallocLayout noalloc_layout(long long nb, long long n_ee, long long n_fv, int n_scenes) {
    const size_t nq = (size_t)(n_ee + n_fv);
    NoallocLayout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += align256(bytes);
        return at;
    };
    L.poses0 = take(48 * (size_t)nb);
    L.poses1 = take(48 * (size_t)nb);
    L.ee = take(8 * (size_t)n_ee);
    L.fv = take(8 * (size_t)n_fv);
    L.toi = take(8 * nq);
    L.hit = take(nq);
    L.counters = take(4 * sizeof(unsigned long long));
    L.qbest = take(8 * nq);
    L.qvisits = take(4 * nq);
    L.qflag = take(4 * nq);
    L.scene_best = take(8 * (size_t)n_scenes);
    L.result = take(2 * sizeof(unsigned long long));
    L.total = o;
    return L;
}

hipError_t noalloc_narrow_phase(CcdArgs a, int grid, unsigned char* base, const NoallocLayout& L, int n_scenes,
                                int per_query, hipStream_t stream) {
    const long long nq = a.n_ee + a.n_fv;
    unsigned long long* sb = (unsigned long long*)(base + L.scene_best);
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)((n_scenes + 255) / 256)), dim3(256), 0, stream, sb,
                       (long long)n_scenes, 0x7ff0000000000000ULL);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (nq > 0 && (e = launch_ccd_single(a, grid, 256, stream)) != hipSuccess) return e;
    hipLaunchKernelGGL(noalloc_result_kernel, dim3(1), dim3(256), 0, stream, a.toi_out, a.hit_out, nq, sb, n_scenes,
                       per_query, (unsigned long long*)(base + L.result));
    return hipGetLastError();
}

}  // namespace rigc
