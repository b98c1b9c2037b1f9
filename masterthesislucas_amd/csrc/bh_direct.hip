// FP64 direct sum on sampled sinks: the exact-sum oracle of the reference (NBodySolverNaive,
// nBodySolverNaive.h:53-140) run on the GPU, used to report the Barnes-Hut error of every
// result and to label training pairs. A block takes 64 sinks (one per lane); its four waves
// split a chunk of sources, and the source chunks of a sink are summed by a second kernel in
// a fixed order, so the result does not depend on scheduling.
#include <mutex>

#include "bh_common.h"
#include "bh_launch.h"

namespace bh {

template <int KERNEL>
__global__ __launch_bounds__(kBlock) void k_direct_f64(const float4* __restrict__ pos4, uint32_t n,
                                                       const uint32_t* __restrict__ sinks,
                                                       uint32_t ns, double G, double eps2,
                                                       double* __restrict__ part, uint32_t S,
                                                       const uint32_t* __restrict__ sys_off,
                                                       uint32_t nsys) {
    __shared__ double red[3][kWavesPerBlock][kWave];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t s = blockIdx.x * kWave + lane;
    const uint32_t sraw = s < ns ? sinks[s] : 0u;
    const bool valid = s < ns && sraw < n;
    const uint32_t si = valid ? sraw : 0u;
    const float4 me = pos4[si];
    const double px = me.x, py = me.y;
    const uint32_t split = blockIdx.y;
    uint64_t r0 = 0, r1 = n;  // the sink's own system (batched) or all particles
    if (sys_off) {
        uint32_t a = 0, b = nsys;
        while (b - a > 1) {
            const uint32_t mid = (a + b) >> 1;
            if (sys_off[mid] <= si) a = mid; else b = mid;
        }
        r0 = sys_off[a];
        r1 = sys_off[a + 1];
    }
    const uint64_t lo = r0 + (r1 - r0) * split / S, hi = r0 + (r1 - r0) * (split + 1) / S;
    double sx = 0.0, sy = 0.0, st = 0.0;
    for (uint64_t j = lo + (uint64_t)w; j < hi; j += kWavesPerBlock) {
        const float4 q = pos4[j];
        if (j == si) continue;
        const double m = q.z;
        if (KERNEL == kKernelGravity) {
            const double dx = (double)q.x - px, dy = (double)q.y - py;
            const double r2 = dx * dx + dy * dy + eps2;
            const double inv = 1.0 / sqrt(r2);
            const double f = G * m * inv * inv * inv;
            sx += f * dx;
            sy += f * dy;
            st += -G * m * inv;
        } else {
            const double dx = px - (double)q.x, dy = py - (double)q.y;
            const double fd = 1.0 / (1.0 + (dx * dx + dy * dy));
            st += m * fd;
            const double f = m * fd * fd;
            sx += f * dx;
            sy += f * dy;
        }
    }
    red[0][w][lane] = sx;
    red[1][w][lane] = sy;
    red[2][w][lane] = st;
    __syncthreads();
    if (w == 0 && s < ns) {
        double a = 0.0, b = 0.0, c = 0.0;
        for (int k = 0; k < kWavesPerBlock; ++k) {
            a += red[0][k][lane];
            b += red[1][k][lane];
            c += red[2][k][lane];
        }
        part[((uint64_t)split * 3 + 0) * ns + s] = a;
        part[((uint64_t)split * 3 + 1) * ns + s] = b;
        part[((uint64_t)split * 3 + 2) * ns + s] = c;
    }
}

// sums the S partials of each sink in split order; a sink index past the particles (only the
// device-pointer entry point can pass one) gives NaN
__global__ __launch_bounds__(kBlock) void k_direct_sum(const double* __restrict__ part, uint32_t S,
                                                       uint32_t ns,
                                                       const uint32_t* __restrict__ sinks_sorted,
                                                       uint32_t n, double* __restrict__ ox,
                                                       double* __restrict__ oy,
                                                       double* __restrict__ ot) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= ns) return;
    double a = 0.0, b = 0.0, c = 0.0;
    for (uint32_t k = 0; k < S; ++k) {
        a += part[((uint64_t)k * 3 + 0) * ns + s];
        b += part[((uint64_t)k * 3 + 1) * ns + s];
        c += part[((uint64_t)k * 3 + 2) * ns + s];
    }
    if (sinks_sorted[s] >= n) a = b = c = __builtin_nan("");
    if (ox) ox[s] = a;
    if (oy) oy[s] = b;
    if (ot) ot[s] = c;
}

// inverse of the sorted-order id array: where[id[i]] = i
__global__ __launch_bounds__(kBlock) void k_invert(const uint32_t* __restrict__ id, uint32_t n,
                                                   uint32_t* __restrict__ where) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) where[id[i]] = i;
}

// sorted position of each selected caller index; an index >= n maps to kNone
__global__ __launch_bounds__(kBlock) void k_take(const uint32_t* __restrict__ where, uint32_t n,
                                                 const uint32_t* __restrict__ sel, uint32_t ns,
                                                 uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= ns) return;
    const uint32_t q = sel[i];
    out[i] = q < n ? where[q] : kNone;
}

static inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

uint32_t direct_splits(uint32_t n, uint32_t ns) {
    const uint32_t sink_blocks = div_up(ns, kWave);
    uint32_t S = div_up(4096, sink_blocks);            // aim for ~4096 blocks
    S = S < 1 ? 1 : (S > 256 ? 256 : S);
    const uint32_t maxS = div_up(n, 256);              // at least 256 sources per split
    return S > maxS ? (maxS ? maxS : 1) : S;
}

void launch_direct_f64(hipStream_t s, int kernel, const float4* pos4, const uint32_t* id,
                       uint32_t n, const uint32_t* sinks_caller, uint32_t ns, double G, double eps2,
                       uint32_t* where, uint32_t* sinks_sorted, double* part, uint32_t S,
                       double* out_x, double* out_y, double* out_t, const uint32_t* sys_off,
                       uint32_t nsys) {
    k_invert<<<div_up(n, kBlock), kBlock, 0, s>>>(id, n, where);
    k_take<<<div_up(ns, kBlock), kBlock, 0, s>>>(where, n, sinks_caller, ns, sinks_sorted);
    dim3 g(div_up(ns, kWave), S);
    if (kernel == kKernelGravity)
        k_direct_f64<kKernelGravity><<<g, kBlock, 0, s>>>(pos4, n, sinks_sorted, ns, G, eps2, part,
                                                          S, sys_off, nsys);
    else
        k_direct_f64<kKernelTsne><<<g, kBlock, 0, s>>>(pos4, n, sinks_sorted, ns, G, eps2, part, S,
                                                       sys_off, nsys);
    k_direct_sum<<<div_up(ns, kBlock), kBlock, 0, s>>>(part, S, ns, sinks_sorted, n, out_x, out_y,
                                                       out_t);
}

// ------------------------------------------------------------------------------------------
// Allocation-free direct sum (bh_direct_fp64_noalloc_async): the sorted sink indices and the
// per-split partial sums live in this file-scope array, one copy per GPU, held until the
// process ends. With S = direct_splits(n, ns) <= ceil(4096 / ceil(ns / 64)), S ns is at most
// 64 * 4096 + ns, so kDirectStaticSinks sinks need 3 (262,144 + kDirectStaticSinks) doubles of
// partials plus kDirectStaticSinks indices: 7.75 MiB. Every kernel writes each scratch entry
// it later reads within the same call, so nothing needs clearing; calls that overlap on one GPU
// share the array and race.
// ------------------------------------------------------------------------------------------
constexpr uint64_t kStaticPartDoubles = 3ull * (64ull * 4096ull + kDirectStaticSinks);
constexpr uint64_t kStaticWords = kStaticPartDoubles + kDirectStaticSinks / 2;  // 8-byte words
static __device__ __align__(256) uint64_t scratch[kStaticWords];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t direct_static_scratch(int device, void** ptr) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

bool direct_static_fits(uint32_t n, uint32_t ns) {
    if (ns > kDirectStaticSinks) return false;
    return 3ull * direct_splits(n, ns) * ns <= kStaticPartDoubles;
}

void launch_direct_f64_static(hipStream_t s, int kernel, const float4* pos4, const uint32_t* id,
                              uint32_t n, const uint32_t* sinks_caller, uint32_t ns, double G,
                              double eps2, uint32_t* where, void* scratch_ptr, double* out_x,
                              double* out_y, double* out_t, const uint32_t* sys_off,
                              uint32_t nsys) {
    double* part = static_cast<double*>(scratch_ptr);
    uint32_t* sinks_sorted = reinterpret_cast<uint32_t*>(part + kStaticPartDoubles);
    launch_direct_f64(s, kernel, pos4, id, n, sinks_caller, ns, G, eps2, where, sinks_sorted, part,
                      direct_splits(n, ns), out_x, out_y, out_t, sys_off, nsys);
}

void launch_sinks_sorted(hipStream_t s, const uint32_t* id, uint32_t n, const uint32_t* sinks_caller,
                         uint32_t ns, uint32_t* where, uint32_t* sinks_sorted) {
    k_invert<<<div_up(n, kBlock), kBlock, 0, s>>>(id, n, where);
    k_take<<<div_up(ns, kBlock), kBlock, 0, s>>>(where, n, sinks_caller, ns, sinks_sorted);
}

}  // namespace bh
