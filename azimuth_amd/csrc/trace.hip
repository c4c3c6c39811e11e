// trace.hip -- batch closest-hit and any-hit kernels over az_ray arrays.
//
// A wave walks a range of rays, a lane taking the next ray when its current
// one finishes; rays and hits move as 32- and 16-byte records; the BVH is
// walked with a per-lane ring stack of RING entries in LDS (older entries in
// scratch).
#include <hip/hip_runtime.h>

#include "azimuth_gpu.h"
#include "device_trace.h"
#include "kernels.h"

namespace az {

// Rays of an az_ray array, hits into an az_hit array.
struct ArrayPool {
    const az_ray *rays;
    az_hit *hits;
    __device__ inline bool fetch(uint32_t id, Ray3 &r, float &tMax) {
        const float4 *rp = reinterpret_cast<const float4 *>(rays + id);
        float4 a = rp[0], b = rp[1];
        r.ox = a.x;
        r.oy = a.y;
        r.oz = a.z;
        r.time = a.w;
        r.dx = b.x;
        r.dy = b.y;
        r.dz = b.z;
        tMax = b.w;
        return false;
    }
    __device__ inline void finish(uint32_t id, bool found, const HitRec &h) {
        uint4 o;
        o.x = found ? h.prim : AZ_MISS;
        o.y = __float_as_uint(found ? h.t : 0.0f);
        o.z = __float_as_uint(found ? h.beta : 0.0f);
        o.w = __float_as_uint(found ? h.gamma : 0.0f);
        reinterpret_cast<uint4 *>(hits)[id] = o;
    }
};

// Adds a wave's traversal counters with one atomic per counter.
__device__ inline void add_counters(unsigned long long *counters, const TraceCounters &c) {
    unsigned long long v[3] = {c.nodes, c.tris, c.prims};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        unsigned long long x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
        if ((threadIdx.x & 63) == 0) atomicAdd(counters + k, x);
    }
}

// WAVES > 0 asks the compiler to fit WAVES waves per SIMD (register budget
// 512/WAVES VGPRs); 0 leaves the allocation to the compiler.  Each wave
// walks perWave consecutive rays with refill (trace_pool); WW selects the
// while-while loop form.
template <bool ANY, bool COUNT, int RING, int BLOCK, int WAVES, bool WW>
__global__ __launch_bounds__(BLOCK, (WAVES > 0 ? WAVES : 1)) void k_trace(DeviceGeometry G, const az_ray *__restrict__ rays, uint32_t n,
                                                az_hit *__restrict__ hits, float eps, int cull,
                                                unsigned long long *__restrict__ counters, uint32_t perWave) {
    __shared__ uint32_t smem[(RING + 4) * BLOCK];
    const uint32_t wave = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    const uint32_t begin = wave * perWave;
    if (begin >= n) return;                          // whole waves only; no barriers follow
    const uint32_t end = min(begin + perWave, n);
    ArrayPool pool{rays, hits};
    TraceCounters c{0u, 0u, 0u};
    trace_pool<ANY ? 1 : 0, COUNT, RING, BLOCK, WW>(G, pool, begin, end, ANY ? 0.0f : eps,
                                                   ANY ? false : (cull != 0), smem + threadIdx.x, c);
    if (COUNT) add_counters(counters, c);
}

template <bool ANY, bool COUNT, int RING, int BLOCK, int WAVES, bool WW>
static hipError_t launch_variant(const DeviceGeometry &G, const az_ray *rays, uint32_t n, az_hit *hits, float eps,
                                 int cull, unsigned long long *counters, int cuCount, hipStream_t s) {
    const uint32_t per = rays_per_wave(n, cuCount, WAVES, kTraceRaysPerLane);
    const uint32_t perBlock = (BLOCK / 64) * per;
    dim3 grid((n + perBlock - 1) / perBlock);
    hipLaunchKernelGGL((k_trace<ANY, COUNT, RING, BLOCK, WAVES, WW>), grid, dim3(BLOCK), 0, s, G, rays, n, hits, eps,
                       cull, counters, per);
    return hipGetLastError();
}

// The compiled (ring, block, waves, loop form) variants: LDS stack rings of
// 8, 16 and 32 entries per lane (older entries spill to scratch).
#define AZ_TRACE_VARIANTS(X) \
    X(16, 64, 0, 0)          \
    X(16, 64, 6, 0)          \
    X(16, 128, 0, 0)         \
    X(16, 128, 6, 0)         \
    X(16, 256, 0, 0)         \
    X(16, 256, 6, 0)         \
    X(16, 64, 6, 1)          \
    X(16, 128, 0, 1)         \
    X(16, 128, 6, 1)         \
    X(16, 256, 0, 1)         \
    X(16, 256, 6, 1)         \
    X(8, 128, 6, 0)          \
    X(32, 128, 6, 0)         \
    X(8, 128, 6, 1)          \
    X(32, 128, 6, 1)

template <bool ANY, bool COUNT>
static hipError_t launch_ring(const TraceLaunch &L, const DeviceGeometry &G, const az_ray *rays, uint32_t n,
                              az_hit *hits, float eps, int cull, unsigned long long *counters, int cuCount,
                              hipStream_t s) {
#define AZ_CASE(R, B, W, F)                                                                     \
    if (L.ring == R && L.block == B && L.waves == W && L.loop == F)                             \
        return launch_variant<ANY, COUNT, R, B, W, F != 0>(G, rays, n, hits, eps, cull, counters, cuCount, s);
    AZ_TRACE_VARIANTS(AZ_CASE)
#undef AZ_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_trace(const TraceLaunch &L, bool any, bool count, const DeviceGeometry &G, const az_ray *rays,
                        uint32_t n, az_hit *hits, float eps, int cull, unsigned long long *counters, int cuCount,
                        hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (any) {
        return count ? launch_ring<true, true>(L, G, rays, n, hits, eps, cull, counters, cuCount, s)
                     : launch_ring<true, false>(L, G, rays, n, hits, eps, cull, counters, cuCount, s);
    }
    return count ? launch_ring<false, true>(L, G, rays, n, hits, eps, cull, counters, cuCount, s)
                 : launch_ring<false, false>(L, G, rays, n, hits, eps, cull, counters, cuCount, s);
}

bool trace_variant_exists(int ring, int block, int waves, int loop) {
#define AZ_CASE(R, B, W, F) \
    if (ring == R && block == B && waves == W && loop == F) return true;
    AZ_TRACE_VARIANTS(AZ_CASE)
#undef AZ_CASE
    return false;
}

int trace_variants(TraceLaunch *out, int cap) {
    int k = 0;
#define AZ_CASE(R, B, W, F)                           \
    if (k < cap) out[k] = TraceLaunch{R, B, W, F};    \
    ++k;
    AZ_TRACE_VARIANTS(AZ_CASE)
#undef AZ_CASE
    return k < cap ? k : cap;
}

// Plain 16-byte-per-lane copy: the calibration kernel for HBM counters.
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * 256;
    for (; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy16(const void *src, void *dst, size_t bytes, hipStream_t s) {
    size_t n = bytes / 16;
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, s, (const uint4 *)src, (uint4 *)dst, n);
    return hipGetLastError();
}

}  // namespace az
