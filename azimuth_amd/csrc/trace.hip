// trace.hip -- batch closest-hit and any-hit kernels over az_ray arrays.
//
// One ray per lane; rays and hits move as 32- and 16-byte records; the BVH is
// walked with a per-lane ring stack of RING entries in LDS (older entries in
// scratch).
#include <hip/hip_runtime.h>

#include "azimuth_gpu.h"
#include "device_trace.h"
#include "kernels.h"

namespace az {

// WAVES > 0 asks the compiler to fit WAVES waves per SIMD (register budget
// 512/WAVES VGPRs); 0 leaves the allocation to the compiler.
template <bool ANY, bool COUNT, int RING, int BLOCK, int WAVES>
__global__ __launch_bounds__(BLOCK, (WAVES > 0 ? WAVES : 1)) void k_trace(DeviceGeometry G, const az_ray *__restrict__ rays, uint32_t n,
                                                az_hit *__restrict__ hits, float eps, int cull,
                                                unsigned long long *__restrict__ counters) {
    __shared__ uint32_t smem[RING * BLOCK];
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const float4 *rp = reinterpret_cast<const float4 *>(rays + i);
    float4 a = rp[0], b = rp[1];
    Ray3 r;
    r.ox = a.x;
    r.oy = a.y;
    r.oz = a.z;
    r.time = a.w;
    r.dx = b.x;
    r.dy = b.y;
    r.dz = b.z;
    HitRec h;
    TraceCounters c{0u, 0u, 0u};
    bool hit = trace_ray<ANY, COUNT, RING, BLOCK>(G, r, b.w, ANY ? 0.0f : eps, ANY ? false : (cull != 0),
                                                  smem + threadIdx.x, h, c);
    uint4 o;
    o.x = hit ? h.prim : AZ_MISS;
    o.y = __float_as_uint(hit ? h.t : 0.0f);
    o.z = __float_as_uint(hit ? h.beta : 0.0f);
    o.w = __float_as_uint(hit ? h.gamma : 0.0f);
    reinterpret_cast<uint4 *>(hits)[i] = o;
    if (COUNT) {
        atomicAdd(counters + 0, (unsigned long long)c.nodes);
        atomicAdd(counters + 1, (unsigned long long)c.tris);
        atomicAdd(counters + 2, (unsigned long long)c.prims);
    }
}

template <bool ANY, bool COUNT, int RING, int BLOCK, int WAVES>
static hipError_t launch_variant(const DeviceGeometry &G, const az_ray *rays, uint32_t n, az_hit *hits, float eps,
                                 int cull, unsigned long long *counters, hipStream_t s) {
    dim3 grid((n + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL((k_trace<ANY, COUNT, RING, BLOCK, WAVES>), grid, dim3(BLOCK), 0, s, G, rays, n, hits, eps,
                       cull, counters);
    return hipGetLastError();
}

// The compiled (ring, block, waves) variants.
#define AZ_TRACE_VARIANTS(X) \
    X(16, 64, 0)             \
    X(16, 64, 6)             \
    X(16, 64, 8)             \
    X(16, 128, 0)            \
    X(16, 128, 6)            \
    X(16, 256, 0)            \
    X(16, 256, 6)            \
    X(32, 64, 0)             \
    X(32, 128, 0)

template <bool ANY, bool COUNT>
static hipError_t launch_ring(const TraceLaunch &L, const DeviceGeometry &G, const az_ray *rays, uint32_t n,
                              az_hit *hits, float eps, int cull, unsigned long long *counters, hipStream_t s) {
#define AZ_CASE(R, B, W)                                                                        \
    if (L.ring == R && L.block == B && L.waves == W)                                            \
        return launch_variant<ANY, COUNT, R, B, W>(G, rays, n, hits, eps, cull, counters, s);
    AZ_TRACE_VARIANTS(AZ_CASE)
#undef AZ_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_trace(const TraceLaunch &L, bool any, bool count, const DeviceGeometry &G, const az_ray *rays,
                        uint32_t n, az_hit *hits, float eps, int cull, unsigned long long *counters,
                        hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (any) {
        return count ? launch_ring<true, true>(L, G, rays, n, hits, eps, cull, counters, s)
                     : launch_ring<true, false>(L, G, rays, n, hits, eps, cull, counters, s);
    }
    return count ? launch_ring<false, true>(L, G, rays, n, hits, eps, cull, counters, s)
                 : launch_ring<false, false>(L, G, rays, n, hits, eps, cull, counters, s);
}

bool trace_variant_exists(int ring, int block, int waves) {
#define AZ_CASE(R, B, W) \
    if (ring == R && block == B && waves == W) return true;
    AZ_TRACE_VARIANTS(AZ_CASE)
#undef AZ_CASE
    return false;
}

int trace_variants(TraceLaunch *out, int cap) {
    int k = 0;
#define AZ_CASE(R, B, W)                 \
    if (k < cap) out[k] = TraceLaunch{R, B, W}; \
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
