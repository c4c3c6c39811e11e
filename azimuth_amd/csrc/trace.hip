// trace.hip -- batch closest-hit and any-hit kernels over az_ray arrays.
//
// One ray per lane; rays and hits move as 32- and 16-byte records; the BVH is
// walked with an LDS stack of DEPTH entries per lane.
#include <hip/hip_runtime.h>

#include "azimuth_gpu.h"
#include "device_trace.h"
#include "kernels.h"

namespace az {

template <bool ANY, bool COUNT, int DEPTH, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace(DeviceGeometry G, const az_ray *__restrict__ rays, uint32_t n,
                                                az_hit *__restrict__ hits, float eps, int cull,
                                                unsigned long long *__restrict__ counters) {
    __shared__ uint32_t smem[DEPTH * BLOCK];
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
    LdsStack<DEPTH, BLOCK> st{smem + threadIdx.x};
    HitRec h;
    TraceCounters c{0u, 0u, 0u};
    bool hit = trace_ray<ANY, COUNT, DEPTH, BLOCK>(G, r, b.w, ANY ? 0.0f : eps, ANY ? false : (cull != 0), st, h, c);
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

template <bool ANY, bool COUNT, int DEPTH, int BLOCK>
static hipError_t launch_variant(const DeviceGeometry &G, const az_ray *rays, uint32_t n, az_hit *hits, float eps,
                                 int cull, unsigned long long *counters, hipStream_t s) {
    dim3 grid((n + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL((k_trace<ANY, COUNT, DEPTH, BLOCK>), grid, dim3(BLOCK), 0, s, G, rays, n, hits, eps, cull,
                       counters);
    return hipGetLastError();
}

template <bool ANY, bool COUNT>
static hipError_t launch_depth(const TraceLaunch &L, const DeviceGeometry &G, const az_ray *rays, uint32_t n,
                               az_hit *hits, float eps, int cull, unsigned long long *counters, hipStream_t s) {
#define AZ_CASE(D, B)                                                                           \
    if (L.depth == D && L.block == B)                                                           \
        return launch_variant<ANY, COUNT, D, B>(G, rays, n, hits, eps, cull, counters, s);
    AZ_CASE(32, 64)
    AZ_CASE(32, 128)
    AZ_CASE(32, 256)
    AZ_CASE(64, 64)
    AZ_CASE(64, 128)
    AZ_CASE(64, 256)
#undef AZ_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_trace(const TraceLaunch &L, bool any, bool count, const DeviceGeometry &G, const az_ray *rays,
                        uint32_t n, az_hit *hits, float eps, int cull, unsigned long long *counters,
                        hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (any) {
        return count ? launch_depth<true, true>(L, G, rays, n, hits, eps, cull, counters, s)
                     : launch_depth<true, false>(L, G, rays, n, hits, eps, cull, counters, s);
    }
    return count ? launch_depth<false, true>(L, G, rays, n, hits, eps, cull, counters, s)
                 : launch_depth<false, false>(L, G, rays, n, hits, eps, cull, counters, s);
}

}  // namespace az

namespace az {
bool trace_variant_exists(int depth, int block) {
    return (depth == 32 || depth == 64) && (block == 64 || block == 128 || block == 256);
}
}  // namespace az
