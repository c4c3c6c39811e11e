// kernels.h -- host-side launchers of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "azimuth_gpu.h"
#include "device_scene.h"

namespace az {

// LDS ring entries per lane and workgroup size of the traversal kernels.  The
// ring plus the scratch spill (kSpill) must cover the deepest stack the
// uploaded BVH can need.
struct TraceLaunch {
    int ring = 16;
    int block = 256;
    int waves = 6;   // minimum waves per SIMD the register allocation targets (0: compiler's choice)
    int loop = 0;    // traversal loop form: 0 one node or leaf per step, 1 while-while
};

bool trace_variant_exists(int ring, int block, int waves, int loop);
// The compiled traversal variants (at most cap of them).
int trace_variants(TraceLaunch *out, int cap);

// Launch parameters of the renderer: camera samples per wavefront chunk (the
// queue tile) and the workgroup size / register target of its traversal
// kernels (extend, connect).  Chosen per scene by az_gpu_autotune_render.
// A chunk of 16M samples gives a level's first traversal launches about ten
// rounds of resident waves (6,144 at 6 waves per SIMD), so the partly filled
// last round costs little.  Its queues take 3 GB of device memory per child
// a shading point can emit (1 to 4) plus 2 GB per light and for the
// environment.
struct RenderTuning {
    uint32_t chunkSamples = 16u << 20;
    int traceBlock = 128;
    int traceWaves = 6;
    int traceLoop = 1;   // while-while
};

// Rays per wave of a traversal launch of n rays: 64 x perLane (each lane
// walks several, refilled from the wave's range) when n fills at least two
// rounds of the device's resident waves, else halved down to 64 (one per
// lane), so that small launches still reach every SIMD.  `waves` is the
// register target in waves per SIMD (0: the compiler's choice, taken as 5).
// The renderer's levels take kRaysPerLane, batch queries kTraceRaysPerLane.
constexpr uint32_t kTraceRaysPerLane = 4;

__host__ __device__ inline uint32_t rays_per_wave(uint32_t n, int cuCount, int waves,
                                                  uint32_t perLane = kRaysPerLane) {
    const uint64_t resident = (uint64_t)cuCount * 4u * (uint64_t)(waves > 0 ? waves : 5);
    uint32_t per = 64u * perLane;
    while (per > 64u && (uint64_t)n < 2u * resident * per) per >>= 1;
    return per;
}

bool render_trace_variant_exists(int block, int waves, int loop);
int render_trace_variants(RenderTuning *out, int cap);   // chunkSamples left at its default

hipError_t launch_trace(const TraceLaunch &L, bool any, bool count, const DeviceGeometry &G, const az_ray *rays,
                        uint32_t n, az_hit *hits, float eps, int cull, unsigned long long *counters, int cuCount,
                        hipStream_t s);

hipError_t launch_copy16(const void *src, void *dst, size_t bytes, hipStream_t s);

}  // namespace az
