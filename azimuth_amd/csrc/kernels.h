// kernels.h -- host-side launchers of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "azimuth_gpu.h"
#include "device_scene.h"

namespace az {

// Stack depth per lane and workgroup size of the traversal kernels.  Depth
// must cover the deepest stack the uploaded BVH can need.
struct TraceLaunch {
    int depth = 32;
    int block = 128;
};

bool trace_variant_exists(int depth, int block);

hipError_t launch_trace(const TraceLaunch &L, bool any, bool count, const DeviceGeometry &G, const az_ray *rays,
                        uint32_t n, az_hit *hits, float eps, int cull, unsigned long long *counters,
                        hipStream_t s);

}  // namespace az
