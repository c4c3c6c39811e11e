// gpu_context.h -- state behind an az_gpu_ctx.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "azimuth_gpu.h"
#include "azimuth_guide.h"
#include "device_render.h"
#include "device_scene.h"
#include "guide.h"
#include "kernels.h"

namespace az {

// A cap on the bytes a group of buffers may hold (0: none).  An allocation
// that would pass it fails with hipErrorOutOfMemory without calling the
// runtime, as a full device would make it fail.
struct AllocLimit {
    size_t limit = 0;
    size_t used = 0;
};

// Owning device allocation.
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    AllocLimit *lim = nullptr;   // not owned; null: uncapped
    // Grows to n plus a quarter, so that queues whose level sizes vary from
    // chunk to chunk settle after a few growths (each one a device-wide
    // synchronising free and a fresh allocation).
    hipError_t alloc(size_t n) {
        if (n <= bytes && ptr) return hipSuccess;
        release();
        n += n / 4;
        if (!n) n = 1;
        if (lim && lim->limit && lim->used + n > lim->limit) return hipErrorOutOfMemory;
        hipError_t e = hipMalloc(&ptr, n);
        if (e == hipSuccess) {
            bytes = n;
            if (lim) lim->used += n;
        } else {
            ptr = nullptr;
        }
        return e;
    }
    void release() {
        if (ptr) {
            (void)hipFree(ptr);
            if (lim) lim->used -= bytes;
        }
        ptr = nullptr;
        bytes = 0;
    }
    ~DeviceBuffer() { release(); }
};

struct GeometryBuffers {
    DeviceBuffer nodes, tris, faceNormals, items, refBoxes;
    void free() {
        nodes.release();
        tris.release();
        faceNormals.release();
        items.release();
        refBoxes.release();
    }
};

// What the geometry upload learned, for the shading upload.
struct UploadInfo {
    std::vector<uint32_t> meshPrimBase, meshBlasRoot, meshTriBase;
    std::vector<std::vector<const entity *>> meshTriEntities;   // per mesh, in primitive-id order
    std::vector<uint32_t> leafOrder;       // per device triangle slot: index within its mesh
    std::vector<const entity *> topEntities;   // per top-level item (device order)
    uint32_t triTotal = 0;
    uint32_t maxBlasStack = 0, maxTlasStack = 0, stackNeeded = 0;
    uint32_t leafCount = 0;
    float boxMin[3] = {0.0f, 0.0f, 0.0f}, boxMax[3] = {1.0f, 1.0f, 1.0f};   // scene bounds
};

}  // namespace az

void az_release_render_buffers(void *buffers);

// Guide calls ordered on a caller's stream (the renderer's).
// Packed lobes (guide_lobes.h) of the queries of non-negative roughness,
// ordered on stream s.
az_status az_guide_infer_on_stream(az_guide *G, const void *d_Queries, u32 N, void *d_Packed, hipStream_t s);
az_status az_guide_train_on_stream(az_guide *G, const void *d_Samples, u32 N, hipStream_t s);
// The device copies of the guide's weights, for kernels that run the network
// themselves (the renderer's fused query and inference).
az::GuideWeights az_guide_device_weights(az_guide *G);

struct az_gpu_ctx {
    int device = 0;
    int cuCount = 256;               // compute units of the device
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, tev0 = nullptr, tev1 = nullptr;
    hipEvent_t lastStart = nullptr, lastStop = nullptr;   // not owned
    std::string lastError;
    bool hasScene = false;
    bool counting = false;
    bool pendingTiming = false;
    uint32_t maxLeafItems = 4;       // BVH build of the next upload: leaf size, node width, SAH costs
    uint32_t bvhWidth = 4;
    uint32_t buildMode = AZ_BUILD_AUTO;   // where the binary BVH of an upload is built
    uint32_t levelMode = AZ_LEVELS_AUTO;  // how a render's bounce levels are launched
    float sahNodeCost = 1.0f, sahItemCost = 1.0f;
    epsilon_values eps{0.0f, 0.000275f};
    az::GeometryBuffers geom;
    az::DeviceGeometry dg{};
    std::vector<az::TopItem> hostItems;
    az::TraceLaunch launch;
    az::RenderTuning renderTuning;
    az::DeviceBuffer counters;
    az::DeviceBuffer outputScratch;  // az_gpu_output_pixels: luminance keys, logs, histogram
    az::RenderState render;
    az::DevScene devScene{};
    void *renderBuffers = nullptr;   // owned; released by az_release_render_buffers
    az_guide *guide = nullptr;       // owned; the learned guide of guided renders (environment NEE)
    az_guide *scatterGuide = nullptr;   // owned; the learned guide of the hemisphere scatter
    az_stats stats{};
    az_render_stats renderStats{};   // the last render's counts (while counting)
    bool recordLaunches = false;
    std::vector<hipEvent_t> launchEvents;   // pairs: start, stop
    size_t launchUsed = 0;                  // pairs in use since the last collect
    ~az_gpu_ctx() {
        if (guide) az_guide_destroy(guide);
        if (scatterGuide) az_guide_destroy(scatterGuide);
        az_release_render_buffers(renderBuffers);
        for (hipEvent_t e : launchEvents) (void)hipEventDestroy(e);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (tev0) (void)hipEventDestroy(tev0);
        if (tev1) (void)hipEventDestroy(tev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

az_status az_upload_geometry(az_gpu_ctx *C, const scene *S, az::UploadInfo &info);
az_status az_upload_shading(az_gpu_ctx *C, const scene *S, const az::UploadInfo &info);
// Device address of the fixed workspace of az_gpu_sum_f32_sequential_noalloc on
// the context's device, which must be current (looked up once per device).
namespace az {
az_status noalloc_workspace(az_gpu_ctx *C, void **Out);
}
