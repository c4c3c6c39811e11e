// render.hip -- the wavefront path tracer behind az_gpu_render.
#include <hip/hip_runtime.h>

#include "azimuth_gpu.h"
#include "gpu_context.h"

az_status az_upload_shading(az_gpu_ctx *C, const scene *S, const az::UploadInfo &info) {
    (void)S;
    (void)info;
    C->render.ready = false;
    return AZ_OK;
}

extern "C" {

az_status az_gpu_render(az_gpu_ctx *C, const camera *, filter_type, b32, const az_render_opts *, pixel *) {
    if (!C) return AZ_ERR_INVALID_ARGUMENT;
    return AZ_ERR_UNSUPPORTED;
}

az_status az_gpu_render_device(az_gpu_ctx *C, const camera *, filter_type, b32, const az_render_opts *, void *) {
    if (!C) return AZ_ERR_INVALID_ARGUMENT;
    return AZ_ERR_UNSUPPORTED;
}

b32 az_render_scene(scene *, camera *, epsilon_values *, u32, b32, pixel *) { return 0; }

}  // extern "C"
