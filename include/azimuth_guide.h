/*
 * azimuth_guide.h -- the learned path guide (no counterpart in the reference
 * renderer): a 128-64-64-64-64-40 MLP from an encoded hit point (16-frequency
 * sin/cos encoding of the position normalized to the scene box, shading
 * normal, roughness) to a mixture of 8 von Mises-Fisher lobes over incident
 * directions.  Inference and training run as BF16 MFMA kernels with FP32
 * accumulation; FP32 master weights are updated by Adam (lr 1e-3, betas 0.9 /
 * 0.999, eps 1e-8).  The renderer mixes it with BSDF sampling by one-sample
 * MIS (az_render_opts.GuideEnable / GuideFraction).
 *
 * Records are float arrays:
 *   query  (8 floats): p01[3], normal[3], roughness, 0
 *   sample (12 floats): p01[3], normal[3], roughness, direction[3], weight, 0
 *   lobes  (40 floats per query): 8 x {weight, mu.x, mu.y, mu.z, kappa}
 * The training loss is the mean over the batch of -weight * log q(direction).
 */
#ifndef AZIMUTH_GUIDE_H
#define AZIMUTH_GUIDE_H

#include "azimuth_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct az_guide az_guide;

u32       az_guide_param_count(void);
/* Seeded He-normal initial weights, zero biases, on HIP device DeviceOrdinal. */
az_status az_guide_create(az_guide **Out, int DeviceOrdinal, u64 Seed);
void      az_guide_destroy(az_guide *Guide);
const char *az_guide_last_error(const az_guide *Guide);
/* FP32 parameters: per layer W [in][out] row-major then b [out]. */
az_status az_guide_set_params(az_guide *Guide, const f32 *Params, u32 Count);
az_status az_guide_get_params(az_guide *Guide, f32 *Params, u32 Count);
/* Accumulated gradient of the last train call(s) made with ApplyUpdate = 0. */
az_status az_guide_get_grad(az_guide *Guide, f32 *Grad, u32 Count);
az_status az_guide_zero_grad(az_guide *Guide);
az_status az_guide_set_learning_rate(az_guide *Guide, f32 LearningRate);
/* Host and device variants; N queries -> N x 40 lobe floats. */
az_status az_guide_infer(az_guide *Guide, const f32 *Queries, u32 N, f32 *Lobes);
az_status az_guide_infer_device(az_guide *Guide, const void *d_Queries, u32 N, f32 *d_Lobes);
/* The renderer's form: N queries -> N x 16 words of packed lobes (8 bytes a
 * lobe: word 0 = half-precision weight | half-precision kappa << 16, word 1 =
 * the mean direction's octahedral map, x | y << 16 as unsigned 16-bit fixed
 * point over [-1, 1]).  Queries of negative roughness are skipped and their
 * 16 words left as the caller's array held them. */
az_status az_guide_infer_packed(az_guide *Guide, const f32 *Queries, u32 N, u32 *Packed);
az_status az_guide_infer_packed_device(az_guide *Guide, const void *d_Queries, u32 N, void *d_Packed);
/* One step on N samples (N a multiple of 16): gradient of the mean loss,
 * then an Adam update when ApplyUpdate is non-zero.  MeanLoss may be NULL. */
az_status az_guide_train(az_guide *Guide, const f32 *Samples, u32 N, b32 ApplyUpdate, f32 *MeanLoss);
az_status az_guide_train_device(az_guide *Guide, const void *d_Samples, u32 N, b32 ApplyUpdate, f32 *MeanLoss);
/* Waits for the guide's queued device work (its own HIP stream). */
az_status az_guide_synchronize(az_guide *Guide);
/* D = A * B for row-major 16x16 matrices through one BF16 MFMA in the
 * operand layouts the guide kernels use (a layout self-check). */
az_status az_guide_mfma_selftest(az_guide *Guide, const f32 *A, const f32 *B, f32 *D);

#ifdef __cplusplus
}
#endif

#endif
