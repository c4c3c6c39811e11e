// guide.h -- device layout of the learned path guide (no reference
// counterpart): records, weights and training buffers.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace az {

constexpr int kGuideIn = 128;
constexpr int kGuideHid = 64;
constexpr int kGuideOut = 40;        // 8 logits, 8 x 3 mean directions, 8 log-concentrations
constexpr int kGuideOutPad = 48;     // outputs padded to whole 16-row MFMA tiles
constexpr int kGuideBiasStride = 64;
constexpr float kGuideKMin = -6.0f;  // log-concentration clamp
constexpr float kGuideKMax = 11.0f;

// A query: position normalized to the scene box, shading normal, roughness.
struct GuideQuery {
    float p[3];
    float n[3];
    float rough;
    float pad;
};

// A training sample: a query plus the sampled direction and its weight
// (target / pdf of the sampling, normalized over the batch).
struct GuideSample {
    float p[3];
    float n[3];
    float rough;
    float dir[3];
    float wgt;
    float pad;
};

struct GuideWeights {
    const short *wt;    // bf16 W^T per layer, [out_pad][in]
    const short *wb;    // bf16 W per layer, [in][out_pad]
    const short *wf;    // bf16 W^T per layer as 16x16x32 MFMA A fragments: tile (mt, t), then lane, then 8 values
    const float *bias;  // 5 x kGuideBiasStride
};

struct GuideTrainBuffers {
    uint32_t batch;
    short *hT;          // per layer: its inputs, features x batch
    short *dyT;         // per layer: gradient of its outputs, out_pad x batch
    float *loss;        // per sample
    float *grad;        // FP32 gradient, laid out like the master parameters
    size_t hOff[5], dyOff[5];
};

}  // namespace az
