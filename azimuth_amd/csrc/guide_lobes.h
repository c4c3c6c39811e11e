// guide_lobes.h -- the packed form of a query's 8 vMF lobes, written by the
// guide's inference kernel for the renderer and read by its shading kernel
// (no reference counterpart).
//
// One lobe is 8 bytes (two 32-bit words):
//   word 0: bits 0..15 the mixture weight, bits 16..31 the concentration
//           kappa, both IEEE half precision (kappa <= e^11 < 65504);
//   word 1: the mean direction in the octahedral map (Meyer et al. 2010,
//           "On floating-point normal vectors"), x in bits 0..15 and y in
//           bits 16..31 as unsigned 16-bit fixed point over [-1, 1].
// A query's 8 lobes are 64 consecutive bytes.  The sampler and the density
// both read the decoded values (the weights divided by their decoded sum), so
// the directions drawn follow exactly the mixture whose density the estimator
// divides by; the rounding only moves the lobes by about 1e-4 radians and the
// weights and concentrations by about 1e-3 of their values.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace az {

constexpr int kGuideLobeBytes = 64;   // packed lobes per query

__device__ inline uint32_t half_bits(float f) { return (uint32_t)__half_as_ushort(__float2half_rn(f)); }
__device__ inline float half_value(uint32_t b) { return __half2float(__ushort_as_half((unsigned short)b)); }

// Weight, kappa and a unit mean direction into one packed lobe.  A zero
// direction (a degenerate network output) is stored as +z.
__device__ inline uint2 lobe_pack(float w, float kappa, float mx, float my, float mz) {
    const float a = fabsf(mx) + fabsf(my) + fabsf(mz);
    float px = 0.0f, py = 0.0f;
    if (a > 0.0f) {
        const float ia = __builtin_amdgcn_rcpf(a);
        px = mx * ia;
        py = my * ia;
        if (mz < 0.0f) {
            const float tx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
            const float ty = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
            px = tx;
            py = ty;
        }
    }
    const uint32_t ux = (uint32_t)fminf(fmaxf(rintf((px * 0.5f + 0.5f) * 65535.0f), 0.0f), 65535.0f);
    const uint32_t uy = (uint32_t)fminf(fmaxf(rintf((py * 0.5f + 0.5f) * 65535.0f), 0.0f), 65535.0f);
    return make_uint2(half_bits(w) | (half_bits(kappa) << 16), ux | (uy << 16));
}

__device__ inline float lobe_weight(uint2 l) { return half_value(l.x & 0xFFFFu); }
__device__ inline float lobe_kappa(uint2 l) { return half_value(l.x >> 16); }

// The decoded unit mean direction.
__device__ inline void lobe_mean(uint2 l, float &x, float &y, float &z) {
    float px = (float)(l.y & 0xFFFFu) * (2.0f / 65535.0f) - 1.0f;
    float py = (float)(l.y >> 16) * (2.0f / 65535.0f) - 1.0f;
    const float pz = 1.0f - fabsf(px) - fabsf(py);
    if (pz < 0.0f) {
        const float tx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
        const float ty = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
        px = tx;
        py = ty;
    }
    const float inv = __builtin_amdgcn_rsqf(px * px + py * py + pz * pz);
    x = px * inv;
    y = py * inv;
    z = pz * inv;
}

}  // namespace az
