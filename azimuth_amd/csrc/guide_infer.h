// guide_infer.h -- device code of the learned guide's inference shared by
// the guide's own kernels (guide.hip) and the renderer's fused query and
// inference kernel (render.hip); no counterpart in the reference renderer.
//
// Network: 128 inputs (16 sin/cos frequencies of the box-normalized position,
// the normal, the roughness), 4 hidden layers of 64 (ReLU), 40 outputs
// (8 logits, 8 raw mean directions, 8 log-concentrations).
//
// Matrix layout.  Activations are kept transposed, features x samples, and a
// wave owns 16 samples.  A layer computes Y^T = W^T X^T with
// v_mfma_f32_16x16x16bf16_1k: A = a 16x16 block of W^T (out x in), B = a
// 16x16 block of X^T (in x samples).  The 16x16 result block in the
// accumulator layout (lane l holds rows 4(l/16)..+3 of column l%16) is the B
// operand layout (lane l holds k = 4(l/16)..+3 of column l%16), so each
// layer's output tiles feed the next layer without moving data.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "guide.h"
#include "guide_lobes.h"

namespace az {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// ---- parameter layout ------------------------------------------------------
__host__ __device__ constexpr int layer_in(int l) { return l == 0 ? kGuideIn : kGuideHid; }
__host__ __device__ constexpr int layer_out(int l) { return l == 4 ? kGuideOut : kGuideHid; }
__host__ __device__ constexpr int layer_out_pad(int l) { return l == 4 ? kGuideOutPad : kGuideHid; }
// FP32 master: W_l [in][out] then b_l [out], layer after layer.
__host__ __device__ constexpr int w_off(int l) {
    return l == 0 ? 0 : w_off(l - 1) + layer_in(l - 1) * layer_out(l - 1) + layer_out(l - 1);
}
__host__ __device__ constexpr int b_off(int l) { return w_off(l) + layer_in(l) * layer_out(l); }
constexpr int kParamCount = w_off(5);
// BF16 copies, padded outputs: W^T [out_pad][in] and W [in][out_pad].
__host__ __device__ constexpr int h_off(int l) {
    return l == 0 ? 0 : h_off(l - 1) + layer_in(l - 1) * layer_out_pad(l - 1);
}
constexpr int kHalfCount = h_off(5);

// ---- bf16 and MFMA ------------------------------------------------------------
__device__ inline short f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (short)(u >> 16);   // inf/nan
    u += 0x7FFFu + ((u >> 16) & 1u);                                 // round to nearest even
    return (short)(u >> 16);
}
__device__ inline float bf2f(short h) { return __uint_as_float(((uint32_t)(uint16_t)h) << 16); }

__device__ inline v4f mfma16(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// Four floats to bf16, round to nearest even (two v_cvt_pk_bf16_f32).
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));
__device__ inline v4s to_bf16x4(v4f x) { return __builtin_bit_cast(v4s, __builtin_convertvector(x, v4bf)); }

__device__ inline v4s load4(const short *p) { return *reinterpret_cast<const v4s *>(p); }

// ReLU of four accumulators, one v_med3_f32 each (the median of x, 0 and
// FLT_MAX; fmaxf in IEEE mode would quiet each input first), then bf16.
__device__ inline v4s relu_bf16(v4f c) {
#pragma unroll
    for (int j = 0; j < 4; ++j) c[j] = __builtin_amdgcn_fmed3f(c[j], 0.0f, 3.402823466e+38f);
    return to_bf16x4(c);
}

// gfx950's 16x16x32 bf16 MFMA: lane l holds A row l % 16 and B column l % 16
// at k = 8 (l / 16) .. +7, the result as in the 16x16x16 form.  A K-block of
// 32 is built from two 16-row tiles in the accumulator layout: lane quarter q
// takes rows 4q..4q+3 of the first tile as k = 8q..8q+3 and of the second as
// k = 8q+4..8q+7, and the weight fragments are packed in the same k order.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef short v8s __attribute__((ext_vector_type(8)));
__device__ inline v4f mfma32(v8s a, v4s b0, v4s b1, v4f c) {
    const v8s b = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0,
                                                   0);
}

// Maximum and sum over the 8 consecutive lanes of a sample with DPP lane
// moves: quad swaps for lane ^ 1 and lane ^ 2, then the half-row mirror,
// which pairs lane i with lane 7 - i of each group of 8.  The operands meet
// as with xor shuffles of 1, 2 and 4 (the last step adds the same two values
// in the other order), so the sums are the same bit for bit; no LDS traffic.
template <int CTRL>
__device__ inline float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ inline float max8(float v) {
    v = fmaxf(v, dpp_f<0xB1>(v));      // quad_perm [1, 0, 3, 2]
    v = fmaxf(v, dpp_f<0x4E>(v));      // quad_perm [2, 3, 0, 1]
    return fmaxf(v, dpp_f<0x141>(v));  // row_half_mirror
}
__device__ inline float sum8(float v) {
    v += dpp_f<0xB1>(v);
    v += dpp_f<0x4E>(v);
    return v + dpp_f<0x141>(v);
}

// ---- input encoding ----------------------------------------------------------------
// Features 0..95 are sin and cos of pi * 2^j * p[c] for 16 frequencies j and
// the 3 coordinates c; the order follows the B-operand lanes: the lanes of
// quarter q = lane / 16 hold features 16 t + 4 q + i (t < 6, i < 4), and with
// m = 4 t + i that feature is frequency 4 q + m / 6 of coordinate (m % 6) / 2,
// sin for even m % 6, cos for odd.  A lane evaluates sinpi/cospi at its
// lowest frequency 2^(4q) and doubles three times (sin 2a = 2 sin a cos a,
// cos 2a = (cos a - sin a)(cos a + sin a)), which stays within a few 1e-7 of
// the direct values, far below bf16 rounding.  Features 96..98 are the
// normal, 99 the roughness, 100..127 zero.
__device__ inline void encode_tiles(float px, float py, float pz, float nx, float ny, float nz, float rough, int lane,
                                    v4s in[8]) {
    const int q = lane >> 4;
    const float scale = (float)(1u << (4 * q));
    const float p[3] = {px, py, pz}, n[3] = {nx, ny, nz};
    float S[4][3], Co[4][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        // sin(pi th) and cos(pi th) with the hardware's sine and cosine of a
        // whole-turn argument (v_sin/v_cos take x in turns, |x| <= 256), fed
        // the fraction of th/2: exact reduction, about 1e-6 absolute error.
        const float turns = __builtin_amdgcn_fractf(0.5f * (scale * p[c]));
        S[0][c] = __builtin_amdgcn_sinf(turns);
        Co[0][c] = __builtin_amdgcn_cosf(turns);
#pragma unroll
        for (int j = 1; j < 4; ++j) {
            const float a = S[j - 1][c], b = Co[j - 1][c];
            S[j][c] = 2.0f * a * b;
            Co[j][c] = (b - a) * (b + a);
        }
    }
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        v4f v;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = 4 * t + i, j = m / 6, combo = m % 6, c = combo >> 1;
            v[i] = (combo & 1) ? Co[j][c] : S[j][c];
        }
        in[t] = to_bf16x4(v);
    }
#pragma unroll
    for (int t = 6; t < 8; ++t) {
        v4f v;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int f = 16 * t + 4 * q + i;
            v[i] = f < 99 ? n[f - 96] : (f == 99 ? rough : 0.0f);
        }
        in[t] = to_bf16x4(v);
    }
}

// Inference: per query (p01, n, rough) -> 8 lobes x {w, mu, kappa}.
// A block stages every layer's weights in LDS once, as A fragments (a wave
// reads one fragment as 64 consecutive 8-byte words: no bank conflicts),
// then loops over groups of 256 queries; a wave takes 64 of them, four
// 16-sample column tiles that share each A fragment.
constexpr int kInferGroups = 4;                         // 16-sample tiles a wave carries through the layers
constexpr int kInferWaves = 12;                         // waves per block sharing one LDS copy of the weights
constexpr int kInferSamples = 16 * kInferGroups;        // per wave iteration

struct alignas(16) InferLds {
    short w[kHalfCount];
    float bias[5 * kGuideBiasStride];
    float outs[kInferWaves][16][kGuideOutPad];
    uint32_t idx[kInferWaves][2 * kInferSamples];   // packed form: each wave's gathered query indices
};

template <int NI, int NO, bool RELU>
__device__ inline void layer_lds(const short *w, const float *b, const v4s (*x)[8], v4s (*y)[8], int lane) {
    const int kq = 4 * (lane >> 4);
#pragma unroll
    for (int mt = 0; mt < NO; ++mt) {
        const v4f bias = *reinterpret_cast<const v4f *>(b + 16 * mt + kq);   // the bias starts the sum
        v4f c[kInferGroups];
#pragma unroll
        for (int g = 0; g < kInferGroups; ++g) c[g] = bias;
#pragma unroll
        for (int t = 0; t < NI / 2; ++t) {
            const v8s a = *reinterpret_cast<const v8s *>(w + ((mt * (NI / 2) + t) * 64 + lane) * 8);
#pragma unroll
            for (int g = 0; g < kInferGroups; ++g) c[g] = mfma32(a, x[g][2 * t], x[g][2 * t + 1], c[g]);
        }
#pragma unroll
        for (int g = 0; g < kInferGroups; ++g) y[g][mt] = RELU ? relu_bf16(c[g]) : to_bf16x4(c[g]);
    }
}

// Orders a wave's LDS writes before its later LDS reads of other lanes'
// words (a wave's LDS operations complete in order; this keeps the compiler
// from moving them).
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The network on one wave's 64 queries: the query of tile g, column col is
// query(16 g + col), stored as result number at(16 g + col); the first cnt of
// them are stored (the others are computed on a copy of query(cnt - 1)).  Plain form: 40 floats per query
// at lobes + 40 s; packed form: 8 packed lobes at lobes + 64 s bytes.
template <bool PACKED, class QueryOf, class At>
__device__ inline void infer_batch(InferLds &L, QueryOf query, uint32_t cnt, At at, void *lobes, int wave, int lane) {
    const int col = lane & 15, kq = 4 * (lane >> 4);
    v4s x[kInferGroups][8], h1[kInferGroups][8], h2[kInferGroups][8];
#pragma unroll
    for (int g = 0; g < kInferGroups; ++g) {
        const uint32_t j = 16 * g + col;
        const GuideQuery Q = query(j < cnt ? j : cnt - 1);
        encode_tiles(Q.p[0], Q.p[1], Q.p[2], Q.n[0], Q.n[1], Q.n[2], Q.rough, lane, x[g]);
    }
    layer_lds<8, 4, true>(L.w + h_off(0), L.bias + 0 * kGuideBiasStride, x, h1, lane);
    layer_lds<4, 4, true>(L.w + h_off(1), L.bias + 1 * kGuideBiasStride, h1, h2, lane);
    layer_lds<4, 4, true>(L.w + h_off(2), L.bias + 2 * kGuideBiasStride, h2, h1, lane);
    layer_lds<4, 4, true>(L.w + h_off(3), L.bias + 3 * kGuideBiasStride, h1, h2, lane);
    const short *w4 = L.w + h_off(4);
    const float *b4 = L.bias + 4 * kGuideBiasStride;
#pragma unroll
    for (int g = 0; g < kInferGroups; ++g) {
        // Output layer of this 16-sample tile, staged in LDS.
#pragma unroll
        for (int mt = 0; mt < 3; ++mt) {
            v4f c = *reinterpret_cast<const v4f *>(b4 + 16 * mt + kq);
#pragma unroll
            for (int t = 0; t < 2; ++t)
                c = mfma32(*reinterpret_cast<const v8s *>(w4 + ((mt * 2 + t) * 64 + lane) * 8), h2[g][2 * t],
                           h2[g][2 * t + 1], c);
            *reinterpret_cast<v4f *>(&L.outs[wave][col][16 * mt + kq]) = c;
        }
        wave_lds_sync();
        // Lobe k of sample sl for the 128 (sample, lobe) items, two per
        // lane: the 8 lanes of a sample form its softmax with DPP moves, and
        // consecutive lanes write consecutive lobes.
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int item = lane + 64 * r, sl = item >> 3, k = item & 7;
            const uint32_t j = 16 * g + sl;
            const float *o = L.outs[wave][sl];
            const float logit = o[k];
            const float m = max8(logit);
            const float e = __builtin_amdgcn_exp2f((logit - m) * 1.44269504f);
            const float sum = sum8(e);
            const float x0 = o[8 + 3 * k], y0 = o[9 + 3 * k], z0 = o[10 + 3 * k];
            const float len2 = x0 * x0 + y0 * y0 + z0 * z0;
            const float inv = len2 > 1e-24f ? __builtin_amdgcn_rsqf(len2) : 0.0f;
            const float kr = fminf(fmaxf(o[32 + k], kGuideKMin), kGuideKMax);
            if (j < cnt) {
                const uint32_t s = at(j);
                const float w = e * __builtin_amdgcn_rcpf(sum), kap = __builtin_amdgcn_exp2f(kr * 1.44269504f);
                if (PACKED) {
                    reinterpret_cast<uint2 *>(lobes)[(size_t)s * 8 + k] = lobe_pack(w, kap, x0 * inv, y0 * inv, z0 * inv);
                } else {
                    float *dst = reinterpret_cast<float *>(lobes) + (size_t)s * 40 + 5 * k;
                    dst[0] = w;
                    dst[1] = x0 * inv;
                    dst[2] = y0 * inv;
                    dst[3] = z0 * inv;
                    dst[4] = kap;
                }
            }
        }
        wave_lds_sync();   // this tile's reads before the next tile's writes
    }
}

}  // namespace az
