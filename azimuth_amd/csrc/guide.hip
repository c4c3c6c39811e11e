// guide.hip -- the learned path guide: an MLP from an encoded hit point to a
// mixture of 8 von Mises-Fisher lobes, with BF16 MFMA inference and training
// kernels (FP32 accumulation, FP32 master weights, Adam).  No counterpart in
// the reference renderer.  The network, its matrix layout and the inference
// device code are in guide_infer.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "azimuth_gpu.h"
#include "azimuth_guide.h"
#include "gpu_context.h"
#include "guide.h"
#include "guide_infer.h"
#include "guide_lobes.h"

namespace az {

// ---- kernels ----------------------------------------------------------------------------

// Lane-layout check of the MFMA builtin: D = A * B for 16x16 matrices given
// row-major in memory, loaded in the operand layouts this file assumes.
__global__ __launch_bounds__(64) void k_mfma_selftest(const float *A, const float *B, float *D) {
    int lane = threadIdx.x;
    int row = lane & 15, kq = 4 * (lane >> 4);
    v4s a, b;
    for (int i = 0; i < 4; ++i) {
        a[i] = f2bf(A[row * 16 + kq + i]);     // A[m = row][k = kq + i]
        b[i] = f2bf(B[(kq + i) * 16 + row]);   // B[k = kq + i][n = row]
    }
    v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
    c = mfma16(a, b, c);
    for (int j = 0; j < 4; ++j) D[(kq + j) * 16 + row] = c[j];   // D[m = kq + j][n = row]
}

// One LDS copy of the weights per block of kInferWaves waves (3 per SIMD at
// the kernel's VGPR count; the copy, the waves' output tiles and index lists
// take 91 KB of the CU's 160 KB).  After the copy the waves run
// independently, each looping over its own 64-query chunks.
// Plain form: every query, 40 floats of lobes each.  Packed form (the
// renderer's): queries of negative roughness (misses and emitters, which
// shading never guides) are skipped, the others gathered into full batches of
// 64 in a per-wave LDS list, and their lobes stored packed; a skipped
// query's lobes are left as they were.
template <bool PACKED>
__global__ __launch_bounds__(64 * kInferWaves) void k_guide_infer(GuideWeights W, const GuideQuery *__restrict__ q,
                                                                 uint32_t n, void *__restrict__ lobes) {
    __shared__ InferLds L;
    for (int k = threadIdx.x; k < kHalfCount / 8; k += 64 * kInferWaves)
        reinterpret_cast<uint4 *>(L.w)[k] = reinterpret_cast<const uint4 *>(W.wf)[k];
    for (int k = threadIdx.x; k < 5 * kGuideBiasStride; k += 64 * kInferWaves) L.bias[k] = W.bias[k];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t chunks = (n + kInferSamples - 1) / kInferSamples;
    const uint32_t first = blockIdx.x * kInferWaves + wave, stride = gridDim.x * kInferWaves;
    if (!PACKED) {
        for (uint32_t it = first; it < chunks; it += stride) {
            const uint32_t base = it * kInferSamples;
            infer_batch<false>(L, [q, base](uint32_t j) { return q[base + j]; }, min(n - base, (uint32_t)kInferSamples),
                               [base](uint32_t j) { return base + j; }, lobes, wave, lane);
        }
        return;
    }
    uint32_t *list = L.idx[wave];
    uint32_t pending = 0;                            // wave-uniform: gathered indices not yet run
    for (uint32_t it = first;; it += stride) {
        const bool more = it < chunks;
        if (more) {
            const uint32_t s = it * kInferSamples + lane;
            const bool valid = s < n && q[s].rough >= 0.0f;
            const uint64_t m = __ballot(valid);
            if (valid)
                list[pending + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = s;
            pending += (uint32_t)__popcll(m);
            wave_lds_sync();
        }
        if (pending >= (uint32_t)kInferSamples || (!more && pending > 0u)) {
            const uint32_t cnt = min(pending, (uint32_t)kInferSamples);
            infer_batch<true>(L, [q, list](uint32_t j) { return q[list[j]]; }, cnt, [list](uint32_t j) { return list[j]; },
                              lobes, wave, lane);
            // The indices past the batch move to the front of the list.
            const uint32_t rest = pending - cnt;
            const uint32_t v = (uint32_t)lane < rest ? list[kInferSamples + lane] : 0u;
            wave_lds_sync();
            if ((uint32_t)lane < rest) list[lane] = v;
            wave_lds_sync();
            pending = rest;
        }
        if (!more) break;
    }
}

// Training pass 1: forward, loss gradient, backward through the activations.
// Writes H_l^T and dY_l^T (features x batch, bf16) for the weight-gradient
// pass and the per-sample loss.  Persistent blocks of kTrainWaves waves hold
// every layer's weights in LDS, the forward copy as 16x16x32 A fragments
// (W^T) and the backward copy as 16x16x16 A fragments (W), each fragment 64
// consecutive lane words (no bank conflicts); each wave then loops over
// groups of 16 samples (n is a multiple of 16, checked on the host).
constexpr int kTrainWaves = 8;

struct alignas(16) TrainLds {
    short wf[kHalfCount];                      // forward: W^T fragments (mt, K-block of 32), lane, 8 values
    short wbf[kHalfCount];                     // backward: W fragments (in tile, out tile), lane, 4 values
    float bias[5 * kGuideBiasStride];
    float outs[kTrainWaves][16][kGuideOutPad];
};

template <int NI, int NO, bool RELU>
__device__ inline void layer_lds1(const short *w, const float *b, const v4s *x, v4f *acc, v4s *y, int lane) {
    const int kq = 4 * (lane >> 4);
#pragma unroll
    for (int mt = 0; mt < NO; ++mt) {
        v4f c = *reinterpret_cast<const v4f *>(b + 16 * mt + kq);     // the bias starts the sum
#pragma unroll
        for (int t = 0; t < NI / 2; ++t)
            c = mfma32(*reinterpret_cast<const v8s *>(w + ((mt * (NI / 2) + t) * 64 + lane) * 8), x[2 * t], x[2 * t + 1],
                       c);
        acc[mt] = c;
        if (y) y[mt] = RELU ? relu_bf16(c) : to_bf16x4(c);
    }
}

// Gradient of layer L's inputs for 16 samples: dX^T = W dY^T on 16x16x16
// MFMA (A = W [in][out_pad] fragments from LDS, B = the NO tiles of dY^T),
// masked by the derivative of the ReLU that produced the inputs (hPrev, the
// bf16 activations), rounded to bf16.
template <int L, int NO>
__device__ inline void back_layer(const short *wbf, int ln, const v4s *dyIn, const v4s *hPrev, v4s *dyOut) {
    constexpr int op = layer_out_pad(L);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kt = 0; kt < NO; ++kt)
            c = mfma16(*reinterpret_cast<const v4s *>(wbf + h_off(L) + ((mt * (op / 16) + kt) * 64 + ln) * 4), dyIn[kt], c);
#pragma unroll
        for (int j = 0; j < 4; ++j) c[j] = bf2f(hPrev[mt][j]) > 0.0f ? c[j] : 0.0f;
        dyOut[mt] = to_bf16x4(c);
    }
}

__global__ __launch_bounds__(64 * kTrainWaves) void k_guide_train_pass1(GuideWeights W, const GuideSample *__restrict__ smp,
                                                                        uint32_t n, float invN, GuideTrainBuffers T) {
    __shared__ TrainLds Ls;
    for (int k = threadIdx.x; k < kHalfCount / 8; k += 64 * kTrainWaves)
        reinterpret_cast<uint4 *>(Ls.wf)[k] = reinterpret_cast<const uint4 *>(W.wf)[k];
    // W [in][out_pad] into 16x16x16 A fragments: fragment (mt, kt) of layer
    // l, lane r + 16 q holds W[16 mt + r][16 kt + 4 q .. + 3].
    for (int k = threadIdx.x; k < kHalfCount / 4; k += 64 * kTrainWaves) {
        int l = 0;
        while (l < 4 && 4 * k >= h_off(l + 1)) ++l;
        const int op = layer_out_pad(l), local = k - h_off(l) / 4, frag = local >> 6, ln = local & 63;
        const int mt = frag / (op / 16), kt = frag % (op / 16);
        const short *src = W.wb + h_off(l) + (16 * mt + (ln & 15)) * op + 16 * kt + 4 * (ln >> 4);
        *reinterpret_cast<uint2 *>(Ls.wbf + 4 * k) = *reinterpret_cast<const uint2 *>(src);
    }
    for (int k = threadIdx.x; k < 5 * kGuideBiasStride; k += 64 * kTrainWaves) Ls.bias[k] = W.bias[k];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, kq = 4 * (lane >> 4);
    float (*outs)[kGuideOutPad] = Ls.outs[wave];
#pragma unroll 1
    for (uint32_t grp = blockIdx.x * kTrainWaves + wave; grp < n / 16; grp += gridDim.x * kTrainWaves) {
        // The fragment addresses are formed from ln inside the loop: an empty
        // asm marks it as changed each iteration, so that the compiler does
        // not hoist one address register per fragment (about 90) out of the
        // loop and spill.
        int ln = lane;
        asm volatile("" : "+v"(ln));
        const uint32_t s = grp * 16 + col;
        const GuideSample g = smp[s];
        v4s x[8], h[4][4];
        v4f acc[4], out[3];
        encode_tiles(g.p[0], g.p[1], g.p[2], g.n[0], g.n[1], g.n[2], g.rough, lane, x);
        layer_lds1<8, 4, true>(Ls.wf + h_off(0), Ls.bias + 0 * kGuideBiasStride, x, acc, h[0], ln);
        layer_lds1<4, 4, true>(Ls.wf + h_off(1), Ls.bias + 1 * kGuideBiasStride, h[0], acc, h[1], ln);
        layer_lds1<4, 4, true>(Ls.wf + h_off(2), Ls.bias + 2 * kGuideBiasStride, h[1], acc, h[2], ln);
        layer_lds1<4, 4, true>(Ls.wf + h_off(3), Ls.bias + 3 * kGuideBiasStride, h[2], acc, h[3], ln);
        layer_lds1<4, 3, false>(Ls.wf + h_off(4), Ls.bias + 4 * kGuideBiasStride, h[3], out, nullptr, ln);
        // Store H^T (the inputs of every layer) for the weight gradients.
        // Row 16 tile + kq + j of a features x batch array: a wave-uniform
        // row base (16 tile + j) plus the lane's 32-bit offset kq batch + s
        // (kq formed inside the loop like the fragment addresses; the host
        // keeps 32 batch below 2^32).
        const uint32_t laneOff = 2u * ((uint32_t)(4 * (ln >> 4)) * T.batch + s);     // bytes
        auto storeT = [&](short *dst, int tile, v4s v) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                *reinterpret_cast<short *>(reinterpret_cast<char *>(dst + (size_t)(16 * tile + j) * T.batch) + laneOff) = v[j];
        };
#pragma unroll
        for (int t = 0; t < 8; ++t) storeT(T.hT + T.hOff[0], t, x[t]);
#pragma unroll
        for (int l = 0; l < 4; ++l)
#pragma unroll
            for (int t = 0; t < 4; ++t) storeT(T.hT + T.hOff[l + 1], t, h[l][t]);
        // The raw outputs staged in LDS, a row of 48 per sample.
#pragma unroll
        for (int t = 0; t < 3; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) outs[col][16 * t + kq + j] = out[t][j];
        wave_lds_sync();
        // Loss and output gradient, one (sample, lobe) item per lane, two
        // items per lane: the 8 lanes of a sample form the softmax of its
        // weights and the log-sum-exp of its mixture with DPP moves.  All
        // reads of the raw outputs come before the gradient overwrites them.
        float dv[2][5];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int item = lane + 64 * r, sl = item >> 3, k = item & 7;
            const float *o = outs[sl];
            const GuideSample &gs = smp[grp * 16 + sl];
            const float w0 = gs.dir[0], w1 = gs.dir[1], w2 = gs.dir[2], wgt = gs.wgt;
            const float logit = o[k];
            const float m = max8(logit);
            const float e = expf(logit - m);
            const float se = sum8(e);
            const float pi = e / se;
            const float rx = o[8 + 3 * k], ry = o[9 + 3 * k], rz = o[10 + 3 * k];
            const float len = sqrtf(rx * rx + ry * ry + rz * rz);
            const float inv = len > 1e-12f ? 1.0f / len : 0.0f, rn = fmaxf(len, 1e-12f);
            const float mx = rx * inv, my = ry * inv, mz = rz * inv;
            const float kr = o[32 + k], kap = expf(fminf(fmaxf(kr, kGuideKMin), kGuideKMax));
            const float c = mx * w0 + my * w1 + mz * w2;
            const float lp = logf(kap) - logf(2.0f * 3.14159265f) - log1pf(-expf(-2.0f * kap)) + kap * (c - 1.0f) +
                             logf(fmaxf(pi, 1e-30f));
            const float lm = max8(lp);
            const float ls = sum8(expf(lp - lm));
            const float lq = lm + logf(ls);
            if (k == 0) T.loss[grp * 16 + sl] = -wgt * lq;
            const float gq = -wgt * invN;                    // dL / dlog q
            const float rr = expf(lp - lq);
            dv[r][0] = gq * (rr - pi);
            const float f = gq * rr * kap / rn;
            dv[r][1] = f * (w0 - c * mx);
            dv[r][2] = f * (w1 - c * my);
            dv[r][3] = f * (w2 - c * mz);
            const bool inside = kr > kGuideKMin && kr < kGuideKMax;
            const float dl = 1.0f / kap - 2.0f / expm1f(2.0f * kap) + (c - 1.0f);
            dv[r][4] = inside ? gq * rr * dl * kap : 0.0f;
        }
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int item = lane + 64 * r, sl = item >> 3, k = item & 7;
            float *o = outs[sl];
            o[k] = dv[r][0];
            o[8 + 3 * k] = dv[r][1];
            o[9 + 3 * k] = dv[r][2];
            o[10 + 3 * k] = dv[r][3];
            o[32 + k] = dv[r][4];
            o[40 + k] = 0.0f;                                  // padding rows of the output tiles
        }
        wave_lds_sync();
        // dY of the output layer as B tiles; bias gradient; store dY^T.
        v4f dy[4];
        v4s dyb[4];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
#pragma unroll
            for (int j = 0; j < 4; ++j) dy[t][j] = outs[col][16 * t + kq + j];
        }
#pragma unroll
        for (int t = 0; t < 3; ++t) {
#pragma unroll
            for (int j = 0; j < 4; ++j) dyb[t][j] = f2bf(dy[t][j]);
            storeT(T.dyT + T.dyOff[4], t, dyb[t]);
        }
        // Backward: dX^T = W dY^T (A = W [in][out_pad]), masked by ReLU'.
        v4s g4[4], g3[4];
        back_layer<4, 3>(Ls.wbf, ln, dyb, h[3], g4);         // dY of layer 3 (hidden 4)
#pragma unroll
        for (int t = 0; t < 4; ++t) storeT(T.dyT + T.dyOff[3], t, g4[t]);
        back_layer<3, 4>(Ls.wbf, ln, g4, h[2], g3);
#pragma unroll
        for (int t = 0; t < 4; ++t) storeT(T.dyT + T.dyOff[2], t, g3[t]);
        back_layer<2, 4>(Ls.wbf, ln, g3, h[1], g4);
#pragma unroll
        for (int t = 0; t < 4; ++t) storeT(T.dyT + T.dyOff[1], t, g4[t]);
        back_layer<1, 4>(Ls.wbf, ln, g4, h[0], g3);
#pragma unroll
        for (int t = 0; t < 4; ++t) storeT(T.dyT + T.dyOff[0], t, g3[t]);
        wave_lds_sync();   // this group's reads of outs before the next group's writes
    }
}

// Training pass 2: the weight and bias gradients of every layer over one
// slice of kDwSlice samples per block (blockIdx.x the slice, blockIdx.y the
// layer): dW_l [in][out] = H_l^T (in x samples) x dY_l (samples x out) on
// 16x16x32 MFMA.  The block stages kDwSub samples of H_l^T and dY_l^T at a
// time in LDS with coalesced 16-byte loads (rows padded to 272 bytes, so
// that the 16 lanes of a 16-byte operand read touch every bank once); each
// wave accumulates its tiles over the slice in registers and writes them to
// the slice's partial (laid out like the master parameters), with the bias
// gradients summed in sample order.  No atomics: k_guide_update adds the
// partials in slice order.
constexpr int kDwSlice = 512;
constexpr int kDwSub = 128;
constexpr int kDwRow = kDwSub + 8;

struct alignas(16) DwLds {
    short h[kGuideIn][kDwRow];
    short d[kGuideHid][kDwRow];
};

__device__ inline v4f mfma32x(v8s a, v8s b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0,
                                                   0);
}

__global__ __launch_bounds__(256) void k_guide_train_dw(GuideTrainBuffers T, float *__restrict__ part) {
    __shared__ DwLds L;
    const int layer = blockIdx.y;
    const int in = layer_in(layer), out = layer_out(layer), op = layer_out_pad(layer);
    const int tilesN = op / 16, tiles = (in / 16) * tilesN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, row = lane & 15, q = lane >> 4;
    const uint32_t s0 = blockIdx.x * kDwSlice, s1 = min(s0 + (uint32_t)kDwSlice, T.batch);
    const short *hT = T.hT + T.hOff[layer];
    const short *dT = T.dyT + T.dyOff[layer];
    v4f acc[8];
    float bsum[8];                                       // bias gradient partials of the tiles with mt = 0
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        acc[j] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
        bsum[j] = 0.0f;
    }
    for (uint32_t b = s0; b < s1; b += kDwSub) {
        __syncthreads();                                 // the previous sub-slice's reads are done
        // Batches are multiples of 16 samples, so an 8-sample chunk is
        // wholly inside the batch or wholly past it (zeros).
        for (int c = threadIdx.x; c < in * (kDwSub / 8); c += 256) {
            const int r = c / (kDwSub / 8), k8 = (c % (kDwSub / 8)) * 8;
            const uint32_t sm = b + (uint32_t)k8;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (sm < s1) v = *reinterpret_cast<const uint4 *>(hT + (size_t)r * T.batch + sm);
            *reinterpret_cast<uint4 *>(&L.h[r][k8]) = v;
        }
        for (int c = threadIdx.x; c < op * (kDwSub / 8); c += 256) {
            const int r = c / (kDwSub / 8), k8 = (c % (kDwSub / 8)) * 8;
            const uint32_t sm = b + (uint32_t)k8;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (sm < s1) v = *reinterpret_cast<const uint4 *>(dT + (size_t)r * T.batch + sm);
            *reinterpret_cast<uint4 *>(&L.d[r][k8]) = v;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int t = wave + 4 * j;
            if (t < tiles) {
                const int mt = t / tilesN, nt = t % tilesN;
#pragma unroll
                for (int kk = 0; kk < kDwSub / 32; ++kk) {
                    const v8s a = *reinterpret_cast<const v8s *>(&L.h[16 * mt + row][32 * kk + 8 * q]);
                    const v8s d = *reinterpret_cast<const v8s *>(&L.d[16 * nt + row][32 * kk + 8 * q]);
                    acc[j] = mfma32x(a, d, acc[j]);
                    if (mt == 0) {
                        float e = 0.0f;
#pragma unroll
                        for (int x = 0; x < 8; ++x) e += bf2f(d[x]);
                        bsum[j] += e;
                    }
                }
            }
        }
    }
    float *dst = part + (size_t)blockIdx.x * kParamCount;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int t = wave + 4 * j;
        if (t < tiles) {
            const int mt = t / tilesN, nt = t % tilesN, o = 16 * nt + row;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int i = 16 * mt + 4 * q + jj;
                if (o < out) dst[w_off(layer) + i * out + o] = acc[j][jj];
            }
            // The bias partial of output o: the four lane quarters' sums.
            float b = bsum[j];
            b += __shfl_xor(b, 16);
            b += __shfl_xor(b, 32);
            if (mt == 0 && q == 0 && o < out) dst[b_off(layer) + o] = b;
        }
    }
}

// grad + the slices' partials (added in slice order); with apply, Adam on the
// FP32 master (gradient zeroed) and the BF16 copies of the updated parameter
// in the three weight layouts (or the FP32 bias copy).  One thread per
// parameter.
// A block takes 32 parameters; its 8 thread rows sum the slices k with
// k % 8 = row (32 consecutive parameters of one slice: one 128-byte read),
// and the first row adds the 8 row sums in order.
__global__ __launch_bounds__(256) void k_guide_update(const float *__restrict__ part, uint32_t slices,
                                                      float *__restrict__ grad, float *__restrict__ p,
                                                      float *__restrict__ m, float *__restrict__ v, short *__restrict__ wt,
                                                      short *__restrict__ wb, short *__restrict__ wf,
                                                      float *__restrict__ bias, int apply, float lr, float b1, float b2,
                                                      float eps, float bc1, float bc2) {
    __shared__ float rows[8][32];
    const int i = blockIdx.x * 32 + (threadIdx.x & 31), sub = threadIdx.x >> 5;
    float ps = 0.0f;
    if (i < kParamCount)
        for (uint32_t k = sub; k < slices; k += 8) ps += part[(size_t)k * kParamCount + i];
    rows[sub][threadIdx.x & 31] = ps;
    __syncthreads();
    if (sub != 0 || i >= kParamCount) return;
    float s = 0.0f;
#pragma unroll
    for (int r = 0; r < 8; ++r) s += rows[r][threadIdx.x];
    const float gr = grad[i] + s;
    if (!apply) {
        grad[i] = gr;
        return;
    }
    const float mm = b1 * m[i] + (1.0f - b1) * gr;
    const float vv = b2 * v[i] + (1.0f - b2) * gr * gr;
    m[i] = mm;
    v[i] = vv;
    const float np = p[i] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    p[i] = np;
    grad[i] = 0.0f;
    int l = 0;
    while (l < 4 && i >= w_off(l + 1)) ++l;
    const int in = layer_in(l), out = layer_out(l), op = layer_out_pad(l), r = i - w_off(l);
    if (r < in * out) {
        const int ii = r / out, o = r % out, c32 = ii % 32;
        const short bv = f2bf(np);
        wt[h_off(l) + o * in + ii] = bv;                 // W^T [out_pad][in]
        wb[h_off(l) + ii * op + o] = bv;                 // W [in][out_pad]
        // 16x16x32 A fragment (o / 16, ii / 32): lane o % 16 + 16 ((ii % 16) / 4),
        // value 4 ((ii % 32) / 16) + ii % 4 (the inverse of k_guide_refresh's map)
        const int frag = (o / 16) * (in / 32) + ii / 32;
        const int ln = (o % 16) + 16 * ((c32 % 16) / 4), e = 4 * (c32 / 16) + (c32 % 4);
        wf[h_off(l) + frag * 512 + ln * 8 + e] = bv;
    } else {
        bias[l * kGuideBiasStride + (r - in * out)] = np;
    }
}

__global__ __launch_bounds__(256) void k_guide_refresh(const float *__restrict__ p, short *__restrict__ wt,
                                                       short *__restrict__ wb, short *__restrict__ wf,
                                                       float *__restrict__ bias) {
    int i = blockIdx.x * 256 + threadIdx.x;
    for (int l = 0; l < 5; ++l) {
        const int in = layer_in(l), out = layer_out(l), op = layer_out_pad(l);
        if (i < in * op) {
            int r = i / in, c = i % in;          // W^T [out_pad][in]
            wt[h_off(l) + i] = f2bf(r < out ? p[w_off(l) + c * out + r] : 0.0f);
            int r2 = i / op, c2 = i % op;        // W [in][out_pad]
            wb[h_off(l) + i] = f2bf(c2 < out ? p[w_off(l) + r2 * out + c2] : 0.0f);
            // 16x16x32 A fragment (mt, t), lane, value e:
            // W^T[16 mt + lane % 16][32 t + 16 (e / 4) + 4 (lane / 16) + e % 4]
            int frag = i >> 9, ln = (i >> 3) & 63, e = i & 7, nt = in / 32;
            int r3 = 16 * (frag / nt) + (ln & 15), c3 = 32 * (frag % nt) + 16 * (e >> 2) + 4 * (ln >> 4) + (e & 3);
            wf[h_off(l) + i] = f2bf(r3 < out ? p[w_off(l) + c3 * out + r3] : 0.0f);
        }
        if (i < kGuideBiasStride) bias[l * kGuideBiasStride + i] = i < out ? p[b_off(l) + i] : 0.0f;
    }
}

}  // namespace az

using namespace az;

// ---- host ----------------------------------------------------------------------------------
struct az_guide {
    int device = 0;
    hipStream_t stream = nullptr;
    DeviceBuffer params, grad, part, m, v, wt, wb, wf, bias, samples, hT, dyT, loss, queries, lobes, scratch;
    uint32_t step = 0;
    uint32_t batchCap = 0;
    int cuCount = 256;
    std::string lastError;
    float lr = 1e-3f;
};

#define AZ_GHIP(g, call)                                                       \
    do {                                                                       \
        hipError_t _e = (call);                                                \
        if (_e != hipSuccess) {                                                \
            (g)->lastError = std::string(#call ": ") + hipGetErrorString(_e);  \
            return AZ_ERR_DEVICE;                                              \
        }                                                                      \
    } while (0)

static GuideWeights weightsOf(az_guide *G) {
    GuideWeights w;
    w.wt = (const short *)G->wt.ptr;
    w.wb = (const short *)G->wb.ptr;
    w.wf = (const short *)G->wf.ptr;
    w.bias = (const float *)G->bias.ptr;
    return w;
}

static az_status refresh(az_guide *G) {
    int n = kGuideIn * kGuideHid;   // the largest layer
    hipLaunchKernelGGL(k_guide_refresh, dim3((n + 255) / 256), dim3(256), 0, G->stream, (const float *)G->params.ptr,
                       (short *)G->wt.ptr, (short *)G->wb.ptr, (short *)G->wf.ptr, (float *)G->bias.ptr);
    AZ_GHIP(G, hipGetLastError());
    return AZ_OK;
}

extern "C" {

u32 az_guide_param_count(void) { return (u32)kParamCount; }

az_status az_guide_create(az_guide **Out, int dev, u64 Seed) {
    if (!Out) return AZ_ERR_INVALID_ARGUMENT;
    *Out = nullptr;
    az_guide *G = new (std::nothrow) az_guide();
    if (!G) return AZ_ERR_OUT_OF_MEMORY;
    G->device = dev;
    if (hipSetDevice(dev) != hipSuccess || hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking) != hipSuccess ||
        hipDeviceGetAttribute(&G->cuCount, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        delete G;
        return AZ_ERR_DEVICE;
    }
    G->cuCount = std::max(1, G->cuCount);
    // He-normal initial weights from a seeded generator (Box-Muller on
    // splitmix64), zero biases.
    std::vector<float> p(kParamCount, 0.0f);
    uint64_t st = Seed * 0x9E3779B97F4A7C15ull + 1;
    auto u01 = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < 5; ++l) {
        double sd = std::sqrt(2.0 / layer_in(l));
        for (int i = 0; i < layer_in(l) * layer_out(l); ++i) {
            double a = u01(), b = u01();
            p[w_off(l) + i] = (float)(sd * std::sqrt(-2.0 * std::log(a)) * std::cos(2.0 * M_PI * b));
        }
    }
    az_status s = az_guide_set_params(G, p.data(), kParamCount);
    if (s != AZ_OK) {
        az_guide_destroy(G);
        return s;
    }
    *Out = G;
    return AZ_OK;
}

void az_guide_destroy(az_guide *G) {
    if (!G) return;
    (void)hipSetDevice(G->device);
    (void)hipStreamSynchronize(G->stream);
    G->params.release();
    if (G->stream) (void)hipStreamDestroy(G->stream);
    delete G;
}

const char *az_guide_last_error(const az_guide *G) { return G ? G->lastError.c_str() : "no guide"; }

az_status az_guide_set_params(az_guide *G, const f32 *Params, u32 Count) {
    if (!G || !Params || Count != (u32)kParamCount) return AZ_ERR_INVALID_ARGUMENT;
    AZ_GHIP(G, hipSetDevice(G->device));
    size_t pb = (size_t)kParamCount * sizeof(float);
    AZ_GHIP(G, G->params.alloc(pb));
    AZ_GHIP(G, G->grad.alloc(pb));
    AZ_GHIP(G, G->m.alloc(pb));
    AZ_GHIP(G, G->v.alloc(pb));
    AZ_GHIP(G, G->wt.alloc((size_t)kHalfCount * sizeof(short)));
    AZ_GHIP(G, G->wb.alloc((size_t)kHalfCount * sizeof(short)));
    AZ_GHIP(G, G->wf.alloc((size_t)kHalfCount * sizeof(short)));
    AZ_GHIP(G, G->bias.alloc((size_t)5 * kGuideBiasStride * sizeof(float)));
    AZ_GHIP(G, hipMemcpyAsync(G->params.ptr, Params, pb, hipMemcpyHostToDevice, G->stream));
    AZ_GHIP(G, hipMemsetAsync(G->grad.ptr, 0, pb, G->stream));
    AZ_GHIP(G, hipMemsetAsync(G->m.ptr, 0, pb, G->stream));
    AZ_GHIP(G, hipMemsetAsync(G->v.ptr, 0, pb, G->stream));
    G->step = 0;
    az_status s = refresh(G);
    if (s != AZ_OK) return s;
    AZ_GHIP(G, hipStreamSynchronize(G->stream));
    return AZ_OK;
}

az_status az_guide_get_params(az_guide *G, f32 *Params, u32 Count) {
    if (!G || !Params || Count != (u32)kParamCount) return AZ_ERR_INVALID_ARGUMENT;
    AZ_GHIP(G, hipMemcpyAsync(Params, G->params.ptr, (size_t)kParamCount * sizeof(float), hipMemcpyDeviceToHost,
                              G->stream));
    AZ_GHIP(G, hipStreamSynchronize(G->stream));
    return AZ_OK;
}

az_status az_guide_get_grad(az_guide *G, f32 *Grad, u32 Count) {
    if (!G || !Grad || Count != (u32)kParamCount) return AZ_ERR_INVALID_ARGUMENT;
    AZ_GHIP(G, hipMemcpyAsync(Grad, G->grad.ptr, (size_t)kParamCount * sizeof(float), hipMemcpyDeviceToHost,
                              G->stream));
    AZ_GHIP(G, hipStreamSynchronize(G->stream));
    return AZ_OK;
}

// Persistent blocks, at most one per CU (the LDS a block holds allows one),
// each wave looping over chunks of kInferSamples queries.
static az_status launchInfer(az_guide *G, const void *d_Queries, u32 N, void *d_Lobes, bool packed) {
    const u32 chunks = (N + kInferSamples - 1) / kInferSamples;
    dim3 grid(std::min<u32>((chunks + kInferWaves - 1) / kInferWaves, (u32)G->cuCount));
    if (packed)
        hipLaunchKernelGGL(k_guide_infer<true>, grid, dim3(64 * kInferWaves), 0, G->stream, weightsOf(G),
                           (const GuideQuery *)d_Queries, N, d_Lobes);
    else
        hipLaunchKernelGGL(k_guide_infer<false>, grid, dim3(64 * kInferWaves), 0, G->stream, weightsOf(G),
                           (const GuideQuery *)d_Queries, N, d_Lobes);
    AZ_GHIP(G, hipGetLastError());
    return AZ_OK;
}

az_status az_guide_infer_device(az_guide *G, const void *d_Queries, u32 N, f32 *d_Lobes) {
    if (!G || (N && (!d_Queries || !d_Lobes))) return AZ_ERR_INVALID_ARGUMENT;
    if (N == 0) return AZ_OK;
    return launchInfer(G, d_Queries, N, d_Lobes, false);
}

az_status az_guide_infer_packed_device(az_guide *G, const void *d_Queries, u32 N, void *d_Packed) {
    if (!G || (N && (!d_Queries || !d_Packed))) return AZ_ERR_INVALID_ARGUMENT;
    if (N == 0) return AZ_OK;
    return launchInfer(G, d_Queries, N, d_Packed, true);
}

az_status az_guide_infer_packed(az_guide *G, const f32 *Queries, u32 N, u32 *Packed) {
    if (!G || (N && (!Queries || !Packed))) return AZ_ERR_INVALID_ARGUMENT;
    if (N == 0) return AZ_OK;
    const size_t pb = (size_t)N * kGuideLobeBytes;
    AZ_GHIP(G, G->queries.alloc((size_t)N * sizeof(GuideQuery)));
    AZ_GHIP(G, G->lobes.alloc(pb));
    AZ_GHIP(G, hipMemcpyAsync(G->queries.ptr, Queries, (size_t)N * sizeof(GuideQuery), hipMemcpyHostToDevice, G->stream));
    // Skipped queries keep what the caller's array held.
    AZ_GHIP(G, hipMemcpyAsync(G->lobes.ptr, Packed, pb, hipMemcpyHostToDevice, G->stream));
    az_status s = launchInfer(G, G->queries.ptr, N, G->lobes.ptr, true);
    if (s != AZ_OK) return s;
    AZ_GHIP(G, hipMemcpyAsync(Packed, G->lobes.ptr, pb, hipMemcpyDeviceToHost, G->stream));
    AZ_GHIP(G, hipStreamSynchronize(G->stream));
    return AZ_OK;
}

az_status az_guide_infer(az_guide *G, const f32 *Queries, u32 N, f32 *Lobes) {
    if (!G || (N && (!Queries || !Lobes))) return AZ_ERR_INVALID_ARGUMENT;
    if (N == 0) return AZ_OK;
    AZ_GHIP(G, G->queries.alloc((size_t)N * sizeof(GuideQuery)));
    AZ_GHIP(G, G->lobes.alloc((size_t)N * 40 * sizeof(float)));
    AZ_GHIP(G, hipMemcpyAsync(G->queries.ptr, Queries, (size_t)N * sizeof(GuideQuery), hipMemcpyHostToDevice, G->stream));
    az_status s = az_guide_infer_device(G, G->queries.ptr, N, (f32 *)G->lobes.ptr);
    if (s != AZ_OK) return s;
    AZ_GHIP(G, hipMemcpyAsync(Lobes, G->lobes.ptr, (size_t)N * 40 * sizeof(float), hipMemcpyDeviceToHost, G->stream));
    AZ_GHIP(G, hipStreamSynchronize(G->stream));
    return AZ_OK;
}

// One training step on N device-resident samples (N a multiple of 16):
// gradients of the mean loss, then Adam unless ApplyUpdate is 0.
az_status az_guide_train_device(az_guide *G, const void *d_Samples, u32 N, b32 ApplyUpdate, f32 *MeanLoss) {
    if (!G || !d_Samples || N == 0 || (N % 16) != 0 || N >= (1u << 27)) return AZ_ERR_INVALID_ARGUMENT;
    AZ_GHIP(G, hipSetDevice(G->device));
    if (N > G->batchCap) {
        AZ_GHIP(G, G->hT.alloc((size_t)N * (kGuideIn + 4 * kGuideHid) * sizeof(short)));
        AZ_GHIP(G, G->dyT.alloc((size_t)N * (4 * kGuideHid + kGuideOutPad) * sizeof(short)));
        AZ_GHIP(G, G->loss.alloc((size_t)N * sizeof(float)));
        G->batchCap = N;
    }
    GuideTrainBuffers T;
    T.batch = N;
    T.hT = (short *)G->hT.ptr;
    T.dyT = (short *)G->dyT.ptr;
    T.loss = (float *)G->loss.ptr;
    T.grad = (float *)G->grad.ptr;
    size_t off = 0;
    for (int l = 0; l < 5; ++l) {
        T.hOff[l] = off;
        off += (size_t)layer_in(l) * N;
    }
    off = 0;
    for (int l = 0; l < 5; ++l) {
        T.dyOff[l] = off;
        off += (size_t)layer_out_pad(l) * N;
    }
    dim3 g1(std::min<u32>((N / 16 + kTrainWaves - 1) / kTrainWaves, (u32)G->cuCount));
    hipLaunchKernelGGL(k_guide_train_pass1, g1, dim3(64 * kTrainWaves), 0, G->stream, weightsOf(G),
                       (const GuideSample *)d_Samples, N, 1.0f / (float)N, T);
    AZ_GHIP(G, hipGetLastError());
    const uint32_t slices = (N + kDwSlice - 1) / kDwSlice;
    AZ_GHIP(G, G->part.alloc((size_t)slices * kParamCount * sizeof(float)));
    hipLaunchKernelGGL(k_guide_train_dw, dim3(slices, 5), dim3(256), 0, G->stream, T, (float *)G->part.ptr);
    AZ_GHIP(G, hipGetLastError());
    float bc1 = 1.0f, bc2 = 1.0f;
    if (ApplyUpdate) {
        G->step++;
        bc1 = 1.0f - powf(0.9f, (float)G->step);
        bc2 = 1.0f - powf(0.999f, (float)G->step);
    }
    hipLaunchKernelGGL(k_guide_update, dim3((kParamCount + 31) / 32), dim3(256), 0, G->stream,
                       (const float *)G->part.ptr, slices, (float *)G->grad.ptr, (float *)G->params.ptr,
                       (float *)G->m.ptr, (float *)G->v.ptr, (short *)G->wt.ptr, (short *)G->wb.ptr, (short *)G->wf.ptr,
                       (float *)G->bias.ptr, ApplyUpdate ? 1 : 0, G->lr, 0.9f, 0.999f, 1e-8f, bc1, bc2);
    AZ_GHIP(G, hipGetLastError());
    if (MeanLoss) {
        std::vector<float> h(N);
        AZ_GHIP(G, hipMemcpyAsync(h.data(), G->loss.ptr, (size_t)N * sizeof(float), hipMemcpyDeviceToHost, G->stream));
        AZ_GHIP(G, hipStreamSynchronize(G->stream));
        double s = 0.0;
        for (float x : h) s += x;
        *MeanLoss = (f32)(s / N);
    }
    return AZ_OK;
}

az_status az_guide_train(az_guide *G, const f32 *Samples, u32 N, b32 ApplyUpdate, f32 *MeanLoss) {
    if (!G || !Samples || N == 0 || (N % 16) != 0) return AZ_ERR_INVALID_ARGUMENT;
    AZ_GHIP(G, G->samples.alloc((size_t)N * sizeof(GuideSample)));
    AZ_GHIP(G, hipMemcpyAsync(G->samples.ptr, Samples, (size_t)N * sizeof(GuideSample), hipMemcpyHostToDevice,
                              G->stream));
    az_status s = az_guide_train_device(G, G->samples.ptr, N, ApplyUpdate, MeanLoss);
    if (s != AZ_OK) return s;
    AZ_GHIP(G, hipStreamSynchronize(G->stream));
    return AZ_OK;
}

az_status az_guide_synchronize(az_guide *G) {
    if (!G) return AZ_ERR_INVALID_ARGUMENT;
    AZ_GHIP(G, hipStreamSynchronize(G->stream));
    return AZ_OK;
}

az_status az_guide_zero_grad(az_guide *G) {
    if (!G) return AZ_ERR_INVALID_ARGUMENT;
    AZ_GHIP(G, hipMemsetAsync(G->grad.ptr, 0, (size_t)kParamCount * sizeof(float), G->stream));
    return AZ_OK;
}

az_status az_guide_set_learning_rate(az_guide *G, f32 Lr) {
    if (!G || !(Lr > 0.0f)) return AZ_ERR_INVALID_ARGUMENT;
    G->lr = Lr;
    return AZ_OK;
}

}  // extern "C"

// The renderer's calls: the guide's kernels ordered on the render stream.
az_status az_guide_infer_on_stream(az_guide *G, const void *d_Queries, u32 N, void *d_Packed, hipStream_t s) {
    hipStream_t own = G->stream;
    G->stream = s;
    az_status st = az_guide_infer_packed_device(G, d_Queries, N, d_Packed);
    G->stream = own;
    return st;
}

GuideWeights az_guide_device_weights(az_guide *G) { return weightsOf(G); }

az_status az_guide_train_on_stream(az_guide *G, const void *d_Samples, u32 N, hipStream_t s) {
    hipStream_t own = G->stream;
    G->stream = s;
    az_status st = az_guide_train_device(G, d_Samples, N, 1, nullptr);
    G->stream = own;
    return st;
}

extern "C" {

az_status az_guide_mfma_selftest(az_guide *G, const f32 *A, const f32 *B, f32 *D) {
    if (!G || !A || !B || !D) return AZ_ERR_INVALID_ARGUMENT;
    AZ_GHIP(G, G->scratch.alloc(3 * 256 * sizeof(float)));
    float *d = (float *)G->scratch.ptr;
    AZ_GHIP(G, hipMemcpyAsync(d, A, 256 * sizeof(float), hipMemcpyHostToDevice, G->stream));
    AZ_GHIP(G, hipMemcpyAsync(d + 256, B, 256 * sizeof(float), hipMemcpyHostToDevice, G->stream));
    hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, G->stream, d, d + 256, d + 512);
    AZ_GHIP(G, hipGetLastError());
    AZ_GHIP(G, hipMemcpyAsync(D, d + 512, 256 * sizeof(float), hipMemcpyDeviceToHost, G->stream));
    AZ_GHIP(G, hipStreamSynchronize(G->stream));
    return AZ_OK;
}

}  // extern "C"
