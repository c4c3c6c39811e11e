// output.hip -- the output stage of OutputPixelsToFile (main.cpp:1620-1797)
// for the 8-bit formats, ending in the byte conversion of OutputPixelsToPNG
// (main.cpp:1461-1477), on the device.
//
//   k_out_prepare : colour / weight, luminance, log(L + 1e-5), and an
//                   order-preserving integer key of L
//   k_out_logsum  : the reference's sum of the logs, in its order (one
//                   sequential f32 chain), then the scale key / log-average
//   k_out_hist, k_out_pick (x3) : radix select of the luminance at the burnout
//                   percentile (11 + 11 + 10 key bits), the order statistic
//                   that the reference reads from its qsort'ed array
//   k_out_map     : colour / weight again, the photographic operator, gamma
//                   and the bytes
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "azimuth_gpu.h"
#include "gpu_context.h"

namespace az {
namespace {

constexpr uint32_t kOutBlock = 256;
constexpr uint32_t kRadixBins = 2048;

struct OutCtl {
    float logSum;       // sum of log(L + 1e-5), in pixel order
    float scale;        // KeyValue / log-average luminance
    float whiteSqr;     // (white luminance * scale)^2
    uint32_t prefix;    // key bits of the wanted order statistic selected so far
    uint32_t rank;      // its rank among the keys that share that prefix
    uint32_t pad[3];
};

// Luminance (math.cpp:1866-1879).
__device__ inline float out_luminance(float r, float g, float b) { return r * 0.27f + g * 0.67f + b * 0.06f; }

// Clamp01 = MaxOf(0, MinOf(v, 1)) (math.cpp:11-26): NaN comes out as 1.
__device__ inline float out_clamp01(float v) {
    const float m = v < 1.0f ? v : 1.0f;
    return 0.0f > m ? 0.0f : m;
}

// powf of the reference's libm, as the correctly rounded binary32 value of
// the double-precision power.
__device__ inline float out_pow(float b, float e) { return (float)pow((double)b, (double)e); }

// Float -> unsigned key with the same order (negative floats reversed).
__device__ inline uint32_t order_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ inline float key_float(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

__global__ __launch_bounds__(kOutBlock) void k_out_prepare(const float4 *__restrict__ px, uint32_t n,
                                                           uint32_t *__restrict__ keys, float *__restrict__ logs) {
    const uint32_t i = blockIdx.x * kOutBlock + threadIdx.x;
    if (i >= n) return;
    const float4 p = px[i];
    const float L = out_luminance(p.x / p.w, p.y / p.w, p.z / p.w);
    keys[i] = order_key(L);
    logs[i] = (float)log((double)(L + 0.00001f));
}

// One wave: the sum in the reference's order (main.cpp:1643-1661), 64 logs
// per coalesced load, added one after another from lane registers; then
// log-average = exp(sum / n) in double (Exp, math.cpp:145-152) and the
// scale.  Also starts the radix select and clears its histogram.
__global__ __launch_bounds__(64) void k_out_logsum(const float *__restrict__ logs, uint32_t n, float key,
                                                   uint32_t rank, OutCtl *__restrict__ ctl,
                                                   uint32_t *__restrict__ hist) {
    const uint32_t lane = threadIdx.x;
    for (uint32_t b = lane; b < kRadixBins; b += 64u) hist[b] = 0u;
    float s = 0.0f;
    uint32_t base = 0u;
    float x = lane < n ? logs[lane] : 0.0f;
    while (base + 64u <= n) {
        const uint32_t nx = base + 64u + lane;
        const float next = nx < n ? logs[nx] : 0.0f;
#pragma unroll
        for (int j = 0; j < 64; ++j) s += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), j));
        x = next;
        base += 64u;
    }
    for (uint32_t j = 0; base + j < n; ++j) s += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), (int)j));
    if (lane == 0) {
        const float avg = (float)exp((double)(s / (float)n));
        ctl->logSum = s;
        ctl->scale = key / avg;
        ctl->whiteSqr = 0.0f;
        ctl->prefix = 0u;
        ctl->rank = rank;
    }
}

// Histogram of key bits [shift, shift + 11) over the keys whose bits above
// them equal the prefix selected so far (hiMask covers those bits).
__global__ __launch_bounds__(kOutBlock) void k_out_hist(const uint32_t *__restrict__ keys, uint32_t n,
                                                        const OutCtl *__restrict__ ctl, uint32_t *__restrict__ hist,
                                                        uint32_t shift, uint32_t hiMask) {
    __shared__ uint32_t h[kRadixBins];
    for (uint32_t b = threadIdx.x; b < kRadixBins; b += kOutBlock) h[b] = 0u;
    __syncthreads();
    const uint32_t prefix = ctl->prefix;
    for (uint32_t i = blockIdx.x * kOutBlock + threadIdx.x; i < n; i += gridDim.x * kOutBlock) {
        const uint32_t k = keys[i];
        if ((k & hiMask) == prefix) atomicAdd(&h[(k >> shift) & (kRadixBins - 1u)], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < kRadixBins; b += kOutBlock)
        if (h[b]) atomicAdd(&hist[b], h[b]);
}

// The bin that holds the wanted rank: its bits join the prefix, the rank
// becomes the rank within it; the histogram is cleared for the next pass.
// After the last pass the prefix is the key of the order statistic.
__global__ __launch_bounds__(kOutBlock) void k_out_pick(uint32_t *__restrict__ hist, OutCtl *__restrict__ ctl,
                                                        uint32_t shift, uint32_t last) {
    __shared__ uint32_t h[kRadixBins];
    for (uint32_t b = threadIdx.x; b < kRadixBins; b += kOutBlock) {
        h[b] = hist[b];
        hist[b] = 0u;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    uint32_t rank = ctl->rank, b = 0u;
    while (b < kRadixBins - 1u && rank >= h[b]) {
        rank -= h[b];
        ++b;
    }
    const uint32_t prefix = ctl->prefix | (b << shift);
    ctl->prefix = prefix;
    ctl->rank = rank;
    if (last) {
        const float w = key_float(prefix) * ctl->scale;
        ctl->whiteSqr = w * w;
    }
}

// Colour / weight, the photographic operator (main.cpp:1712-1731), gamma
// (main.cpp:1734-1778) and the bytes (main.cpp:1468-1474).
__global__ __launch_bounds__(kOutBlock) void k_out_map(const float4 *__restrict__ px, uint32_t n,
                                                       const OutCtl *__restrict__ ctl, uint32_t tonemapped,
                                                       float saturation, float invGamma, uint32_t srgb,
                                                       uint8_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * kOutBlock + threadIdx.x;
    if (i >= n) return;
    const float4 p = px[i];
    float c[3] = {p.x / p.w, p.y / p.w, p.z / p.w};
    if (tonemapped) {
        const float L = out_luminance(c[0], c[1], c[2]);
        if (L != 0.0f) {
            const float scale = ctl->scale, whiteSqr = ctl->whiteSqr;
            const float Ls = L * scale;
            const float Ld = (Ls * (1.0f + (Ls / whiteSqr))) / (1.0f + Ls);
            for (int k = 0; k < 3; ++k) c[k] = out_clamp01(out_pow(c[k] / L, saturation) * Ld);
        }
    }
    for (int k = 0; k < 3; ++k) {
        float v = c[k];
        if (srgb) v = v > 0.0031308f ? 1.055f * out_pow(v, invGamma) - 0.055f : v * 12.92f;
        else v = out_pow(v, invGamma);
        out[(size_t)3 * i + k] = (uint8_t)(out_clamp01(v) * 255.0f);
    }
}

}  // namespace
}  // namespace az

using namespace az;

#define AZ_HIPO(ctx, call)                                                  \
    do {                                                                    \
        hipError_t _e = (call);                                             \
        if (_e != hipSuccess) {                                             \
            (ctx)->lastError = std::string(#call ": ") + hipGetErrorString(_e); \
            return AZ_ERR_DEVICE;                                           \
        }                                                                   \
    } while (0)

az_status az_gpu_output_pixels(az_gpu_ctx *C, const pixel *d_Pixels, u32 Width, u32 Height, const tonemap *Tonemap,
                               f32 InvGamma, b32 SRGB, u8 *d_Out) {
    if (!C) return AZ_ERR_INVALID_ARGUMENT;
    const uint64_t n64 = (uint64_t)Width * Height;
    if (n64 >= 0x80000000ull) {
        C->lastError = "image of 2^31 pixels or more";
        return AZ_ERR_INVALID_ARGUMENT;
    }
    const uint32_t n = (uint32_t)n64;
    if (n == 0) return AZ_OK;
    if (!d_Pixels || !d_Out) return AZ_ERR_INVALID_ARGUMENT;
    if (Tonemap && Tonemap->Operator != TonemapOperator_None && Tonemap->Operator != TonemapOperator_Photographic)
        return AZ_ERR_INVALID_ARGUMENT;
    const bool tm = Tonemap && Tonemap->Operator == TonemapOperator_Photographic;
    AZ_HIPO(C, hipSetDevice(C->device));
    hipStream_t st = C->stream;
    const dim3 grid((n + kOutBlock - 1) / kOutBlock);
    AZ_HIPO(C, hipEventRecord(C->ev0, st));
    OutCtl *ctl = nullptr;
    if (tm) {
        // The rank the reference reads (RoundAsU32 of f32 arithmetic,
        // main.cpp:1697-1700), checked to fall inside the image.
        const float burnout = Tonemap->BurnoutPercentage;
        const float at = (1.0f - burnout * 0.01f) * (n - 1) + 0.5f;
        if (!(burnout >= 0.0f && burnout <= 100.0f) || !(at >= 0.0f) || !(at < (float)n)) {
            C->lastError = "BurnoutPercentage selects no pixel (it must lie in [0, 100])";
            return AZ_ERR_INVALID_ARGUMENT;
        }
        const uint32_t rank = (uint32_t)at;
        const size_t need = sizeof(OutCtl) + kRadixBins * sizeof(uint32_t) + (size_t)n * 8;
        AZ_HIPO(C, C->outputScratch.alloc(need));
        char *base = (char *)C->outputScratch.ptr;
        ctl = (OutCtl *)base;
        uint32_t *hist = (uint32_t *)(base + sizeof(OutCtl));
        uint32_t *keys = hist + kRadixBins;
        float *logs = (float *)(keys + n);
        hipLaunchKernelGGL(k_out_prepare, grid, dim3(kOutBlock), 0, st, (const float4 *)d_Pixels, n, keys, logs);
        hipLaunchKernelGGL(k_out_logsum, dim3(1), dim3(64), 0, st, (const float *)logs, n, Tonemap->KeyValue, rank, ctl,
                           hist);
        const uint32_t histBlocks = std::min<uint32_t>(grid.x, (uint32_t)C->cuCount * 4u);
        const uint32_t shifts[3] = {21u, 10u, 0u};
        const uint32_t masks[3] = {0u, 0xFFE00000u, 0xFFFFFC00u};
        for (int p = 0; p < 3; ++p) {
            hipLaunchKernelGGL(k_out_hist, dim3(histBlocks), dim3(kOutBlock), 0, st, (const uint32_t *)keys, n,
                               (const OutCtl *)ctl, hist, shifts[p], masks[p]);
            hipLaunchKernelGGL(k_out_pick, dim3(1), dim3(kOutBlock), 0, st, hist, ctl, shifts[p], p == 2 ? 1u : 0u);
        }
    }
    hipLaunchKernelGGL(k_out_map, grid, dim3(kOutBlock), 0, st, (const float4 *)d_Pixels, n, (const OutCtl *)ctl,
                       tm ? 1u : 0u, tm ? Tonemap->Saturation : 1.0f, InvGamma, SRGB ? 1u : 0u, (uint8_t *)d_Out);
    AZ_HIPO(C, hipGetLastError());
    AZ_HIPO(C, hipEventRecord(C->ev1, st));
    C->lastStart = C->ev0;
    C->lastStop = C->ev1;
    C->pendingTiming = true;
    return AZ_OK;
}
