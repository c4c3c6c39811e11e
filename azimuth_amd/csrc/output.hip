// output.hip -- the output stage of OutputPixelsToFile (main.cpp:1620-1797)
// for the 8-bit formats, ending in the byte conversion of OutputPixelsToPNG
// (main.cpp:1461-1477), on the device.
//
//   k_out_prepare : colour / weight, luminance, log(L + 1e-5), and an
//                   order-preserving integer key of L
//   k_out_chunk_sums, k_out_chunk_scan, k_out_chunk_fold, k_out_logsum :
//                   the reference's sum of the logs, bit for bit as its
//                   sequential f32 loop computes it, then the scale
//                   key / log-average
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

// ---- the reference's sequential f32 sum, in parallel -------------------------------------
// The reference adds the n logs one after another into an f32 (main.cpp:
// 1643-1661).  While the running sum s stays inside one binade
// 2^e <= |s| < 2^(e+1) (ulp u = 2^(e-23)), each step fl(s + x) equals
// s + u q with q = x / u rounded to the nearest integer, an exact tie going
// to the q that leaves the integer mantissa (s / u + q) even.  A chunk of
// logs therefore moves s / u by an integer that depends on e and, at ties,
// on the parity of s / u: one lane per chunk computes it for both parities
// (with the lowest and highest partial offsets) at the binade its start is
// expected in, from a double prefix of the chunk sums.  A single wave then
// walks the chunks: where s is in the expected binade and every partial sum
// stays strictly inside it, s jumps by the chunk's offset (each
// intermediate value is then exactly the reference's); any other chunk (s
// near zero, crossing a power of two, a non-finite log) is added element by
// element.  The result is the reference's sum bit for bit.
constexpr uint32_t kChunk = 1024;
constexpr int kNoBinade = -100000;

struct ChunkFold {
    long long q[2];                  // offset of s / u over the chunk, by start parity
    long long lo[2], hi[2];          // lowest and highest partial offset
    int ok[2];                       // 0: a log is not finite or too large for the grid
    int binade;                      // e the fold assumes, or kNoBinade
    int pad;
};

// Approximate chunk sums in double (their order does not matter: they only
// predict the binade).
__global__ __launch_bounds__(kOutBlock) void k_out_chunk_sums(const float *__restrict__ logs, uint32_t n,
                                                              double *__restrict__ sums) {
    __shared__ double part[kOutBlock / 64];
    const uint32_t base = blockIdx.x * kChunk;
    double v = 0.0;
    for (uint32_t j = threadIdx.x; j < kChunk && base + j < n; j += kOutBlock) v += (double)logs[base + j];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) sums[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Exclusive prefix of the chunk sums (one block, in tiles) -> the binade each
// chunk is expected to start in.
__global__ __launch_bounds__(kOutBlock) void k_out_chunk_scan(const double *__restrict__ sums, uint32_t nch,
                                                              ChunkFold *__restrict__ folds) {
    __shared__ double tile[kOutBlock];
    double carry = 0.0;
    for (uint32_t t0 = 0; t0 < nch; t0 += kOutBlock) {
        const uint32_t k = t0 + threadIdx.x;
        tile[threadIdx.x] = k < nch ? sums[k] : 0.0;
        __syncthreads();
        for (uint32_t o = 1; o < kOutBlock; o <<= 1) {
            const double add = threadIdx.x >= o ? tile[threadIdx.x - o] : 0.0;
            __syncthreads();
            tile[threadIdx.x] += add;
            __syncthreads();
        }
        if (k < nch) {
            const float s = (float)(carry + (threadIdx.x ? tile[threadIdx.x - 1] : 0.0));
            int e = kNoBinade;
            if (s != 0.0f && isfinite(s) && fabsf(s) >= 1.17549435e-38f) {
                int ex;
                frexpf(s, &ex);
                e = ex - 1;
            }
            folds[k].binade = e;
        }
        carry += tile[kOutBlock - 1];
        __syncthreads();
    }
}

// One lane per chunk: the integer offsets of the chunk at its binade.
__global__ __launch_bounds__(kOutBlock) void k_out_chunk_fold(const float *__restrict__ logs, uint32_t n, uint32_t nch,
                                                              ChunkFold *__restrict__ folds) {
    const uint32_t k = blockIdx.x * kOutBlock + threadIdx.x;
    if (k >= nch) return;
    ChunkFold f = folds[k];
    const int e = f.binade;
    const uint32_t base = k * kChunk, end = min(n, base + kChunk);
    long long P[2] = {0, 0}, lo[2] = {0, 0}, hi[2] = {0, 0};
    bool ok = e != kNoBinade;
    bool first = true;
    for (uint32_t j = base; ok && j < end; ++j) {
        const double y = ldexp((double)logs[j], 23 - e);    // exact: a power-of-two scale
        if (!(fabs(y) <= 1099511627776.0)) {                  // not finite, or beyond 2^40
            ok = false;
            break;
        }
        const double fl = floor(y), fr = y - fl;
        const long long qf = (long long)fl;
        for (int p = 0; p < 2; ++p) {
            long long q = qf;
            if (fr > 0.5) q = qf + 1;
            else if (fr == 0.5) q = (((long long)p + P[p] + qf) & 1ll) ? qf + 1 : qf;
            P[p] += q;
            lo[p] = first ? P[p] : min(lo[p], P[p]);
            hi[p] = first ? P[p] : max(hi[p], P[p]);
        }
        first = false;
    }
    for (int p = 0; p < 2; ++p) {
        f.q[p] = P[p];
        f.lo[p] = lo[p];
        f.hi[p] = hi[p];
        f.ok[p] = ok ? 1 : 0;
    }
    folds[k] = f;
}

// One wave: walks the chunks (fold where it is exact, else the sequential
// chain: 64 logs per coalesced load, added one after another from lane
// registers); then log-average = exp(sum / n) in double (Exp,
// math.cpp:145-152) and the scale.  Also starts the radix select and clears
// its histogram.
__global__ __launch_bounds__(64) void k_out_logsum(const float *__restrict__ logs, uint32_t n,
                                                   const ChunkFold *__restrict__ folds, float key, uint32_t rank,
                                                   OutCtl *__restrict__ ctl, uint32_t *__restrict__ hist,
                                                   float *__restrict__ sumOut) {
    const uint32_t lane = threadIdx.x;
    if (hist)
        for (uint32_t b = lane; b < kRadixBins; b += 64u) hist[b] = 0u;
    const uint32_t nch = (n + kChunk - 1) / kChunk;
    const long long lo24 = 1ll << 23, hi24 = 1ll << 24;
    float s = 0.0f;
    for (uint32_t k = 0; k < nch; ++k) {
        bool done = false;
        if (s != 0.0f && isfinite(s) && fabsf(s) >= 1.17549435e-38f) {
            int ex;
            frexpf(s, &ex);
            const int e = ex - 1;
            const ChunkFold &f = folds[k];
            if (e == f.binade) {
                const long long S = (long long)ldexp((double)s, 23 - e);   // the integer mantissa, signed
                const int p = (int)(S & 1ll);
                if (f.ok[p]) {
                    const bool inside = S > 0 ? (S + f.lo[p] > lo24 && S + f.hi[p] < hi24)
                                              : (S + f.hi[p] < -lo24 && S + f.lo[p] > -hi24);
                    if (inside) {
                        s = (float)ldexp((double)(S + f.q[p]), e - 23);
                        done = true;
                    }
                }
            }
        }
        if (!done) {
            const uint32_t base = k * kChunk, end = min(n, base + kChunk);
            for (uint32_t b = base; b < end; b += 64u) {
                const float x = b + lane < end ? logs[b + lane] : 0.0f;
                const uint32_t m = min(64u, end - b);
                if (m == 64u) {
#pragma unroll
                    for (int j = 0; j < 64; ++j) s += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), j));
                } else {
                    for (uint32_t j = 0; j < m; ++j)
                        s += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), (int)j));
                }
            }
        }
    }
    if (lane == 0 && sumOut) *sumOut = s;
    if (lane == 0 && ctl) {
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

// The kernels of the sequential sum over x[0..n): the sum to sumOut and, with
// ctl, the scale and the start of the radix select.
void launch_sequential_sum(const float *x, uint32_t n, ChunkFold *folds, double *sums, float key, uint32_t rank,
                           OutCtl *ctl, uint32_t *hist, float *sumOut, hipStream_t st) {
    const uint32_t nch = (n + kChunk - 1) / kChunk;
    hipLaunchKernelGGL(k_out_chunk_sums, dim3(nch), dim3(kOutBlock), 0, st, x, n, sums);
    hipLaunchKernelGGL(k_out_chunk_scan, dim3(1), dim3(kOutBlock), 0, st, (const double *)sums, nch, folds);
    hipLaunchKernelGGL(k_out_chunk_fold, dim3((nch + kOutBlock - 1) / kOutBlock), dim3(kOutBlock), 0, st, x, n, nch,
                       folds);
    hipLaunchKernelGGL(k_out_logsum, dim3(1), dim3(64), 0, st, x, n, (const ChunkFold *)folds, key, rank, ctl, hist,
                       sumOut);
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
        const uint32_t nch = (n + kChunk - 1) / kChunk;
        const size_t need = sizeof(OutCtl) + kRadixBins * sizeof(uint32_t) + (size_t)nch * (sizeof(ChunkFold) + 8) +
                            (size_t)n * 8;
        AZ_HIPO(C, C->outputScratch.alloc(need));
        char *base = (char *)C->outputScratch.ptr;
        ctl = (OutCtl *)base;
        ChunkFold *folds = (ChunkFold *)(base + sizeof(OutCtl));
        double *sums = (double *)(folds + nch);
        uint32_t *hist = (uint32_t *)(sums + nch);
        uint32_t *keys = hist + kRadixBins;
        float *logs = (float *)(keys + n);
        hipLaunchKernelGGL(k_out_prepare, grid, dim3(kOutBlock), 0, st, (const float4 *)d_Pixels, n, keys, logs);
        launch_sequential_sum(logs, n, folds, sums, Tonemap->KeyValue, rank, ctl, hist, nullptr, st);
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

az_status az_gpu_sum_f32_sequential(az_gpu_ctx *C, const f32 *d_X, u32 N, f32 *d_Sum) {
    if (!C || !d_Sum || (N && !d_X)) return AZ_ERR_INVALID_ARGUMENT;
    if (N >= 0x80000000u) {
        C->lastError = "more than 2^31 - 1 values";
        return AZ_ERR_INVALID_ARGUMENT;
    }
    AZ_HIPO(C, hipSetDevice(C->device));
    hipStream_t st = C->stream;
    if (N == 0) {
        AZ_HIPO(C, hipMemsetAsync(d_Sum, 0, sizeof(f32), st));
        return AZ_OK;
    }
    const uint32_t nch = (N + kChunk - 1) / kChunk;
    AZ_HIPO(C, C->outputScratch.alloc((size_t)nch * (sizeof(ChunkFold) + 8)));
    ChunkFold *folds = (ChunkFold *)C->outputScratch.ptr;
    double *sums = (double *)(folds + nch);
    launch_sequential_sum(d_X, N, folds, sums, 0.0f, 0u, nullptr, nullptr, d_Sum, st);
    AZ_HIPO(C, hipGetLastError());
    return AZ_OK;
}
