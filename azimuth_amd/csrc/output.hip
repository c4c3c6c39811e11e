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
#include <mutex>
#include <string>
#include <vector>

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

struct ChunkFold {                   // 32 bytes: eight words a lane each in the walk
    int q[2];                        // offset of the integer mantissa over the chunk, by start parity
    int lo[2], hi[2];                // lowest and highest partial offset
    int binade;                      // e the fold assumes, or kNoBinade
    uint32_t ok;                     // bit p: the fold for start parity p holds (values finite, offsets < 2^25)
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

// A run of values as a map of the integer mantissa: for start parity p the
// offset d[p] it adds and its lowest and highest partial offsets.  An empty
// run is the identity (lo = +big, hi = -big).
struct RunMap {
    long long d[2], lo[2], hi[2];
    bool ok;
};

// The map of run a followed by run b.
__device__ inline RunMap compose(const RunMap &a, const RunMap &b) {
    RunMap r;
    for (int p = 0; p < 2; ++p) {
        const int pb = (int)((p + a.d[p]) & 1ll);
        r.d[p] = a.d[p] + b.d[pb];
        r.lo[p] = min(a.lo[p], a.d[p] + b.lo[pb]);
        r.hi[p] = max(a.hi[p], a.d[p] + b.hi[pb]);
    }
    r.ok = a.ok && b.ok;
    return r;
}

// One wave per chunk: each lane maps its 16 values (exact in binary32: the
// scale by 2^(23 - e) only moves the exponent, and q = round(y) is exact for
// |y| <= 2^40), then an order-keeping tree of compositions gives the chunk's
// map in lane 0.
__global__ __launch_bounds__(kOutBlock) void k_out_chunk_fold(const float *__restrict__ logs, uint32_t n, uint32_t nch,
                                                              ChunkFold *__restrict__ folds) {
    const uint32_t k = blockIdx.x * (kOutBlock / 64u) + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (k >= nch) return;
    const int e = folds[k].binade;
    const uint32_t base = k * kChunk, end = min(n, base + kChunk);
    constexpr long long kBig = 1ll << 60;
    RunMap m;
    for (int p = 0; p < 2; ++p) {
        m.d[p] = 0;
        m.lo[p] = kBig;
        m.hi[p] = -kBig;
    }
    m.ok = e != kNoBinade;
    const uint32_t j0 = base + 16u * lane;
    for (uint32_t t = 0; t < 16u && m.ok && j0 + t < end; ++t) {
        const float y = ldexpf(logs[j0 + t], 23 - e);
        if (!(fabsf(y) <= 1099511627776.0f)) {             // not finite, or beyond 2^40
            m.ok = false;
            break;
        }
        const float fl = floorf(y), fr = y - fl;
        const long long qf = (long long)fl;
        for (int p = 0; p < 2; ++p) {
            long long q = qf;
            if (fr > 0.5f) q = qf + 1;
            else if (fr == 0.5f) q = (((long long)p + m.d[p] + qf) & 1ll) ? qf + 1 : qf;
            m.d[p] += q;
            m.lo[p] = min(m.lo[p], m.d[p]);
            m.hi[p] = max(m.hi[p], m.d[p]);
        }
    }
    for (uint32_t s = 1; s < 64u; s <<= 1) {
        RunMap o;
        for (int p = 0; p < 2; ++p) {
            o.d[p] = __shfl_down(m.d[p], s);
            o.lo[p] = __shfl_down(m.lo[p], s);
            o.hi[p] = __shfl_down(m.hi[p], s);
        }
        o.ok = __shfl_down((int)m.ok, s) != 0;
        if ((lane & (2u * s - 1u)) == 0u && lane + s < 64u) m = compose(m, o);
    }
    if (lane == 0) {
        ChunkFold f;
        f.binade = e;
        f.ok = 0u;
        for (int p = 0; p < 2; ++p) {
            const bool fits = m.ok && m.lo[p] > -(1ll << 25) && m.hi[p] < (1ll << 25) && m.d[p] > -(1ll << 25) &&
                              m.d[p] < (1ll << 25);
            f.q[p] = fits ? (int)m.d[p] : 0;
            f.lo[p] = fits ? (int)m.lo[p] : 0;
            f.hi[p] = fits ? (int)m.hi[p] : 0;
            f.ok |= fits ? (1u << p) : 0u;
        }
        folds[k] = f;
    }
}

// One wave walks the chunks, their folds staged in LDS 1024 at a time (one
// wait on global memory per tile; the per-chunk reads are LDS broadcasts):
// where s is a normal float in the fold's
// binade and every partial mantissa stays strictly inside (2^23, 2^24), the
// chunk is one integer step of the mantissa; any other chunk is added
// element by element (64 logs per coalesced load, one after another from
// lane registers).  Then log-average = exp(sum / n) in double (Exp,
// math.cpp:145-152) and the scale; also starts the radix select and clears
// its histogram.
__global__ __launch_bounds__(64) void k_out_logsum(const float *__restrict__ logs, uint32_t n,
                                                   const ChunkFold *__restrict__ folds, float key, uint32_t rank,
                                                   OutCtl *__restrict__ ctl, uint32_t *__restrict__ hist,
                                                   float *__restrict__ sumOut) {
    const uint32_t lane = threadIdx.x;
    if (hist)
        for (uint32_t b = lane; b < kRadixBins; b += 64u) hist[b] = 0u;
    const uint32_t nch = (n + kChunk - 1) / kChunk;
    constexpr uint32_t kTile = 1024;
    __shared__ ChunkFold tile[kTile];
    float s = 0.0f;
    for (uint32_t k0 = 0; k0 < nch; k0 += kTile) {
        const uint32_t tn = min(kTile, nch - k0);
        const uint4 *src = (const uint4 *)(folds + k0);
        uint4 *dst = (uint4 *)tile;
        for (uint32_t i = lane; i < 2u * tn; i += 64u) dst[i] = src[i];
        __syncthreads();
        // Each fold is read whole (two 16-byte LDS reads), one chunk ahead.
        uint4 na = dst[0], nb = dst[1];
        for (uint32_t k = k0; k < k0 + tn; ++k) {
            const uint4 fa = na, fb = nb;   // q[0], q[1], lo[0], lo[1] | hi[0], hi[1], binade, ok
            if (k + 1u < k0 + tn) {
                na = dst[2u * (k + 1u - k0)];
                nb = dst[2u * (k + 1u - k0) + 1u];
            }
            bool done = false;
            const uint32_t bits = __float_as_uint(s);
            const uint32_t E = (bits >> 23) & 0xFFu;
            if (E != 0u && E != 0xFFu && (int)E - 127 == (int)fb.z) {
                const int M = (int)((bits & 0x7FFFFFu) | 0x800000u);
                const int S = (bits >> 31) ? -M : M;
                const int p = S & 1;
                if ((fb.w >> p) & 1u) {
                    const int q = (int)(p ? fa.y : fa.x);
                    const int lo = (int)(p ? fa.w : fa.z);
                    const int hi = (int)(p ? fb.y : fb.x);
                    const bool inside = S > 0 ? (S + lo > (1 << 23) && S + hi < (1 << 24))
                                              : (S + hi < -(1 << 23) && S + lo > -(1 << 24));
                    if (inside) {
                        const int S2 = S + q;
                        const uint32_t M2 = (uint32_t)(S2 < 0 ? -S2 : S2);
                        s = __uint_as_float((bits & 0xFF800000u) | (M2 & 0x7FFFFFu));
                        done = true;
                    }
                }
            }
            if (!done && k * kChunk + kChunk <= n) {
                // A whole chunk: its 1024 logs in 16 registers a lane (one
                // wait), then added in order.
                const uint32_t base = k * kChunk;
                float v[kChunk / 64];
#pragma unroll
                for (uint32_t t = 0; t < kChunk / 64; ++t) v[t] = logs[base + 64u * t + lane];
#pragma unroll
                for (uint32_t t = 0; t < kChunk / 64; ++t)
#pragma unroll
                    for (int j = 0; j < 64; ++j) s += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v[t]), j));
                done = true;
            }
            if (!done) {
                const uint32_t base = k * kChunk, end = min(n, base + kChunk);
                for (uint32_t b = base; b < end; b += 64u) {
                    const float x = b + lane < end ? logs[b + lane] : 0.0f;
                    const uint32_t m = min(64u, end - b);
                    if (m == 64u) {
#pragma unroll
                        for (int j = 0; j < 64; ++j)
                            s += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), j));
                    } else {
                        for (uint32_t j = 0; j < m; ++j)
                            s += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), (int)j));
                    }
                }
            }
        }
        __syncthreads();
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
// After the last pass the prefix is the key of the order statistic.  Each
// thread sums 8 bins, a block scan finds the thread whose bins hold the
// rank, and that thread walks its 8.
__global__ __launch_bounds__(kOutBlock) void k_out_pick(uint32_t *__restrict__ hist, OutCtl *__restrict__ ctl,
                                                        uint32_t shift, uint32_t last) {
    __shared__ uint32_t h[kRadixBins];
    __shared__ uint32_t part[kOutBlock];
    constexpr uint32_t kPer = kRadixBins / kOutBlock;
    const uint32_t t = threadIdx.x;
    uint32_t local = 0u;
    for (uint32_t i = 0; i < kPer; ++i) {
        const uint32_t b = t * kPer + i;
        const uint32_t c = hist[b];
        hist[b] = 0u;
        h[b] = c;
        local += c;
    }
    part[t] = local;
    __syncthreads();
    for (uint32_t o = 1; o < kOutBlock; o <<= 1) {
        const uint32_t add = t >= o ? part[t - o] : 0u;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    const uint32_t rank = ctl->rank;
    const uint32_t before = t ? part[t - 1] : 0u;
    if (rank < before || rank >= part[t]) return;
    uint32_t r = rank - before, b = t * kPer;
    while (b < t * kPer + kPer - 1u && r >= h[b]) {
        r -= h[b];
        ++b;
    }
    const uint32_t prefix = ctl->prefix | (b << shift);
    ctl->prefix = prefix;
    ctl->rank = r;
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
    hipLaunchKernelGGL(k_out_chunk_fold, dim3((nch + kOutBlock / 64u - 1) / (kOutBlock / 64u)), dim3(kOutBlock), 0, st,
                       x, n, nch, folds);
    hipLaunchKernelGGL(k_out_logsum, dim3(1), dim3(64), 0, st, x, n, (const ChunkFold *)folds, key, rank, ctl, hist,
                       sumOut);
}

// Workspace of az_gpu_sum_f32_sequential_noalloc: the chunk folds and chunk
// sums for up to kNoAllocMaxValues values (2.5 MiB), one copy on each device,
// held until the process ends.  Host code reaches it only through the address
// hipGetSymbolAddress gives for the current device (noalloc_workspace).
constexpr uint32_t kNoAllocMaxValues = 1u << 26;
constexpr uint32_t kNoAllocChunks = kNoAllocMaxValues / kChunk;
static __device__ __align__(256) unsigned char scratch[(size_t)kNoAllocChunks * (sizeof(ChunkFold) + sizeof(double))];

__global__ void k_store_zero(float *p) { *p = 0.0f; }

std::mutex gWorkspaceMu;
std::vector<void *> gWorkspace;      // device address of `scratch`, by device ordinal

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

// The device address of the workspace on the context's device (current),
// looked up once per device.
az_status az::noalloc_workspace(az_gpu_ctx *C, void **Out) {
    std::lock_guard<std::mutex> lock(gWorkspaceMu);
    if ((size_t)C->device >= gWorkspace.size()) gWorkspace.resize((size_t)C->device + 1, nullptr);
    if (!gWorkspace[C->device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            C->lastError = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return AZ_ERR_DEVICE;
        }
        gWorkspace[C->device] = p;
    }
    *Out = gWorkspace[C->device];
    return AZ_OK;
}

az_status az_gpu_sum_f32_sequential_noalloc(az_gpu_ctx *C, const f32 *d_X, u32 N, f32 *d_Sum) {
    if (!C || !d_Sum || (N && !d_X)) return AZ_ERR_INVALID_ARGUMENT;
    if (N > kNoAllocMaxValues) {
        C->lastError = "az_gpu_sum_f32_sequential_noalloc: more than 2^26 values (its fixed workspace)";
        return AZ_ERR_INVALID_ARGUMENT;
    }
    AZ_HIPO(C, hipSetDevice(C->device));
    void *ws = nullptr;
    const az_status s = noalloc_workspace(C, &ws);
    if (s != AZ_OK) return s;
    hipStream_t st = C->stream;
    if (N == 0) {
        hipLaunchKernelGGL(k_store_zero, dim3(1), dim3(1), 0, st, (float *)d_Sum);
    } else {
        // Every chunk's fold and sum is written before it is read: nothing to clear.
        ChunkFold *folds = (ChunkFold *)ws;
        double *sums = (double *)(folds + kNoAllocChunks);
        launch_sequential_sum(d_X, N, folds, sums, 0.0f, 0u, nullptr, nullptr, d_Sum, st);
    }
    AZ_HIPO(C, hipGetLastError());
    return AZ_OK;
}
