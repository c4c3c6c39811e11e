// render.hip -- the wavefront path tracer behind az_gpu_render.
//
// The reference renders each sample with a depth-first recursion that splits
// at every vertex into NEE shadow rays, one hemisphere scatter ray, a mirror
// ray and, at dielectrics, a refraction and a reflection ray
// (TraceRay, main.cpp:966-1133).  Here that tree is expanded breadth first:
// every level of the tree is one queue of PathRay records in HBM, and each
// level runs three kernels:
//   extend  : closest hit of every queued ray (device_trace.h)
//   shade   : hit reconstruction, emission/ambient, NEE shadow-ray records,
//             children for the next level with their throughput weights
//   connect : occlusion of the shadow rays; unoccluded ones add their
//             contribution to their sample
// followed, per chunk of samples, by a splat into the pixel buffer (box or
// Gaussian, main.cpp:1209-1254).  Random numbers are addressed by
// (seed, pixel, sample, tree node, draw), so the breadth-first order draws the
// same numbers as the depth-first CPU oracle.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "azimuth_gpu.h"
#include "device_trace.h"
#include "gpu_context.h"
#include "azimuth_guide.h"
#include "guide_infer.h"
#include "guide_lobes.h"
#include "guide.h"
#include "host_parallel.h"
#include "shade_math.h"

namespace az {

constexpr int kRenderBlock = 256;
constexpr int kRenderRing = 16;

// ---- segmented queues -------------------------------------------------------------
// Appends to one counter serialize in the memory system (about 11 ns each on
// MI355X whatever the wave count), so a producer kernel appends through
// kSegs counters, each on its own 128-byte line: block b writes into segment
// b % nseg, a fixed slice of the queue whose capacity covers the most its
// blocks can emit.  A scan then turns the segment counts into offsets, and a
// consumer maps its dense index i to the slot of the i-th record.
constexpr uint32_t kSegs = 256;
constexpr uint32_t kSegStride = 32;          // uint32 per counter: 128 bytes apart

struct SegIn {
    const uint32_t *offsets;                 // nseg + 1 exclusive prefix of the segment counts
    uint32_t nseg, segCap;                   // nseg == 1: the queue is dense
};

// Slot of dense index i: the segment s with offsets[s] <= i < offsets[s + 1].
__device__ inline uint32_t seg_slot(const uint32_t *off, uint32_t nseg, uint32_t segCap, uint32_t i) {
    if (nseg == 1u) return i;
    uint32_t lo = 0u, hi = nseg;
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo * segCap + (i - off[lo]);
}

// Stages a queue's offsets in LDS; every thread of the block must call it.
__device__ inline const uint32_t *stage_offsets(const SegIn &in, uint32_t *soff) {
    if (in.nseg == 1u) return nullptr;
    for (uint32_t k = threadIdx.x; k <= in.nseg; k += blockDim.x) soff[k] = in.offsets[k];
    __syncthreads();
    return soff;
}

// One slot per lane that wants one: wave-aggregated append to a counter.
__device__ inline uint32_t wave_append(uint32_t *counter, bool want) {
    uint64_t mask = __ballot(want);
    if (!want) return 0xFFFFFFFFu;
    uint32_t lo = (uint32_t)mask, hi = (uint32_t)(mask >> 32);
    uint32_t below = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
    int leader = __builtin_ctzll(mask);
    uint32_t base = 0;
    if (below == 0) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + below;
}

// Per-sample colour accumulators are 64-bit fixed-point sums in units of
// 2^-32.  The contributions of a sample's path-tree vertices and shadow
// rays arrive from different lanes in no fixed order; integer addition makes
// the sums, and so the image, independent of that order (a render is
// reproducible bit for bit).  A contribution is rounded to the nearest unit
// (2.3e-10) and saturated at +-2^30 (NaN counts as 0).
typedef unsigned long long AccWord;
constexpr double kAccUnit = 4294967296.0;   // units per 1.0

__device__ inline AccWord acc_fixed(float v) {
    if (!(v == v)) return 0ull;                                           // NaN adds nothing
    const double x = fmin(fmax((double)v * kAccUnit, -0x1p62), 0x1p62);   // +-inf saturate
    return (AccWord)(long long)rint(x);
}

__device__ inline float acc_value(AccWord w) { return (float)((double)(long long)w / kAccUnit); }

__device__ inline void add3(AccWord *a, V3 c) {
    atomicAdd(a + 0, acc_fixed(c.x));
    atomicAdd(a + 1, acc_fixed(c.y));
    atomicAdd(a + 2, acc_fixed(c.z));
}

// Luminance weights of the guide's training targets.
__device__ inline float guide_lum(V3 c) { return 0.27f * c.x + 0.67f * c.y + 0.06f * c.z; }

// A scatter guide training record's place in its sample's path tree: the
// luminance that its scatter ray's subtree delivered to the sample (fixed
// point, as the colour accumulators), the record of the nearest scatter
// vertex above it (kEmpty at the first), and the depth of its vertex.
struct ScatterLink {
    AccWord sum;
    uint32_t parent;
    uint32_t depth;
};

// A contribution c to the sample, delivered along the subtree of scatter
// record rec (kEmpty: none).
__device__ inline void credit(ScatterLink *links, uint32_t rec, V3 c) {
    if (rec != kEmpty) atomicAdd(&links[rec].sum, acc_fixed(guide_lum(c)));
}

// ---- generate: camera rays (main.cpp:1135-1178) ----------------------------
// Kernel bodies take their block index: the same code runs as a one-scene
// kernel (block = blockIdx.x) and as part of a batched launch over several
// scenes, whose blocks each belong to one scene (block = index within it).
__device__ __forceinline__ void generate_body(const DevScene &S, const CameraSetup &cam, const ChunkParams &cp,
                                     PathRay *__restrict__ q, AccWord *__restrict__ acc, uint32_t n, uint32_t block) {
    uint32_t i = block * kRenderBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t pixLocal = i / cp.spp, sl = i % cp.spp;
    uint32_t row = cp.row0 + pixLocal / cam.width, col = pixLocal % cam.width;
    uint32_t pixel = row * cam.width + col;
    uint32_t s = cp.s0 + sl;
    uint32_t sx = s % cam.sppSqrt, sy = s / cam.sppSqrt;
    float offx = 0.5f * cam.invSppSqrt + sx * cam.invSppSqrt;
    float offy = 0.5f * cam.invSppSqrt + sy * cam.invSppSqrt;
    float scx = col + offx, scy = row + offy;
    Rng rng = make_rng(S.pcgTable, cp.seed, pixel, s, 0u);
    float lx = rng.uniform() - 0.5f;
    float ly = rng.uniform() - 0.5f;
    lx *= cam.aperture;
    ly *= cam.aperture;
    V3 ax = v3(cam.ax), ay = v3(cam.ay);
    V3 lox = ax * lx, loy = ay * ly;
    V3 dir = normalize(v3(cam.base) + scx * v3(cam.stepX) - scy * v3(cam.stepY) - lox - loy);
    V3 org = v3(cam.pos) + lox + loy;
    float time = rng.uniform();
    PathRay r;
    r.ox = org.x;
    r.oy = org.y;
    r.oz = org.z;
    r.time = time;
    r.dx = dir.x;
    r.dy = dir.y;
    r.dz = dir.z;
    r.node = child_node(0u, 0u);
    r.wx = r.wy = r.wz = 1.0f;
    r.sample = i;
    r.cx = r.cy = r.cz = 0.0f;
    r.flags = 0u;
    q[i] = r;
    acc[3 * i + 0] = 0ull;
    acc[3 * i + 1] = 0ull;
    acc[3 * i + 2] = 0ull;
}

__global__ __launch_bounds__(kRenderBlock) void k_generate(DevScene S, CameraSetup cam, ChunkParams cp,
                                                           PathRay *__restrict__ q, AccWord *__restrict__ acc,
                                                           uint32_t n) {
    generate_body(S, cam, cp, q, acc, n, blockIdx.x);
}

// ---- extend: closest hit of every queued ray ---------------------------------
// The traversal kernels of the renderer come in a few (workgroup size,
// register target) variants, chosen per scene by the autotuner.
#define AZ_RENDER_TRACE_VARIANTS(X) \
    X(256, 6, 1) X(128, 6, 1) X(256, 4, 1) X(64, 8, 1) X(256, 6, 0) X(128, 6, 0)

// Path rays of a segmented queue; hits by dense index.
struct ExtendPool {
    const PathRay *q;
    const uint32_t *off;
    SegIn in;
    PathHit *hits;
    __device__ inline bool fetch(uint32_t id, Ray3 &r, float &tMax) {
        const PathRay &pr = q[seg_slot(off, in.nseg, in.segCap, id)];
        r.ox = pr.ox;
        r.oy = pr.oy;
        r.oz = pr.oz;
        r.time = pr.time;
        r.dx = pr.dx;
        r.dy = pr.dy;
        r.dz = pr.dz;
        tMax = kFltMax;
        return false;
    }
    __device__ inline void finish(uint32_t id, bool found, const HitRec &h) {
        PathHit o;
        o.prim = found ? h.prim : kEmpty;
        o.item = h.item;
        o.t = h.t;
        o.tl = h.tl;
        o.beta = h.beta;
        o.gamma = h.gamma;
        o.pad0 = 0.0f;
        o.pad1 = 0.0f;
        hits[id] = o;
    }
};

// A wave's traversal counters added to counters[0..2] (nodes, triangles,
// top-level items) with one atomic each.
__device__ inline void add_render_counters(unsigned long long *counters, const TraceCounters &c) {
    const unsigned long long v[3] = {c.nodes, c.tris, c.prims};
    for (int k = 0; k < 3; ++k) {
        unsigned long long x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
        if ((threadIdx.x & 63) == 0) atomicAdd(counters + k, x);
    }
}

template <int BLOCK, bool WW, bool COUNT = false>
__device__ __forceinline__ void extend_body(const DeviceGeometry &G, const PathRay *__restrict__ q, SegIn qin, uint32_t n,
                                   PathHit *__restrict__ hits, float eps, int cull, uint32_t perWave, uint32_t block,
                                   unsigned long long *counters = nullptr) {
    __shared__ uint32_t smem[(kRenderRing + 4) * BLOCK];
    __shared__ uint32_t soff[kSegs + 1];
    const uint32_t *off = stage_offsets(qin, soff);
    const uint32_t begin = (block * (BLOCK / 64) + (threadIdx.x >> 6)) * perWave;
    const uint32_t end = min(begin + perWave, n);
    ExtendPool pool{q, off, qin, hits};
    TraceCounters c{0u, 0u, 0u};
    trace_pool<0, COUNT, kRenderRing, BLOCK, WW>(G, pool, begin, end, eps, cull != 0, smem + threadIdx.x, c);
    if (COUNT) add_render_counters(counters, c);
}

// Rays per wave for a level whose count is read on the device (nDev), over
// a grid sized for the largest count at 64 rays per wave: the host's choice
// for n (at least 64), or more if the grid's waves would not cover n.
template <int BLOCK, int WAVES>
__device__ inline uint32_t device_rays_per_wave(uint32_t n, int cuCount) {
    const uint32_t waves = gridDim.x * (BLOCK / 64);
    return max(rays_per_wave(n, cuCount, WAVES), (n + waves - 1) / waves);
}

// nDev, when set, holds the level's ray count (written by the previous
// level's segment scan) and n the largest count, for which the grid is
// sized; the waves past the count return at once.
template <int BLOCK, int WAVES, bool WW>
__global__ __launch_bounds__(BLOCK, WAVES) void k_extend(DeviceGeometry G, const PathRay *__restrict__ q, SegIn qin,
                                                        uint32_t n, PathHit *__restrict__ hits, float eps, int cull,
                                                        uint32_t perWave, const uint32_t *__restrict__ nDev, int cuCount) {
    if (nDev) {
        n = min(*nDev, n);
        perWave = device_rays_per_wave<BLOCK, WAVES>(n, cuCount);
        if ((uint64_t)blockIdx.x * (BLOCK / 64) * perWave >= n) return;
    }
    extend_body<BLOCK, WW>(G, q, qin, n, hits, eps, cull, perWave, blockIdx.x);
}

// The same walk counting nodes, triangles and items (untimed renders that
// measure a render's algorithmic bytes; host-driven levels only).
template <int BLOCK, int WAVES, bool WW>
__global__ __launch_bounds__(BLOCK, WAVES) void k_extend_count(DeviceGeometry G, const PathRay *__restrict__ q,
                                                              SegIn qin, uint32_t n, PathHit *__restrict__ hits,
                                                              float eps, int cull, uint32_t perWave,
                                                              unsigned long long *__restrict__ counters) {
    extend_body<BLOCK, WW, true>(G, q, qin, n, hits, eps, cull, perWave, blockIdx.x, counters);
}

// ---- connect: shadow rays ------------------------------------------------------
// Shadow rays of a segmented queue: any hit below tmax for point, area,
// spot, directional and environment lights (main.cpp:697-712, 751-752); the
// closest hit, which must be the light itself, for entity lights
// (main.cpp:704-708).  An unoccluded ray adds its contribution to its sample.
struct ConnectPool {
    const ShadowRay *sq;
    const uint32_t *off;
    SegIn in;
    const DeviceGeometry *G;
    AccWord *acc;
    GuideSample *recs;
    ScatterLink *links;
    __device__ inline const ShadowRay &at(uint32_t id) const { return sq[seg_slot(off, in.nseg, in.segCap, id)]; }
    __device__ inline bool fetch(uint32_t id, Ray3 &r, float &tMax) {
        const ShadowRay &s = at(id);
        r.ox = s.ox;
        r.oy = s.oy;
        r.oz = s.oz;
        r.time = s.time;
        r.dx = s.dx;
        r.dy = s.dy;
        r.dz = s.dz;
        tMax = s.tmax;
        return s.mode == kShadowAny;
    }
    __device__ inline void finish(uint32_t id, bool found, const HitRec &h) {
        const ShadowRay &s = at(id);
        const bool reaches = s.mode == kShadowAny ? !found : (!found || G->items[h.item].entityId == s.entityId);
        if (reaches) {
            add3(acc + 3 * s.sample, v3(s.cx, s.cy, s.cz));
            credit(links, s.srec, v3(s.cx, s.cy, s.cz));
        }
        // A guide training sample whose direction is blocked saw no radiance.
        if (s.rec != kEmpty && !reaches) recs[s.rec].wgt = 0.0f;
    }
};

// ANYONLY: every shadow ray is an occlusion query (no entity lights).
template <int BLOCK, bool WW, bool COUNT = false, bool ANYONLY = false>
__device__ __forceinline__ void connect_body(const DeviceGeometry &G, const ShadowRay *__restrict__ sq, SegIn sin, uint32_t n,
                                    AccWord *__restrict__ acc, GuideSample *__restrict__ recs,
                                    ScatterLink *__restrict__ links, uint32_t perWave, uint32_t block,
                                    unsigned long long *counters = nullptr) {
    __shared__ uint32_t smem[(kRenderRing + 4) * BLOCK];
    __shared__ uint32_t soff[kSegs + 1];
    const uint32_t *off = stage_offsets(sin, soff);
    const uint32_t begin = (block * (BLOCK / 64) + (threadIdx.x >> 6)) * perWave;
    const uint32_t end = min(begin + perWave, n);
    ConnectPool pool{sq, off, sin, &G, acc, recs, links};
    TraceCounters c{0u, 0u, 0u};
    trace_pool<ANYONLY ? 3 : 2, COUNT, kRenderRing, BLOCK, WW>(G, pool, begin, end, 0.0f, false, smem + threadIdx.x, c);
    if (COUNT) add_render_counters(counters, c);
}

template <int BLOCK, int WAVES, bool WW, bool ANYONLY>
__global__ __launch_bounds__(BLOCK, WAVES) void k_connect(DeviceGeometry G, const ShadowRay *__restrict__ sq, SegIn sin,
                                                         uint32_t n, AccWord *__restrict__ acc,
                                                         GuideSample *__restrict__ recs,
                                                         ScatterLink *__restrict__ links, uint32_t perWave,
                                                         const uint32_t *__restrict__ nDev, int cuCount) {
    if (nDev) {
        n = min(*nDev, n);
        perWave = device_rays_per_wave<BLOCK, WAVES>(n, cuCount);
        if ((uint64_t)blockIdx.x * (BLOCK / 64) * perWave >= n) return;
    }
    connect_body<BLOCK, WW, false, ANYONLY>(G, sq, sin, n, acc, recs, links, perWave, blockIdx.x);
}

template <int BLOCK, int WAVES, bool WW>
__global__ __launch_bounds__(BLOCK, WAVES) void k_connect_count(DeviceGeometry G, const ShadowRay *__restrict__ sq,
                                                               SegIn sin, uint32_t n, AccWord *__restrict__ acc,
                                                               GuideSample *__restrict__ recs,
                                                               ScatterLink *__restrict__ links, uint32_t perWave,
                                                               unsigned long long *__restrict__ counters) {
    connect_body<BLOCK, WW, true>(G, sq, sin, n, acc, recs, links, perWave, blockIdx.x, counters);
}

// n is the level's ray count, or with nDev set the largest count the level
// can hold (the count itself is read from nDev on the device).
static hipError_t launch_extend(const RenderTuning &T, int cuCount, const DeviceGeometry &G, const PathRay *q,
                                SegIn qin, uint32_t n, PathHit *hits, float eps, int cull, hipStream_t st,
                                const uint32_t *nDev = nullptr, unsigned long long *counters = nullptr) {
    // With the count on the device, the grid has a wave per 64 rays of the
    // largest count, so that every count gets the split it gets on the host.
    const uint32_t per = nDev ? 64u : rays_per_wave(n, cuCount, T.traceWaves);
    if (counters) {   // counting: the 128-lane, 6-wave walk in the tuned loop form
        const uint32_t perC = rays_per_wave(n, cuCount, 6), perBlock = 2 * perC;
        if (T.traceLoop)
            hipLaunchKernelGGL((k_extend_count<128, 6, true>), dim3((n + perBlock - 1) / perBlock), dim3(128), 0, st, G,
                               q, qin, n, hits, eps, cull, perC, counters);
        else
            hipLaunchKernelGGL((k_extend_count<128, 6, false>), dim3((n + perBlock - 1) / perBlock), dim3(128), 0, st, G,
                               q, qin, n, hits, eps, cull, perC, counters);
        return hipGetLastError();
    }
#define AZ_CASE(B, W, F)                                                                                 \
    if (T.traceBlock == B && T.traceWaves == W && T.traceLoop == F) {                                  \
        const uint32_t perBlock = (B / 64) * per;                                                       \
        hipLaunchKernelGGL((k_extend<B, W, F != 0>), dim3((n + perBlock - 1) / perBlock), dim3(B), 0, st, G, q, qin, \
                           n, hits, eps, cull, per, nDev, cuCount);                                     \
        return hipGetLastError();                                                                       \
    }
    AZ_RENDER_TRACE_VARIANTS(AZ_CASE)
#undef AZ_CASE
    return hipErrorInvalidValue;
}

// anyOnly: the scene has no entity lights, so every shadow ray is an
// occlusion query (k_connect's any-hit-only walk; the interior render 486 ->
// 483 ms).
static hipError_t launch_connect(bool anyOnly, const RenderTuning &T, int cuCount, const DeviceGeometry &G,
                                 const ShadowRay *sq,
                                 SegIn sin, uint32_t n, AccWord *acc, GuideSample *recs, ScatterLink *links,
                                 hipStream_t st, const uint32_t *nDev = nullptr, unsigned long long *counters = nullptr) {
    const uint32_t per = nDev ? 64u : rays_per_wave(n, cuCount, T.traceWaves);
    if (counters) {
        const uint32_t perC = rays_per_wave(n, cuCount, 6), perBlock = 2 * perC;
        if (T.traceLoop)
            hipLaunchKernelGGL((k_connect_count<128, 6, true>), dim3((n + perBlock - 1) / perBlock), dim3(128), 0, st, G,
                               sq, sin, n, acc, recs, links, perC, counters);
        else
            hipLaunchKernelGGL((k_connect_count<128, 6, false>), dim3((n + perBlock - 1) / perBlock), dim3(128), 0, st,
                               G, sq, sin, n, acc, recs, links, perC, counters);
        return hipGetLastError();
    }
#define AZ_CASE(B, W, F)                                                                                  \
    if (T.traceBlock == B && T.traceWaves == W && T.traceLoop == F) {                                   \
        const uint32_t perBlock = (B / 64) * per;                                                        \
        if (anyOnly)                                                                                     \
            hipLaunchKernelGGL((k_connect<B, W, F != 0, true>), dim3((n + perBlock - 1) / perBlock), dim3(B), 0, st, \
                               G, sq, sin, n, acc, recs, links, per, nDev, cuCount);                     \
        else                                                                                             \
            hipLaunchKernelGGL((k_connect<B, W, F != 0, false>), dim3((n + perBlock - 1) / perBlock), dim3(B), 0,    \
                               st, G, sq, sin, n, acc, recs, links, per, nDev, cuCount);                 \
        return hipGetLastError();                                                                        \
    }
    AZ_RENDER_TRACE_VARIANTS(AZ_CASE)
#undef AZ_CASE
    return hipErrorInvalidValue;
}

int render_trace_variants(RenderTuning *out, int cap) {
    int k = 0;
#define AZ_CASE(B, W, F)                           \
    if (k < cap) {                                 \
        out[k] = RenderTuning{};                   \
        out[k].traceBlock = B;                     \
        out[k].traceWaves = W;                     \
        out[k].traceLoop = F;                      \
    }                                              \
    ++k;
    AZ_RENDER_TRACE_VARIANTS(AZ_CASE)
#undef AZ_CASE
    return k < cap ? k : cap;
}

bool render_trace_variant_exists(int block, int waves, int loop) {
#define AZ_CASE(B, W, F) \
    if (block == B && waves == W && loop == F) return true;
    AZ_RENDER_TRACE_VARIANTS(AZ_CASE)
#undef AZ_CASE
    return false;
}

// ---- shade ------------------------------------------------------------------------
// Learned guide inputs and training outputs of the shade kernel.
struct GuideIO {
    const uint2 *lobes;              // per queued ray: 8 packed lobes (guide_lobes.h), or null (plain sampling)
    GuideSample *recs;               // training samples (kSegs segments), or null when not training
    uint32_t *recCounters;           // kSegs counters, kSegStride apart
    uint32_t recSegCap;
    float alpha;                     // probability of drawing from the guide
    // Scatter guide: lobes per queued ray (null: plain scatter sampling);
    // training records, their links and kSegs counters (null when not
    // training); per slot of the current queue, the record whose subtree the
    // ray is in (null at the first level).
    const uint2 *scatLobes;
    GuideSample *srecs;
    ScatterLink *links;
    uint32_t *srecCounters;
    uint32_t srecSegCap;
    const uint32_t *rayRec;
    // Per record slot, the record's path vertex (sample << 32 | random-stream
    // node): the order in which the records are drawn from, whatever order
    // the waves appended them in.
    uint64_t *recKeys;
    uint64_t *srecKeys;
    uint32_t keyNodeBits;            // the node's bits in a key: sample << keyNodeBits | node
};

// Guide query of a hit: position in the scene box, normal, roughness.
__device__ inline float guide_roughness(const DevMaterial &M) {
    bool spec = M.ks[0] != 0.0f || M.ks[1] != 0.0f || M.ks[2] != 0.0f;
    if (!spec) return (M.flags & kMatMirror) && M.kd[0] == 0.0f && M.kd[1] == 0.0f && M.kd[2] == 0.0f ? 0.0f : 1.0f;
    return 2.0f / (M.exponent + 2.0f);
}

__device__ inline void guide_position(const DevScene &S, V3 P, float *p) {
    p[0] = fminf(fmaxf((P.x - S.boxMin[0]) * S.boxInv[0], 0.0f), 1.0f);
    p[1] = fminf(fmaxf((P.y - S.boxMin[1]) * S.boxInv[1], 0.0f), 1.0f);
    p[2] = fminf(fmaxf((P.z - S.boxMin[2]) * S.boxInv[2], 0.0f), 1.0f);
}

// e^x on the hardware's base-2 exponential (about 1e-7 relative for the
// arguments the mixture takes, x <= 0).
__device__ inline float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504f); }

// Mixture pdf q(w) of 8 packed vMF lobes: the weights divided by their sum,
// each lobe kappa / (2 pi (1 - e^-2kappa)) exp(kappa (mu.w - 1)) (Jakob 2012).
// 1 - e^-2kappa by its series below 2 kappa = 1/32, where the subtraction
// would cancel.  A pdf that differs from the sampler's density by a few
// 1e-7 relative biases the estimate by as little.
__device__ inline float guide_pdf(const uint2 *lb, V3 w) {
    float q = 0.0f, wsum = 0.0f;
    for (int k = 0; k < 8; ++k) {
        const uint2 l = lb[k];
        const float wk = lobe_weight(l), kap = lobe_kappa(l);
        float mx, my, mz;
        lobe_mean(l, mx, my, mz);
        const float c = mx * w.x + my * w.y + mz * w.z;
        const float x = 2.0f * kap;
        const float den = x < 0.03125f ? x * (1.0f - x * (0.5f - x * (1.0f / 6.0f - x * (1.0f / 24.0f))))
                                       : 1.0f - fast_exp(-x);
        wsum += wk;
        q += wk * kap * __builtin_amdgcn_rcpf(2.0f * kPiF * den) * fast_exp(kap * (c - 1.0f));
    }
    return wsum > 0.0f ? q / wsum : 0.0f;
}

// Draws a direction from the mixture with three uniforms: lobe choice by the
// weights, then Jakob's inversion for the cosine to the mean and a uniform
// azimuth.
__device__ inline V3 guide_sample(const uint2 *lb, float u0, float u1, float u2) {
    float wsum = 0.0f;
    for (int k = 0; k < 8; ++k) wsum += lobe_weight(lb[k]);
    const float t = u0 * wsum;
    int k = 0;
    float c = lobe_weight(lb[0]);
    while (k < 7 && t >= c) {
        ++k;
        c += lobe_weight(lb[k]);
    }
    const uint2 l = lb[k];
    float mx, my, mz;
    lobe_mean(l, mx, my, mz);
    V3 mu = v3(mx, my, mz);
    float kap = lobe_kappa(l);
    float xi = 1.0f - u1;                                            // (0, 1]
    float wz = 1.0f + logf(xi + (1.0f - xi) * expf(-2.0f * kap)) / kap;
    wz = fminf(fmaxf(wz, -1.0f), 1.0f);
    float s = sqrtf(fmaxf(0.0f, 1.0f - wz * wz));
    V3 tb, b;
    orthonormal_basis(mu, tb, b);
    float ph = 2.0f * kPiF * u2;
    return normalize(mu * wz + tb * (s * cosf(ph)) + b * (s * sinf(ph)));
}

// Outputs of shade: the next level's rays and the shadow rays, both
// segmented (block b appends to segment b % nseg).
// ---- environment luminance sampling ----------------------------------------------------
// The environment map is looked up at u = (pi - atan2(z, x)) / 2pi,
// v = acos(y) / pi (main.cpp:726-733); a (u, v) cell of the W x H grid spans
// the solid angle 2 pi^2 sin(theta) du dv.  Cells are drawn by their
// luminance (3x3 texel maximum, so that nearest and bilinear lookups are
// both covered, plus a small floor) times sin(theta), then a uniform point
// in the cell: pdf(w) = P(cell) W H / (2 pi^2 sin(theta)).

// Largest k < n with cdf[k] <= u (cdf[0] = 0, cdf[n] = 1).
__device__ inline uint32_t cdf_find(const float *cdf, uint32_t n, float u) {
    uint32_t lo = 0u, hi = n;
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ inline V3 env_lum_sample(const DevScene &S, float u0, float u1, float &pdf) {
    const uint32_t W = S.envW, H = S.envH;
    const float *rowCdf = S.envCdf;
    const uint32_t r = cdf_find(rowCdf, H, u0);
    const float pr = rowCdf[r + 1] - rowCdf[r];
    const float fv = fminf(fmaxf((u0 - rowCdf[r]) / pr, 0.0f), 0.99999994f);
    const float *colCdf = S.envCdf + (H + 1) + (size_t)r * (W + 1);
    const uint32_t c = cdf_find(colCdf, W, u1);
    const float pc = colCdf[c + 1] - colCdf[c];
    const float fu = fminf(fmaxf((u1 - colCdf[c]) / pc, 0.0f), 0.99999994f);
    const float u = ((float)c + fu) / (float)W, v = ((float)r + fv) / (float)H;
    const float th = kPiF * v, ph = kPiF - 2.0f * kPiF * u;
    const float st = sinf(th);
    pdf = pr * pc * (float)W * (float)H / (2.0f * kPiF * kPiF * fmaxf(st, 1e-7f));
    return v3(st * cosf(ph), cosf(th), st * sinf(ph));
}

__device__ inline float env_lum_pdf(const DevScene &S, V3 L) {
    const uint32_t W = S.envW, H = S.envH;
    const float u = (-atan2f(L.z, L.x) + kPiF) / (2.0f * kPiF);
    const float v = acosf(fminf(fmaxf(L.y, -1.0f), 1.0f)) / kPiF;
    const uint32_t r = min((uint32_t)fmaxf(v * (float)H, 0.0f), H - 1u);
    const uint32_t c = min((uint32_t)fmaxf(u * (float)W, 0.0f), W - 1u);
    const float *colCdf = S.envCdf + (H + 1) + (size_t)r * (W + 1);
    const float p = (S.envCdf[r + 1] - S.envCdf[r]) * (colCdf[c + 1] - colCdf[c]);
    const float st = sqrtf(fmaxf(0.0f, 1.0f - L.y * L.y));
    return p * (float)W * (float)H / (2.0f * kPiF * kPiF * fmaxf(st, 1e-7f));
}

// The luminance sampler and its density on caller-given inputs (for tests
// and for callers that draw their own environment directions).
__global__ __launch_bounds__(kRenderBlock) void k_env_lum_sample(DevScene S, const float *__restrict__ u, uint32_t n,
                                                                 float *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float pdf;
    const V3 d = env_lum_sample(S, u[2 * i], u[2 * i + 1], pdf);
    out[4 * i] = d.x;
    out[4 * i + 1] = d.y;
    out[4 * i + 2] = d.z;
    out[4 * i + 3] = pdf;
}

__global__ __launch_bounds__(kRenderBlock) void k_env_lum_pdf(DevScene S, const float *__restrict__ dirs, uint32_t n,
                                                              float *__restrict__ pdf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pdf[i] = env_lum_pdf(S, v3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
}

struct ShadeOut {
    PathRay *next;
    ShadowRay *shadow;
    uint32_t *childCounters, *shadowCounters;     // kSegs counters each, kSegStride apart
    uint32_t nseg, childSegCap, shadowSegCap;
    uint32_t *nextRec;               // per slot of next: the child's scatter guide record (null: not kept)
};

__device__ inline void emit_child(const ShadeOut &O, uint32_t seg, bool want, const PathRay &c, uint32_t rec) {
    uint32_t idx = wave_append(O.childCounters + seg * kSegStride, want);
    if (want && idx < O.childSegCap) {
        const size_t slot = (size_t)seg * O.childSegCap + idx;
        O.next[slot] = c;
        if (O.nextRec) O.nextRec[slot] = rec;
    }
}

__device__ inline void emit_shadow(const ShadeOut &O, uint32_t seg, bool want, const ShadowRay &s) {
    uint32_t idx = wave_append(O.shadowCounters + seg * kSegStride, want);
    if (want && idx < O.shadowSegCap) O.shadow[(size_t)seg * O.shadowSegCap + idx] = s;
}

__device__ inline ShadowRay make_shadow(V3 o, V3 d, float time, float tmax, V3 c, uint32_t sample, uint32_t mode,
                                        uint32_t entityId) {
    ShadowRay s;
    s.ox = o.x;
    s.oy = o.y;
    s.oz = o.z;
    s.time = time;
    s.dx = d.x;
    s.dy = d.y;
    s.dz = d.z;
    s.tmax = tmax;
    s.cx = c.x;
    s.cy = c.y;
    s.cz = c.z;
    s.sample = sample;
    s.mode = mode;
    s.entityId = entityId;
    s.rec = kEmpty;
    s.srec = kEmpty;
    return s;
}

__device__ inline PathRay make_child(V3 o, V3 d, float time, uint32_t node, V3 w, uint32_t sample, V3 coef,
                                     uint32_t depth, bool inside) {
    PathRay r;
    r.ox = o.x;
    r.oy = o.y;
    r.oz = o.z;
    r.time = time;
    r.dx = d.x;
    r.dy = d.y;
    r.dz = d.z;
    r.node = node;
    r.wx = w.x;
    r.wy = w.y;
    r.wz = w.z;
    r.sample = sample;
    r.cx = coef.x;
    r.cy = coef.y;
    r.cz = coef.z;
    r.flags = (depth & 0xFFu) | (inside ? 0x100u : 0u);
    return r;
}

// Bump mapping of main.cpp:323-377 on the hit normal N.
__device__ inline V3 bump(const DevScene &S, const DevTexture &T, V3 N, V3 P, float u, float v, V3 Pu, V3 Pv) {
    if (T.type == 0) {
        V3 a = sample_texture_2d(S, T, u, v, false);
        V3 b = sample_texture_2d(S, T, u + T.invW, v, false);
        V3 c = sample_texture_2d(S, T, u, v + T.invH, false);
        V3 du = b - a, dv = c - a;
        float Du = ((du.x + du.y + du.z) / 3.0f) * T.bumpMul;
        float Dv = ((dv.x + dv.y + dv.z) / 3.0f) * T.bumpMul;
        V3 dPu = Pu + Du * N;
        V3 dPv = Pv + Dv * N;
        return normalize(cross(dPv, dPu));
    }
    // sample_perlin at P and at P + e along each axis; the offset points
    // mostly lie in P's lattice cell, whose corner gradients (a chain of
    // dependent permutation loads) are then fetched once.
    const float e = 0.001f;
    const V3 Q = P * T.scaling;
    const int32_t I = (int32_t)floorf(Q.x), J = (int32_t)floorf(Q.y), K = (int32_t)floorf(Q.z);
    V3 gc[8];
    perlin_cell(S, T, I, J, K, gc);
    float D = perlin_value(T, gc, I, J, K, Q);
    float Dd[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const V3 Qa = v3(a == 0 ? P.x + e : P.x, a == 1 ? P.y + e : P.y, a == 2 ? P.z + e : P.z) * T.scaling;
        const int32_t Ia = (int32_t)floorf(Qa.x), Ja = (int32_t)floorf(Qa.y), Ka = (int32_t)floorf(Qa.z);
        if (Ia == I && Ja == J && Ka == K) {
            Dd[a] = perlin_value(T, gc, I, J, K, Qa);
        } else {
            V3 ga[8];
            perlin_cell(S, T, Ia, Ja, Ka, ga);
            Dd[a] = perlin_value(T, ga, Ia, Ja, Ka, Qa);
        }
    }
    const float Dx = Dd[0], Dy = Dd[1], Dz = Dd[2];
    V3 g = v3((Dx - D) / e, (Dy - D) / e, (Dz - D) / e);
    N = N - g * T.bumpMul;
    return normalize(N);
}

// Shading class of a hit for the in-block sort of k_shade: misses, emitters,
// then material flags (mirror, dielectric) and BRDF model, which select the
// branches the vertex code takes; an item's class is set at upload
// (item_shade_class), so the sort key costs one load after the hit's.
constexpr uint32_t kShadeClasses = 32;

inline uint32_t item_shade_class(bool emitter, const DevMaterial &M) {
    if (emitter) return 1u;
    const uint32_t bt = (uint32_t)std::min(std::max(M.brdfType, 0), 6);
    return 2u + (M.flags & 3u) + 4u * bt;
}

__device__ inline uint32_t shade_class(const DevScene &S, const PathHit &h) {
    return h.prim == kEmpty ? 0u : S.itemShade[h.item].shadeClass;
}

// Hit point, shading normal and uv of a closest hit, as RayEntityIntersection
// leaves them in ray_hit (main.cpp:64-146, 147-243, 244-281, 323-382).
__device__ inline void reconstruct(const DevScene &S, const DeviceGeometry &G, const PathRay &r, const PathHit &h,
                                   const DevItemShade &sh, V3 &P, V3 &N, float &u, float &v) {
    const TopItem &it = G.items[h.item];
    float mx = r.time * it.motion[0], my = r.time * it.motion[1], mz = r.time * it.motion[2];
    float sox = r.ox - mx, soy = r.oy - my, soz = r.oz - mz;
    float lox, loy, loz, ldx, ldy, ldz, sc;
    local_ray(it.inv, sox, soy, soz, r.dx, r.dy, r.dz, lox, loy, loz, ldx, ldy, ldz, sc);
    const V3 motion = v3(mx, my, mz);
    const bool textured = sh.texture != 0u;
    const DevTexture *tex = textured ? &S.textures[sh.texture] : nullptr;
    if (sh.kind == kItemSphere) {
        V3 hp = v3(lox + (ldx * h.tl), loy + (ldy * h.tl), loz + (ldz * h.tl));
        V3 lhp = hp - v3(it.center);
        V3 hn = normalize(lhp);
        P = point_to_world(hp, it.fwd) + motion;
        float theta = Acos(hn.y);
        float phi = Atan2F(hn.z, hn.x);
        float twoPi = kPiF * 2.0f;
        u = (-phi + kPiF) / twoPi;
        v = theta / kPiF;
        if (textured) {
            N = hn;
            if (tex->bump) {
                float pyPi = kPiF * lhp.y, rPi = kPiF * it.radius;
                float sph, cph;
                SinCosF(phi, sph, cph);
                V3 Pv = v3(cph * pyPi, -SinF(theta) * rPi, sph * pyPi);
                V3 Pu = v3(twoPi * lhp.z, 0.0f, -twoPi * lhp.x);
                N = bump(S, *tex, N, P, u, v, Pu, Pv);
            }
            N = normal_to_world(N, it.inv);
        } else {
            N = normal_to_world(hn, it.inv);
        }
    } else {
        const DevPrimAttr &pa = S.prims[h.prim];
        V3 Pl = v3(lox + (ldx * h.tl), loy + (ldy * h.tl), loz + (ldz * h.tl));
        float beta = h.beta, gamma = h.gamma;
        float alpha = 1.0f - beta - gamma;
        V3 Nl = pa.smooth ? alpha * v3(pa.n0) + beta * v3(pa.n1) + gamma * v3(pa.n2) : v3(pa.fn);
        u = (alpha * pa.uv0[0]) + (beta * pa.uv1[0]) + (gamma * pa.uv2[0]);
        v = (alpha * pa.uv0[1]) + (beta * pa.uv1[1]) + (gamma * pa.uv2[1]);
        if (textured && tex->bump) {
            float u0 = pa.uv0[0], v0 = pa.uv0[1], u1 = pa.uv1[0], v1 = pa.uv1[1], u2 = pa.uv2[0], v2 = pa.uv2[1];
            V3 P0 = v3(pa.p0), P1 = v3(pa.p1), P2 = v3(pa.p2);
            float div = ((u1 - u0) * (v2 - v0)) - ((v1 - v0) * (u2 - u0));
            div = 1.0f / div;
            V3 Pu = (v2 - v0) * (P1 - P0) + (v0 - v1) * (P2 - P0);
            Pu = Pu * div;
            V3 Pv = (u0 - u2) * (P1 - P0) + (u1 - u0) * (P2 - P0);
            Pv = Pv * div;
            Nl = bump(S, *tex, Nl, Pl, u, v, Pu, Pv);
        }
        P = point_to_world(Pl, it.fwd) + motion;
        N = normal_to_world(Nl, it.inv);
    }

}

// Direction to, distance of, falloff and shadow-ray reach of a point, spot or
// directional light at P, as CalculateLighting computes them for these types
// (main.cpp:611-627, 643-672, 697-699); they draw no random numbers.
__device__ inline void delta_light(const DevLight &L, V3 P, float eps, V3 &Ld, float &ldist, float &falloff,
                                   float &tmax) {
    ldist = 1.0f;
    falloff = 1.0f;
    tmax = 1.0f - eps;
    if (L.type == kLightPoint) {
        Ld = v3(L.pos) - P;
        ldist = length(Ld);
    } else if (L.type == kLightDirectional) {
        Ld = normalize(-v3(L.dir));
        tmax = INFINITY;
    } else {
        Ld = v3(L.pos) - P;
        ldist = length(Ld);
        float ct = dot(normalize(-Ld), normalize(v3(L.dir)));
        if (ct < L.cosCov) falloff = 0.0f;
        else if (ct > L.cosFall) falloff = 1.0f;
        else {
            falloff = (ct - L.cosCov) / (L.cosFall - L.cosCov);
            falloff = PowIF<4>(falloff);
            if (falloff < 0.0f) falloff = 0.0f;
            else if (falloff > 1.0f) falloff = 1.0f;
        }
    }
    if (ldist == 0.0f) ldist = 0.001f;
}

// A cheap test for the same light: false only where the full computation
// gives the vertex nothing, that is where the light lies behind the surface
// (cosI <= 0) or outside a spot's cone (falloff 0).  Estimates by the hardware
// reciprocal square root, within about 1e-6 of the exact cosines, decide with
// a margin of 1e-4 (N is a unit vector); vectors too short or too long for
// that estimate pass.
__device__ inline bool delta_light_may_add(const DevLight &L, V3 P, V3 N) {
    const float m = 1e-4f;
    const V3 Ld = L.type == kLightDirectional ? -v3(L.dir) : v3(L.pos) - P;
    const float l2 = dot(Ld, Ld);
    if (!(l2 > 1e-30f && l2 < 1e30f)) return true;
    const float il = __builtin_amdgcn_rsqf(l2);
    if (dot(N, Ld) * il < -m) return false;
    if (L.type == kLightSpot) {
        const V3 sd = v3(L.dir);
        const float s2 = dot(sd, sd);
        if (s2 > 1e-30f && s2 < 1e30f && -dot(Ld, sd) * il * __builtin_amdgcn_rsqf(s2) < L.cosCov - m) return false;
    }
    return true;
}

__device__ inline bool is_delta_light(int32_t type) {
    return type == kLightPoint || type == kLightSpot || type == kLightDirectional;
}

// Shadow rays of a wave's (lane, light) pairs for point, spot and directional
// lights.  Each lane's cand has bit li set for the lights li < 32 that may
// add to its vertex (delta_light_may_add: every light left out would have
// added exactly nothing, and a vertex inside a dielectric takes none).  The wave's
// lanes share the pairs out (64 at a time) and compute each in full from the
// owner lane's vertex, so that lanes outside a spot's cone or behind the
// surface do not hold the wave in the BRDF and the shadow-ray append (a
// 1080p render with 8 spot lights: shading 189 -> 146 ms).  Every lane
// still shading must call it; pairs of lanes that left are never made.
constexpr uint32_t kPairCap = 1024;   // pairs staged per wave and pass

__device__ __forceinline__ void delta_light_pairs(const DevScene &S, const ShadeOut &O, uint32_t seg, uint32_t cand,
                                                  uint16_t *pairs, uint32_t matIdx, V3 kd, V3 P, V3 N, V3 D, V3 w,
                                                  float time, uint32_t sample, uint32_t rin) {
    const uint32_t lane = __lane_id();
    const uint64_t act = __ballot(true);
    const uint32_t nAct = (uint32_t)__popcll(act);
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(act >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0u));
    // Exclusive prefix of the lanes' pair counts (at most 32: six bits).
    const uint32_t cnt = (uint32_t)__popc(cand);
    uint32_t pre = 0u, total = 0u;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        const uint64_t m = __ballot((cnt >> b) & 1u);
        pre += __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u)) << b;
        total += (uint32_t)__popcll(m) << b;
    }
    const float eps = S.epsShadow;
    for (uint32_t w0 = 0u; w0 < total; w0 += kPairCap) {
        {
            uint32_t bits = cand, idx = pre;
            while (bits) {
                const uint32_t li = (uint32_t)__builtin_ctz(bits);
                bits &= bits - 1u;
                if (idx >= w0 && idx < w0 + kPairCap) pairs[idx - w0] = (uint16_t)((lane << 5) | li);
                ++idx;
            }
        }
        wave_lds_sync();
        const uint32_t n = min(total - w0, kPairCap);
        for (uint32_t r0 = 0u; r0 < n; r0 += nAct) {
            const bool has = r0 + rank < n;
            const uint32_t pr = has ? pairs[r0 + rank] : 0u;
            const int o = (int)(pr >> 5);
            // The owner's vertex.
            const uint32_t mo = __shfl(matIdx, o);
            const V3 Po = v3(__shfl(P.x, o), __shfl(P.y, o), __shfl(P.z, o));
            const V3 No = v3(__shfl(N.x, o), __shfl(N.y, o), __shfl(N.z, o));
            const V3 Do = v3(__shfl(D.x, o), __shfl(D.y, o), __shfl(D.z, o));
            const V3 wo = v3(__shfl(w.x, o), __shfl(w.y, o), __shfl(w.z, o));
            const V3 kdo = v3(__shfl(kd.x, o), __shfl(kd.y, o), __shfl(kd.z, o));
            const float to = __shfl(time, o);
            const uint32_t so = __shfl(sample, o);
            const uint32_t ro = __shfl(rin, o);
            bool want = false;
            V3 c = v3(0.0f, 0.0f, 0.0f), Ld = v3(0.0f, 0.0f, 0.0f);
            float tmax = 0.0f;
            uint32_t ent = 0u;
            if (has) {
                const DevLight &L = S.lights[pr & 31u];
                const DevMaterial M = S.materials[mo];
                float ldist, falloff;
                delta_light(L, Po, eps, Ld, ldist, falloff, tmax);
                want = light_term(M, kdo, v3(L.intensity), Ld, Do, No, ldist, falloff, 1.0f, c);
                if (want) c = had(wo, c);
                want = want && (c.x != 0.0f || c.y != 0.0f || c.z != 0.0f);
                ent = L.entityId;
            }
            ShadowRay sr = make_shadow(Po + (Ld * eps), Ld, to, tmax, c, so, kShadowAny, ent);
            sr.srec = ro;
            emit_shadow(O, seg, want, sr);
        }
        wave_lds_sync();
    }
}

// Which learned-guide code a shading kernel holds: none (plain renders and
// batches), the environment guide's (its lobes and training records), or
// every guide's (with the scatter guide, its records and their links).
// Guide code runs where GI's pointers are set; fewer live values leave the
// register allocation of the lighter kernels less spilling.
enum ShadeGuide { kShadeNoGuide = 0, kShadeEnvGuide = 1, kShadeAllGuides = 2 };

template <int GUIDE>
__device__ __forceinline__ void shade_body(const DevScene &S, const DeviceGeometry &G, const CameraSetup &cam,
                                  const ChunkParams &cp, const PathRay *__restrict__ q, SegIn qin,
                                  const PathHit *__restrict__ hits, uint32_t n, AccWord *__restrict__ acc,
                                  const ShadeOut &O, const GuideIO &GI, uint32_t block) {
    __shared__ uint32_t soff[kSegs + 1];
    __shared__ uint32_t cls[kShadeClasses];
    __shared__ uint16_t order[kRenderBlock];
    __shared__ uint16_t pairs[kRenderBlock / 64][kPairCap];
    // Rays of the block sorted by shading class (a counting sort in LDS), so
    // that the lanes of a wave take the same branches of the vertex code.
    // The key's loads are issued before the queue offsets are staged, so
    // that the two wait together.
    const uint32_t i0 = block * kRenderBlock + threadIdx.x;
    const uint32_t key = i0 < n ? shade_class(S, hits[i0]) : kShadeClasses - 1u;
    const uint32_t *off = stage_offsets(qin, soff);
    {
        if (threadIdx.x < kShadeClasses) cls[threadIdx.x] = 0u;
        __syncthreads();
        atomicAdd(&cls[key], 1u);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t run = 0;
            for (uint32_t k = 0; k < kShadeClasses; ++k) {
                const uint32_t c = cls[k];
                cls[k] = run;
                run += c;
            }
        }
        __syncthreads();
        order[atomicAdd(&cls[key], 1u)] = (uint16_t)threadIdx.x;
        __syncthreads();
    }
    const uint32_t i = block * kRenderBlock + order[threadIdx.x];
    if (i >= n) return;
    const uint32_t seg = block % O.nseg;
    const uint32_t slot = seg_slot(off, qin.nseg, qin.segCap, i);
    const PathRay r = q[slot];
    const PathHit h = hits[i];
    const uint32_t depth = r.flags & 0xFFu;
    const uint32_t rin = GUIDE == kShadeAllGuides && GI.rayRec && depth > 0u ? GI.rayRec[slot] : kEmpty;
    const bool inside = (r.flags & 0x100u) != 0u;
    const uint32_t pixLocal = r.sample / cp.spp;
    const uint32_t pixel = cp.row0 * cam.width + pixLocal;
    const uint32_t s = cp.s0 + r.sample % cp.spp;
    AccWord *a = acc + 3 * r.sample;
    V3 w = v3(r.wx, r.wy, r.wz);
    V3 D = v3(r.dx, r.dy, r.dz);

    if (h.prim == kEmpty) {
        if (depth == 0u) {
            // Camera-ray miss: SampleBackgroundColor (main.cpp:1180-1207).
            V3 bg;
            if (S.useEnv) {
                float th = Acos(D.y);
                float ph = Atan2F(D.z, D.x);
                float u = (-ph + kPiF) / (2.0f * kPiF), v = th / kPiF;
                bg = sample_texture(S, S.env, v3(u, v, 1.0f), u, v);
            } else if (S.useBackgroundTexture) {
                uint32_t row = pixel / cam.width, col = pixel % cam.width;
                float u = (float)col / (float)((int32_t)cam.width - 1), v = (float)row / (float)((int32_t)cam.height - 1);
                bg = sample_texture(S, S.background, v3(u, v, 1.0f), u, v);
            } else {
                bg = v3(S.backgroundColor);
            }
            add3(a, bg);
        }
        return;
    }
    // Beer attenuation over this ray's segment (main.cpp:916-928, 951-958).
    if (r.cx != 0.0f || r.cy != 0.0f || r.cz != 0.0f)
        w = had(w, v3(Exp(r.cx * h.t), Exp(r.cy * h.t), Exp(r.cz * h.t)));

    const DevItemShade sh = S.itemShade[h.item];
    if (sh.isEmitter) {                                   // main.cpp:984-985
        const V3 e = had(w, v3(sh.radiance));
        add3(a, e);
        credit(GI.links, rin, e);
        return;
    }

    V3 P, N;
    float u, v;
    reconstruct(S, G, r, h, sh, P, N, u, v);
    const bool textured = sh.texture != 0u;
    const DevTexture *tex = textured ? &S.textures[sh.texture] : nullptr;

    // ---- the vertex (main.cpp:986-1128) ------------------------------------------------
    const DevMaterial M = S.materials[sh.material];
    V3 kd = v3(M.kd);
    bool ignoreAmbient = false;
    if (textured) {
        V3 ts = sample_texture(S, *tex, P, u, v);
        if (tex->blend == 0) {            // ReplaceAll: DC starts at 1, no ambient
            kd = ts;
            add3(a, had(w, ts));
            credit(GI.links, rin, had(w, ts));
            ignoreAmbient = true;
        } else if (tex->blend == 1) {
            kd = ts;
        } else {
            kd = kd + ts;
            kd = kd * 0.5f;
        }
    }
    Rng rng = make_rng(S.pcgTable, cp.seed, pixel, s, r.node);
    const float eps = S.epsShadow;

    // Lights (main.cpp:609-717).  Point, spot and directional lights 0-31
    // are only marked here (delta_light_pairs computes them); area and
    // entity lights, which draw random numbers, and any later lights run
    // here in light order.
    uint32_t cand = 0u;
    for (uint32_t li = 0; li < S.lightCount; ++li) {
        const DevLight &L = S.lights[li];
        if (li < 32u && is_delta_light(L.type)) {
            if (!inside && delta_light_may_add(L, P, N)) cand |= 1u << li;
            continue;
        }
        V3 Ld = v3(0.0f, 0.0f, 0.0f), Li = v3(L.intensity);
        float falloff = 1.0f, ldist = 1.0f, invPdf = 1.0f, tmax = 1.0f - eps;
        uint32_t mode = kShadowAny;
        if (L.type == kLightPoint) {
            Ld = v3(L.pos) - P;
            ldist = length(Ld);
        } else if (L.type == kLightArea) {
            float r0 = rng.uniform();
            float r1 = rng.uniform();
            V3 lp = v3(L.pos) + v3(L.e0) * r0 + v3(L.e1) * r1;
            falloff = dot(normalize(lp - P), v3(L.normal));
            if (falloff < 0.0f) falloff *= -1.0f;
            Ld = lp - P;
            ldist = length(Ld);
        } else if (L.type == kLightDirectional) {
            Ld = normalize(-v3(L.dir));
            tmax = INFINITY;
        } else if (L.type == kLightSpot) {
            Ld = v3(L.pos) - P;
            ldist = length(Ld);
            float ct = dot(normalize(-Ld), normalize(v3(L.dir)));
            if (ct < L.cosCov) falloff = 0.0f;
            else if (ct > L.cosFall) falloff = 1.0f;
            else {
                falloff = (ct - L.cosCov) / (L.cosFall - L.cosCov);
                falloff = PowIF<4>(falloff);
                if (falloff < 0.0f) falloff = 0.0f;
                else if (falloff > 1.0f) falloff = 1.0f;
            }
        } else {
            // SampleEntityLight (main.cpp:392-459).
            V3 hpl = point_to_world(P, L.inv);
            V3 dl;
            if (L.entityKind == (int32_t)kItemSphere) {
                V3 Dv = v3(L.center) - hpl;
                float rod = L.radius / length(Dv);
                float cosMax = sqrtf(1.0f - sq(rod));
                float cosT = 1.0f + rng.uniform() * (cosMax - 1.0f);
                float sinT = sqrtf(1.0f - sq(cosT));
                float phi = 2.0f * kPiF * rng.uniform();
                float cp_, sp_;
                SinCosF(phi, sp_, cp_);
                V3 W = normalize(Dv), U, Vv;
                orthonormal_basis(W, U, Vv);
                dl = normalize(U * sinT * cp_ + Vv * sinT * sp_ + W * cosT);
                invPdf = 2.0f * kPiF * (1.0f - cosMax);
            } else {
                uint32_t idx = (uint32_t)(rng.uniform() * (L.triCount));
                idx = idx < L.triCount - 1u ? idx : L.triCount - 1u;
                const DevLightTri &T = S.lightTris[L.triOffset + idx];
                float r1 = sqrtf(rng.uniform());
                float r2 = rng.uniform();
                V3 Q = (1.0f - r2) * v3(T.p0) + r2 * v3(T.p1);
                V3 Pp = (1.0f - r1) * v3(T.p2) + r1 * Q;
                V3 wi = Pp - hpl;
                dl = normalize(wi);
                float ca = dot(dl, v3(T.n));
                if (ca < 0.0f) ca *= -1.0f;
                invPdf = (L.totalArea * ca) / sq(length(wi));
            }
            Ld = vector_to_world(dl, L.fwd);
            Li = v3(L.radiance);
            mode = kShadowEntity;
        }
        if (ldist == 0.0f) ldist = 0.001f;
        V3 c;
        bool want = !inside && light_term(M, kd, Li, Ld, D, N, ldist, falloff, invPdf, c);
        if (want) c = had(w, c);
        // A contribution of exactly zero (outside a spot cone, a black
        // material) adds nothing whatever the occlusion: no shadow ray.
        want = want && (c.x != 0.0f || c.y != 0.0f || c.z != 0.0f);
        ShadowRay sr = make_shadow(P + (Ld * eps), Ld, r.time, tmax, c, r.sample, mode, L.entityId);
        sr.srec = rin;
        emit_shadow(O, seg, want, sr);
    }
    if (__ballot(cand != 0u))
        delta_light_pairs(S, O, seg, cand, pairs[threadIdx.x >> 6], sh.material, kd, P, N, D, w, r.time, r.sample, rin);
    // Environment NEE, cosine-distributed (main.cpp:719-778).  With the
    // guide, the direction comes from the guide with probability alpha and
    // from the cosine distribution otherwise, weighted by the mixture pdf
    // (one-sample MIS, balance heuristic; Veach 1997), which keeps the
    // estimator unbiased: the cosine term covers every direction with
    // non-zero contribution.
    if (S.useEnv) {
        const uint2 *lb = GUIDE != kShadeNoGuide && GI.lobes ? GI.lobes + (size_t)i * 8 : nullptr;
        const bool lum = cp.envLuminance != 0u;
        V3 L;
        bool fromGuide = false;
        float pBase = 0.0f;
        if (lb) fromGuide = rng.uniform() < GI.alpha;
        if (fromGuide) {
            float u0 = rng.uniform(), u1 = rng.uniform(), u2 = rng.uniform();
            L = guide_sample(lb, u0, u1, u2);
        } else if (lum) {
            float u0 = rng.uniform(), u1 = rng.uniform();
            L = env_lum_sample(S, u0, u1, pBase);
        } else {
            float phi = 2.0f * kPiF * rng.uniform();
            float cp_, sp_;
            SinCosF(phi, sp_, cp_);
            float theta = Asin(sqrtf(rng.uniform()));
            float ct, st;
            SinCosF(theta, st, ct);
            V3 U, Vv;
            orthonormal_basis(N, U, Vv);
            L = normalize(U * st * cp_ + Vv * st * sp_ + N * ct);
        }
        float eu = (-Atan2F(L.z, L.x) + kPiF) / (2.0f * kPiF);
        float ev = Acos(L.y) / kPiF;
        V3 I = sample_texture(S, S.env, v3(eu, ev, 1.0f), eu, ev);
        float ca = dot(L, N);
        float invPdf;
        if (lb) {
            // One-sample MIS over the guide and the base distribution.
            float base = lum ? (fromGuide ? env_lum_pdf(S, L) : pBase) : fmaxf(ca, 0.0f) / kPiF;
            float pdf = GI.alpha * guide_pdf(lb, L) + (1.0f - GI.alpha) * base;
            invPdf = pdf > 0.0f ? 1.0f / pdf : 0.0f;
        } else if (lum) {
            invPdf = pBase > 0.0f ? 1.0f / pBase : 0.0f;
        } else {
            if (ca == 0.0f) ca = 0.000001f;
            invPdf = kPiF / ca;
        }
        V3 c;
        bool want = light_term(M, kd, I, L, D, N, 1.0f, 1.0f, invPdf, c);
        if (want) c = had(w, c);
        want = want && (c.x != 0.0f || c.y != 0.0f || c.z != 0.0f);
        ShadowRay sr = make_shadow(P + (L * eps), L, r.time, INFINITY, c, r.sample, kShadowAny, 0u);
        sr.srec = rin;
        if (GUIDE != kShadeNoGuide && GI.recs) {
            // Training sample: direction, weight = luminance(L_env) cos / pdf;
            // connect zeroes the weight when the direction is blocked.
            const uint32_t rseg = block % kSegs;
            uint32_t idx = wave_append(GI.recCounters + rseg * kSegStride, want);
            if (want && idx < GI.recSegCap) {
                idx += rseg * GI.recSegCap;
                GuideSample g;
                guide_position(S, P, g.p);
                g.n[0] = N.x;
                g.n[1] = N.y;
                g.n[2] = N.z;
                g.rough = guide_roughness(M);
                g.dir[0] = L.x;
                g.dir[1] = L.y;
                g.dir[2] = L.z;
                g.wgt = guide_lum(I) * fmaxf(ca, 0.0f) * invPdf;
                g.pad = 0.0f;
                GI.recs[idx] = g;
                GI.recKeys[idx] = ((uint64_t)r.sample << GI.keyNodeBits) | r.node;
                sr.rec = idx;
            }
        }
        emit_shadow(O, seg, want, sr);
    }
    const bool childOk = depth + 1u <= S.maxDepth;
    const V3 zero = v3(0.0f, 0.0f, 0.0f);
    // Hemisphere scatter (main.cpp:1024-1083), or, as an option, from the
    // BRDF lobe mixture.  With the scatter guide, the direction comes from
    // the guide with probability alpha and from that distribution otherwise,
    // weighted by the mixture pdf (one-sample MIS, balance heuristic), which
    // keeps the estimator unbiased: the base distribution covers every
    // direction with non-zero BRDF.
    if (S.pathTrace) {
        const uint2 *sl = GUIDE == kShadeAllGuides && GI.scatLobes ? GI.scatLobes + (size_t)i * 8 : nullptr;
        bool fromGuide = false;
        if (sl) fromGuide = rng.uniform() < GI.alpha;
        V3 Sd;
        float invPdf, pBase = 0.0f;
        if (fromGuide) {
            const float u0 = rng.uniform(), u1 = rng.uniform(), u2 = rng.uniform();
            Sd = guide_sample(sl, u0, u1, u2);
        } else if (cp.brdfSampling) {
            const float u0 = rng.uniform(), u1 = rng.uniform(), u2 = rng.uniform();
            Sd = sample_brdf_mixture(M, kd, N, -D, u0, u1, u2, pBase);
        } else {
            float theta = S.importance ? Asin(sqrtf(rng.uniform())) : Acos(rng.uniform());
            float ct, st;
            SinCosF(theta, st, ct);
            float phi = 2.0f * kPiF * rng.uniform();
            float cp_, sp_;
            SinCosF(phi, sp_, cp_);
            V3 U, Vv;
            orthonormal_basis(N, U, Vv);
            Sd = normalize(U * st * cp_ + Vv * st * sp_ + N * ct);
            pBase = S.importance ? fmaxf(dot(Sd, N), 0.0f) / kPiF : 1.0f / (2.0f * kPiF);
        }
        if (sl) {
            if (fromGuide) {
                const float cs = dot(Sd, N);
                pBase = cp.brdfSampling ? brdf_mixture_pdf(M, kd, N, -D, Sd)
                                        : (cs > 0.0f ? (S.importance ? cs / kPiF : 1.0f / (2.0f * kPiF)) : 0.0f);
            }
            const float pdf = GI.alpha * guide_pdf(sl, Sd) + (1.0f - GI.alpha) * pBase;
            invPdf = pdf > 0.0f ? 1.0f / pdf : 0.0f;
        } else if (cp.brdfSampling) {
            invPdf = pBase > 0.0f ? 1.0f / pBase : 0.0f;
        } else if (S.importance) {
            float cs = dot(Sd, N);
            if (cs == 0.0f) cs = 0.000001f;
            invPdf = kPiF / cs;
        } else {
            invPdf = kPiF * 2.0f;
        }
        V3 cw;
        bool want = childOk && invPdf > 0.0f && scatter_weight(M, kd, Sd, D, N, invPdf, cw);
        if (want) cw = had(w, cw);
        uint32_t crec = rin;
        if (GUIDE == kShadeAllGuides && GI.srecs) {
            // Training sample of the scatter guide: direction and
            // cos / (pdf lum(child throughput)); the luminance that the
            // child's subtree delivers to the sample, divided by the child
            // throughput, estimates the incident radiance, and multiplies
            // this weight when the chunk is done.
            const uint32_t rseg = block % kSegs;
            uint32_t idx = wave_append(GI.srecCounters + rseg * kSegStride, want);
            if (want && idx < GI.srecSegCap) {
                idx += rseg * GI.srecSegCap;
                const float lw = guide_lum(cw);
                GuideSample g;
                guide_position(S, P, g.p);
                g.n[0] = N.x;
                g.n[1] = N.y;
                g.n[2] = N.z;
                g.rough = guide_roughness(M);
                g.dir[0] = Sd.x;
                g.dir[1] = Sd.y;
                g.dir[2] = Sd.z;
                g.wgt = lw > 0.0f ? fmaxf(dot(Sd, N), 0.0f) * invPdf / lw : 0.0f;
                g.pad = 0.0f;
                GI.srecs[idx] = g;
                GI.srecKeys[idx] = ((uint64_t)r.sample << GI.keyNodeBits) | r.node;
                GI.links[idx] = ScatterLink{0ull, rin, depth};
                crec = idx;
            }
        }
        emit_child(O, seg, want, make_child(P + (Sd * eps), Sd, r.time, child_node(r.node, 0u), cw, r.sample, zero,
                                       depth + 1u, inside), crec);
    }
    // Ambient (main.cpp:1088-1099).
    const bool transparent = (M.flags & kMatTransparent) != 0u;
    if ((!inside || !transparent) && !ignoreAmbient) {
        V3 am = had(w, had(v3(M.ka), v3(S.ambient)));
        if (am.x != 0.0f || am.y != 0.0f || am.z != 0.0f) {
            add3(a, am);
            credit(GI.links, rin, am);
        }
    }
    // Mirror (main.cpp:781-854).
    if (M.flags & kMatMirror) {
        V3 R = reflect(-D, N);
        if (M.gloss != 0.0f) {
            V3 rp = v3(fabsf(R.x), fabsf(R.y), fabsf(R.z));
            if (rp.x < rp.y) {
                if (rp.x < rp.z) { rp = R; rp.x = rp.x < 0.0f ? -1.0f : 1.0f; }
                else { rp = R; rp.z = rp.z < 0.0f ? -1.0f : 1.0f; }
            } else {
                if (rp.y < rp.z) { rp = R; rp.y = rp.y < 0.0f ? -1.0f : 1.0f; }
                else { rp = R; rp.z = rp.z < 0.0f ? -1.0f : 1.0f; }
            }
            rp = normalize(rp);
            V3 uu = cross(rp, R);
            V3 vv = cross(uu, R);
            float x1 = rng.uniform();
            float x2 = rng.uniform();
            R = R + 2.0f * M.gloss * (uu * (x1 - 0.5f) + vv * (x2 - 0.5f));
            R = normalize(R);
        }
        emit_child(O, seg, childOk, make_child(P + (R * 0.000825f), R, r.time, child_node(r.node, 1u),
                                          had(w, v3(M.km)), r.sample, v3(r.cx, r.cy, r.cz), depth + 1u, inside),
                   rin);
    }
    // Dielectric (main.cpp:856-964): both refraction and reflection are traced.
    if (transparent) {
        V3 tr = v3(M.tr);
        float to = M.ior, from = 1.0f;
        if (inside) {
            from = M.ior;
            to = 1.0f;
        }
        V3 Nn = N;
        float cosT = -dot(D, Nn);
        if (cosT < 0.0f) {
            Nn = -N;
            cosT = -dot(D, Nn);
        }
        float ratio = from / to;
        float delta = 1.0f - sq(ratio) * (1.0f - sq(cosT));
        float refl;
        bool refract = delta > 0.0f;
        PathRay rc{};
        if (!refract) {
            refl = 1.0f;
        } else {
            float cosPhi = sqrtf(delta);
            float cosAir = inside ? cosPhi : cosT;
            float inv = 1.0f - cosAir;
            float r0 = sq((from - to) / (from + to));
            refl = r0 + (1.0f - r0) * PowIF<5>(inv);
            V3 T = (D + Nn * cosT) * ratio - Nn * cosPhi;
            T = normalize(T);
            V3 coef = inside ? zero : v3(tr.x - 1.0f, tr.y - 1.0f, tr.z - 1.0f);
            rc = make_child(P + (T * 0.000100f), T, r.time, child_node(r.node, 2u), w * (1.0f - refl), r.sample, coef,
                            depth + 1u, !inside);
        }
        emit_child(O, seg, childOk && refract, rc, rin);
        V3 R = reflect(-D, Nn);
        V3 coef = inside ? v3(tr.x - 1.0f, tr.y - 1.0f, tr.z - 1.0f) : zero;
        emit_child(O, seg, childOk, make_child(P + (R * 0.000100f), R, r.time, child_node(r.node, 3u), w * refl, r.sample,
                                          coef, depth + 1u, inside),
                   rin);
    }
}

// nDev and cpd, when set, hold the level's ray count and the chunk's
// parameters in device memory (levels replayed from a graph); the grid is
// then sized for the largest count.
template <int GUIDE>
__global__ __launch_bounds__(kRenderBlock, 4) void k_shade(DevScene S, DeviceGeometry G, CameraSetup cam, ChunkParams cp,
                                                        const PathRay *__restrict__ q, SegIn qin,
                                                        const PathHit *__restrict__ hits, uint32_t n,
                                                        AccWord *__restrict__ acc, ShadeOut O, GuideIO GI,
                                                        const uint32_t *__restrict__ nDev,
                                                        const ChunkParams *__restrict__ cpd) {
    if (nDev) {
        n = min(*nDev, n);
        if (blockIdx.x * kRenderBlock >= n) return;
    }
    shade_body<GUIDE>(S, G, cam, cpd ? *cpd : cp, q, qin, hits, n, acc, O, GI, blockIdx.x);
}

// The chunk's parameters and its camera-ray count (the count of level 0),
// for levels replayed from a graph.
__global__ void k_set_chunk(ChunkParams cp, ChunkParams *__restrict__ cpd, uint32_t *__restrict__ total, uint32_t n) {
    if (threadIdx.x == 0) {
        *cpd = cp;
        total[0] = n;
    }
}

// ---- learned guide: queries and training batches ------------------------------------------
__global__ __launch_bounds__(kRenderBlock) void k_guide_query(DevScene S, DeviceGeometry G, const PathRay *__restrict__ q,
                                                              SegIn qin, const PathHit *__restrict__ hits, uint32_t n,
                                                              GuideQuery *__restrict__ out) {
    __shared__ uint32_t soff[kSegs + 1];
    const uint32_t *off = stage_offsets(qin, soff);
    uint32_t i = blockIdx.x * kRenderBlock + threadIdx.x;
    if (i >= n) return;
    const PathRay r = q[seg_slot(off, qin.nseg, qin.segCap, i)];
    const PathHit h = hits[i];
    // Misses and emitters are never guided: roughness -1 makes the
    // inference skip them.
    GuideQuery g;
    memset(&g, 0, sizeof(g));
    g.rough = -1.0f;
    if (h.prim != kEmpty) {
        const DevItemShade sh = S.itemShade[h.item];
        if (!sh.isEmitter) {
            V3 P, N;
            float u, v;
            reconstruct(S, G, r, h, sh, P, N, u, v);
            guide_position(S, P, g.p);
            g.n[0] = N.x;
            g.n[1] = N.y;
            g.n[2] = N.z;
            g.rough = guide_roughness(S.materials[sh.material]);
        }
    }
    out[i] = g;
}

// The environment guide's queries and inference in one pass over a level's
// hits.  Each wave takes 64 rays at a time and gathers, in an LDS list, the
// indices of those that hit a surface that does not emit (misses and
// emitters are never guided); for every 64 gathered it builds their queries
// (one ray per lane, as k_guide_query builds them) into LDS and runs the
// network on them (infer_batch, guide_infer.h), writing each ray's packed
// lobes at its index.  The memory-bound query building of some waves runs
// beside the MFMA work of others, and the queries never go through HBM.
struct alignas(16) FusedInferLds {
    InferLds net;
    GuideQuery q[kInferWaves][kInferSamples];
    uint32_t soff[kSegs + 1];
};

__global__ __launch_bounds__(64 * kInferWaves) void k_guide_query_infer(DevScene S, DeviceGeometry G, GuideWeights W,
                                                                       const PathRay *__restrict__ queue, SegIn qin,
                                                                       const PathHit *__restrict__ hits, uint32_t n,
                                                                       uint2 *__restrict__ lobes) {
    __shared__ FusedInferLds L;
    for (int k = threadIdx.x; k < kHalfCount / 8; k += 64 * kInferWaves)
        reinterpret_cast<uint4 *>(L.net.w)[k] = reinterpret_cast<const uint4 *>(W.wf)[k];
    for (int k = threadIdx.x; k < 5 * kGuideBiasStride; k += 64 * kInferWaves) L.net.bias[k] = W.bias[k];
    const uint32_t *off = nullptr;
    if (qin.nseg != 1u) {
        for (uint32_t k = threadIdx.x; k <= qin.nseg; k += blockDim.x) L.soff[k] = qin.offsets[k];
        off = L.soff;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t chunks = (n + kInferSamples - 1) / kInferSamples;
    const uint32_t first = blockIdx.x * kInferWaves + wave, stride = gridDim.x * kInferWaves;
    uint32_t *list = L.net.idx[wave];
    GuideQuery *qs = L.q[wave];
    uint32_t pending = 0;                            // wave-uniform: gathered indices not yet run
    for (uint32_t it = first;; it += stride) {
        const bool more = it < chunks;
        if (more) {
            const uint32_t s = it * kInferSamples + lane;
            bool valid = false;
            if (s < n) {
                const PathHit h = hits[s];
                valid = h.prim != kEmpty && !S.itemShade[h.item].isEmitter;
            }
            const uint64_t m = __ballot(valid);
            if (valid)
                list[pending + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = s;
            pending += (uint32_t)__popcll(m);
            wave_lds_sync();
        }
        if (pending >= (uint32_t)kInferSamples || (!more && pending > 0u)) {
            const uint32_t cnt = min(pending, (uint32_t)kInferSamples);
            if ((uint32_t)lane < cnt) {
                const uint32_t s = list[lane];
                const PathRay r = queue[seg_slot(off, qin.nseg, qin.segCap, s)];
                const PathHit h = hits[s];
                const DevItemShade sh = S.itemShade[h.item];
                V3 P, N;
                float u, v;
                reconstruct(S, G, r, h, sh, P, N, u, v);
                GuideQuery g;
                guide_position(S, P, g.p);
                g.n[0] = N.x;
                g.n[1] = N.y;
                g.n[2] = N.z;
                g.rough = guide_roughness(S.materials[sh.material]);
                g.pad = 0.0f;
                qs[lane] = g;
            }
            wave_lds_sync();
            infer_batch<true>(L.net, [qs](uint32_t j) { return qs[j]; }, cnt, [list](uint32_t j) { return list[j]; },
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

// Training batches are drawn in proportion to the records' weights (a
// record of weight w stands for w times the loss -log q of its direction),
// each drawn sample then carrying weight 1: the same gradient in
// expectation, with far less variance when a few records (directions
// towards the sun) carry most of the weight; records of weight zero
// (occluded, or whose path delivered nothing) are never drawn.  The prefix
// of the weights in dense record order, in three passes (block scans, one
// block over the block sums, offsets):
// The records of a chunk in the order of their keys (sample, then random-
// stream node), so that the weighted draws below, and the guide trained on
// them, do not depend on the order in which waves appended the records:
// the dense index's slot and key, for a radix sort of the pairs.
__global__ __launch_bounds__(kRenderBlock) void k_rec_keys(const uint64_t *__restrict__ keys, SegIn rin,
                                                           uint32_t count, uint64_t *__restrict__ kout,
                                                           uint32_t *__restrict__ sout) {
    __shared__ uint32_t soff[kSegs + 1];
    const uint32_t *off = stage_offsets(rin, soff);
    const uint32_t i = blockIdx.x * kRenderBlock + threadIdx.x;
    if (i >= count) return;
    const uint32_t slot = seg_slot(off, rin.nseg, rin.segCap, i);
    kout[i] = keys[slot];
    sout[i] = slot;
}

__global__ __launch_bounds__(kRenderBlock) void k_rec_prefix_blocks(const GuideSample *__restrict__ recs,
                                                                    const uint32_t *__restrict__ order,
                                                                    uint32_t count, float *__restrict__ cum,
                                                                    float *__restrict__ sums) {
    __shared__ float s[kRenderBlock];
    const uint32_t i = blockIdx.x * kRenderBlock + threadIdx.x;
    float w = 0.0f;
    if (i < count) {
        w = recs[order[i]].wgt;
        w = w > 0.0f && w < INFINITY ? w : 0.0f;          // NaN, negative and infinite weights are not drawn
    }
    s[threadIdx.x] = w;
    __syncthreads();
    for (uint32_t o = 1; o < kRenderBlock; o <<= 1) {
        const float a = threadIdx.x >= o ? s[threadIdx.x - o] : 0.0f;
        __syncthreads();
        s[threadIdx.x] += a;
        __syncthreads();
    }
    if (i < count) cum[i] = s[threadIdx.x];
    if (threadIdx.x == kRenderBlock - 1) sums[blockIdx.x] = s[threadIdx.x];
}

__global__ __launch_bounds__(kRenderBlock) void k_rec_prefix_sums(float *__restrict__ sums, uint32_t nb) {
    __shared__ float s[kRenderBlock];
    float carry = 0.0f;
    for (uint32_t b0 = 0; b0 < nb; b0 += kRenderBlock) {
        const uint32_t i = b0 + threadIdx.x;
        const float v = i < nb ? sums[i] : 0.0f;
        s[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t o = 1; o < kRenderBlock; o <<= 1) {
            const float a = threadIdx.x >= o ? s[threadIdx.x - o] : 0.0f;
            __syncthreads();
            s[threadIdx.x] += a;
            __syncthreads();
        }
        if (i < nb) sums[i] = carry + s[threadIdx.x] - v;   // exclusive
        carry += s[kRenderBlock - 1];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kRenderBlock) void k_rec_prefix_add(float *__restrict__ cum, uint32_t count,
                                                                 const float *__restrict__ sums) {
    const uint32_t i = blockIdx.x * kRenderBlock + threadIdx.x;
    if (i < count) cum[i] += sums[blockIdx.x];
}

// A training batch of n records drawn by weight from the count recorded
// ones (hash of (seed, slot), splitmix64 finalizer, then the first record
// whose prefix exceeds u times the total); all weights zero: drawn
// uniformly with their weights.
__global__ __launch_bounds__(kRenderBlock) void k_guide_gather_weighted(const GuideSample *__restrict__ recs,
                                                                        const uint32_t *__restrict__ order,
                                                                        uint32_t count, const float *__restrict__ cum,
                                                                        GuideSample *__restrict__ batch, uint32_t n,
                                                                        uint64_t seed) {
    uint32_t i = blockIdx.x * kRenderBlock + threadIdx.x;
    if (i >= n) return;
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const float total = cum[count - 1];
    if (!(total > 0.0f)) {
        batch[i] = recs[order[(uint32_t)(z % count)]];
        return;
    }
    const float target = (float)(z >> 40) * 0x1p-24f * total;
    uint32_t lo = 0u, hi = count - 1u;                      // first j with cum[j] > target
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cum[mid] > target) hi = mid;
        else lo = mid + 1u;
    }
    GuideSample g = recs[order[lo]];
    g.wgt = 1.0f;
    batch[i] = g;
}

// Scatter guide records of depth `depth` pass the luminance their subtree
// delivered on to their parent record.  Launched for depth = maxDepth - 1
// down to 1 after a chunk's paths are done, so that a record holds its whole
// subtree's sum when it passes it on (its descendants are deeper; no record
// of this depth is a parent of another).
__global__ __launch_bounds__(kRenderBlock) void k_scatter_propagate(ScatterLink *__restrict__ links, SegIn rin,
                                                                    uint32_t count, uint32_t depth) {
    __shared__ uint32_t soff[kSegs + 1];
    const uint32_t *off = stage_offsets(rin, soff);
    const uint32_t i = blockIdx.x * kRenderBlock + threadIdx.x;
    if (i >= count) return;
    const ScatterLink L = links[seg_slot(off, rin.nseg, rin.segCap, i)];
    if (L.depth == depth && L.parent != kEmpty && L.sum != 0ull) atomicAdd(&links[L.parent].sum, L.sum);
}

// Training weight of each scatter record: its factor times the luminance its
// subtree delivered.
__global__ __launch_bounds__(kRenderBlock) void k_scatter_finalize(GuideSample *__restrict__ recs,
                                                                   const ScatterLink *__restrict__ links, SegIn rin,
                                                                   uint32_t count) {
    __shared__ uint32_t soff[kSegs + 1];
    const uint32_t *off = stage_offsets(rin, soff);
    const uint32_t i = blockIdx.x * kRenderBlock + threadIdx.x;
    if (i >= count) return;
    const uint32_t slot = seg_slot(off, rin.nseg, rin.segCap, i);
    recs[slot].wgt *= acc_value(links[slot].sum);
}

// Segment counts -> offsets (one block per queue, one thread per segment):
// offsets[s] = sum of the clamped counts of segments < s, offsets[nseg] and
// *total = their sum.  The counters are cleared for the next producer.
struct SegScan {
    uint32_t *counters[2];
    uint32_t *offsets[2];
    uint32_t segCap[2];
    uint32_t nseg[2];
    uint32_t *total;                 // total[q] for queue q
};

__device__ __forceinline__ void seg_scan_body(const SegScan &P, uint32_t q) {
    __shared__ uint32_t part[kSegs / 64];
    const uint32_t t = threadIdx.x, nseg = P.nseg[q];
    uint32_t c = 0u;
    if (t < nseg) {
        c = min(P.counters[q][t * kSegStride], P.segCap[q]);
        P.counters[q][t * kSegStride] = 0u;
    }
    // Inclusive scan within the wave, then across the block's waves.
    const uint32_t lane = t & 63u, wave = t >> 6;
    uint32_t x = c;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63u) part[wave] = x;
    __syncthreads();
    uint32_t before = 0u;
    for (uint32_t w = 0; w < wave; ++w) before += part[w];
    const uint32_t incl = before + x;
    if (t < nseg) P.offsets[q][t] = incl - c;
    if (t == nseg - 1u) {
        P.offsets[q][nseg] = incl;
        P.total[q] = incl;
    }
}

__global__ __launch_bounds__(kSegs) void k_seg_scan(SegScan P) { seg_scan_body(P, blockIdx.x); }

// Clears n words (a fill kernel rather than a memset node, so that the same
// code can be captured in a graph).
__global__ void k_clear_u32(uint32_t *p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---- splat (main.cpp:1209-1254) -------------------------------------------------------
template <typename T>
__device__ __forceinline__ void splat_box_body(const CameraSetup &cam, const ChunkParams &cp, const AccWord *__restrict__ acc,
                                      T *__restrict__ px, uint32_t block) {
    uint32_t p = block * kRenderBlock + threadIdx.x;
    if (p >= cp.rows * cam.width) return;
    uint32_t pixel = cp.row0 * cam.width + p;
    T cr = px[4 * pixel], cg = px[4 * pixel + 1], cb = px[4 * pixel + 2], cw = px[4 * pixel + 3];
    for (uint32_t k = 0; k < cp.spp; ++k) {
        const AccWord *c = acc + 3 * ((size_t)p * cp.spp + k);
        cr += (T)acc_value(c[0]);
        cg += (T)acc_value(c[1]);
        cb += (T)acc_value(c[2]);
        cw += (T)1.0f;
    }
    px[4 * pixel] = cr;
    px[4 * pixel + 1] = cg;
    px[4 * pixel + 2] = cb;
    px[4 * pixel + 3] = cw;
}

template <typename T>
__global__ __launch_bounds__(kRenderBlock) void k_splat_box(CameraSetup cam, ChunkParams cp, const AccWord *__restrict__ acc,
                                                            T *__restrict__ px) {
    splat_box_body<T>(cam, cp, acc, px, blockIdx.x);
}

// Gaussian, in scatter form over tiles: a block's 256 threads are a 16x16
// tile of source pixels; each reads its samples of the chunk once and sums
// them, weighted, for each of its 3x3 destination neighbours into LDS; then
// each of the tile's inner 14x14 pixels adds the nine partial sums aimed at
// it.  Tiles step by 14 pixels, so every destination is an inner pixel of
// exactly one tile.  The filter radius (1.5: one pixel each way) is fixed.
constexpr uint32_t kSplatSrc = 16, kSplatDst = kSplatSrc - 2;
static_assert(kSplatSrc * kSplatSrc == kRenderBlock, "one source pixel per thread");

__host__ __device__ inline uint32_t splat_gauss_blocks(uint32_t width, uint32_t dstRows) {
    return ((width + kSplatDst - 1) / kSplatDst) * ((dstRows + kSplatDst - 1) / kSplatDst);
}

template <typename T>
__device__ __forceinline__ void splat_gauss_body(const CameraSetup &cam, const ChunkParams &cp, const AccWord *__restrict__ acc,
                                        T *__restrict__ px, uint32_t dstRow0, uint32_t dstRows, uint32_t block) {
    __shared__ T part[kRenderBlock][9][4];
    const uint32_t tilesX = (cam.width + kSplatDst - 1) / kSplatDst;
    const int32_t lx = (int32_t)(threadIdx.x % kSplatSrc), ly = (int32_t)(threadIdx.x / kSplatSrc);
    const int32_t col = (int32_t)((block % tilesX) * kSplatDst) - 1 + lx;
    const int32_t row = (int32_t)(dstRow0 + (block / tilesX) * kSplatDst) - 1 + ly;
    const int32_t r0 = (int32_t)cp.row0, r1 = (int32_t)(cp.row0 + cp.rows) - 1;
    T sum[9][4];
#pragma unroll
    for (int j = 0; j < 9; ++j) sum[j][0] = sum[j][1] = sum[j][2] = sum[j][3] = (T)0.0f;
    if (row >= r0 && row <= r1 && col >= 0 && col < (int32_t)cam.width) {
        const AccWord *c = acc + 3 * ((size_t)((uint32_t)(row - r0) * cam.width + (uint32_t)col) * cp.spp);
        for (uint32_t k = 0; k < cp.spp; ++k) {
            const uint32_t s = cp.s0 + k;
            const uint32_t sx = s % cam.sppSqrt, sy = s / cam.sppSqrt;
            const float scx = col + (0.5f * cam.invSppSqrt + sx * cam.invSppSqrt);
            const float scy = row + (0.5f * cam.invSppSqrt + sy * cam.invSppSqrt);
            const float cr = acc_value(c[3 * k]), cg = acc_value(c[3 * k + 1]), cb = acc_value(c[3 * k + 2]);
#pragma unroll
            for (int oy = -1; oy <= 1; ++oy)
#pragma unroll
                for (int ox = -1; ox <= 1; ++ox) {
                    const float dx = scx - ((col + ox) + 0.5f), dy = scy - ((row + oy) + 0.5f);
                    const float g = cam.oneOverTwoPiSigma2 * Exp(cam.negOneOverTwoSigma2 * (sq(dx) + sq(dy)));
                    T *a = sum[(oy + 1) * 3 + (ox + 1)];
                    a[0] += (T)(cr * g);
                    a[1] += (T)(cg * g);
                    a[2] += (T)(cb * g);
                    a[3] += (T)g;
                }
        }
    }
#pragma unroll
    for (int j = 0; j < 9; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) part[threadIdx.x][j][q] = sum[j][q];
    __syncthreads();
    if (lx < 1 || ly < 1 || lx > (int32_t)kSplatDst || ly > (int32_t)kSplatDst) return;
    if (col >= (int32_t)cam.width || row >= (int32_t)(dstRow0 + dstRows)) return;
    const uint32_t pixel = (uint32_t)row * cam.width + (uint32_t)col;
    T o[4] = {px[4 * pixel], px[4 * pixel + 1], px[4 * pixel + 2], px[4 * pixel + 3]};
    // Source (col - ox, row - oy) sent its offset-(ox, oy) sum here.
#pragma unroll
    for (int oy = -1; oy <= 1; ++oy)
#pragma unroll
        for (int ox = -1; ox <= 1; ++ox) {
            const T *a = part[(ly - oy) * (int32_t)kSplatSrc + (lx - ox)][(oy + 1) * 3 + (ox + 1)];
            o[0] += a[0];
            o[1] += a[1];
            o[2] += a[2];
            o[3] += a[3];
        }
    px[4 * pixel] = o[0];
    px[4 * pixel + 1] = o[1];
    px[4 * pixel + 2] = o[2];
    px[4 * pixel + 3] = o[3];
}

template <typename T>
__global__ __launch_bounds__(kRenderBlock) void k_splat_gauss(CameraSetup cam, ChunkParams cp,
                                                              const AccWord *__restrict__ acc, T *__restrict__ px,
                                                              uint32_t dstRow0, uint32_t dstRows) {
    splat_gauss_body<T>(cam, cp, acc, px, dstRow0, dstRows, blockIdx.x);
}

// ---- batched launches: several scenes per launch ---------------------------------------
// The blocks of a batched launch are split into consecutive ranges, one per
// scene (BatchMap.start), so each block works on one scene; the per-scene
// arguments live in device memory, read with scalar loads.
constexpr uint32_t kMaxBatch = 32;

struct BatchMap {
    uint32_t count;
    uint32_t start[kMaxBatch + 1];
};

__device__ inline uint32_t batch_scene(const BatchMap &M, uint32_t b) {
    uint32_t s = 0;
    while (s + 1 < M.count && b >= M.start[s + 1]) ++s;
    return s;
}

struct GenArgs {
    DevScene S;
    CameraSetup cam;
    ChunkParams cp;
    PathRay *q;
    AccWord *acc;
    uint32_t n;
};

struct ExtArgs {
    DeviceGeometry G;
    const PathRay *q;
    SegIn qin;
    uint32_t n;
    PathHit *hits;
    float eps;
    int cull;
};

struct ShadeArgs {
    DevScene S;
    DeviceGeometry G;
    CameraSetup cam;
    ChunkParams cp;
    const PathRay *q;
    SegIn qin;
    const PathHit *hits;
    uint32_t n;
    AccWord *acc;
    ShadeOut O;
    GuideIO GI;
};

struct ConnArgs {
    DeviceGeometry G;
    const ShadowRay *sq;
    SegIn sin;
    uint32_t n;
    AccWord *acc;
    GuideSample *recs;
};

struct SplatArgs {
    CameraSetup cam;
    ChunkParams cp;
    const AccWord *acc;
    void *px;
    uint32_t dstRow0, dstRows;
};

__global__ __launch_bounds__(kRenderBlock) void kb_generate(const GenArgs *__restrict__ A, BatchMap M) {
    const uint32_t s = batch_scene(M, blockIdx.x);
    const GenArgs &a = A[s];
    generate_body(a.S, a.cam, a.cp, a.q, a.acc, a.n, blockIdx.x - M.start[s]);
}

template <int BLOCK, int WAVES, bool WW>
__global__ __launch_bounds__(BLOCK, WAVES) void kb_extend(const ExtArgs *__restrict__ A, BatchMap M) {
    const uint32_t s = batch_scene(M, blockIdx.x);
    const ExtArgs &a = A[s];
    extend_body<BLOCK, WW>(a.G, a.q, a.qin, a.n, a.hits, a.eps, a.cull, 64u * kRaysPerLane, blockIdx.x - M.start[s]);
}

__global__ __launch_bounds__(kRenderBlock, 4) void kb_shade(const ShadeArgs *__restrict__ A, BatchMap M) {
    const uint32_t s = batch_scene(M, blockIdx.x);
    const ShadeArgs &a = A[s];
    shade_body<kShadeNoGuide>(a.S, a.G, a.cam, a.cp, a.q, a.qin, a.hits, a.n, a.acc, a.O, a.GI, blockIdx.x - M.start[s]);
}

__global__ __launch_bounds__(kSegs) void kb_seg_scan(const SegScan *__restrict__ P) {
    seg_scan_body(P[blockIdx.x >> 1], blockIdx.x & 1u);
}

template <int BLOCK, int WAVES, bool WW>
__global__ __launch_bounds__(BLOCK, WAVES) void kb_connect(const ConnArgs *__restrict__ A, BatchMap M) {
    const uint32_t s = batch_scene(M, blockIdx.x);
    const ConnArgs &a = A[s];
    connect_body<BLOCK, WW>(a.G, a.sq, a.sin, a.n, a.acc, a.recs, nullptr, 64u * kRaysPerLane, blockIdx.x - M.start[s]);
}

template <typename T>
__global__ __launch_bounds__(kRenderBlock) void kb_splat(const SplatArgs *__restrict__ A, BatchMap M) {
    const uint32_t s = batch_scene(M, blockIdx.x);
    const SplatArgs &a = A[s];
    if (a.cp.gaussian)
        splat_gauss_body<T>(a.cam, a.cp, a.acc, (T *)a.px, a.dstRow0, a.dstRows, blockIdx.x - M.start[s]);
    else
        splat_box_body<T>(a.cam, a.cp, a.acc, (T *)a.px, blockIdx.x - M.start[s]);
}

static hipError_t launch_extend_batch(const RenderTuning &T, const ExtArgs *A, const BatchMap &M, uint32_t blocks,
                                      hipStream_t st) {
#define AZ_CASE(B, W, F)                                                                                  \
    if (T.traceBlock == B && T.traceWaves == W && T.traceLoop == F) {                                   \
        hipLaunchKernelGGL((kb_extend<B, W, F != 0>), dim3(blocks), dim3(B), 0, st, A, M);             \
        return hipGetLastError();                                                                        \
    }
    AZ_RENDER_TRACE_VARIANTS(AZ_CASE)
#undef AZ_CASE
    return hipErrorInvalidValue;
}

static hipError_t launch_connect_batch(const RenderTuning &T, const ConnArgs *A, const BatchMap &M, uint32_t blocks,
                                       hipStream_t st) {
#define AZ_CASE(B, W, F)                                                                                  \
    if (T.traceBlock == B && T.traceWaves == W && T.traceLoop == F) {                                   \
        hipLaunchKernelGGL((kb_connect<B, W, F != 0>), dim3(blocks), dim3(B), 0, st, A, M);            \
        return hipGetLastError();                                                                        \
    }
    AZ_RENDER_TRACE_VARIANTS(AZ_CASE)
#undef AZ_CASE
    return hipErrorInvalidValue;
}

}  // namespace az

using namespace az;

// ---- host side ---------------------------------------------------------------------------
namespace {

#define AZ_HIPR(ctx, call)                                                  \
    do {                                                                    \
        hipError_t _e = (call);                                             \
        if (_e != hipSuccess) {                                             \
            (ctx)->lastError = std::string(#call ": ") + hipGetErrorString(_e); \
            if (_e == hipErrorOutOfMemory) (void)hipGetLastError();        \
            return _e == hipErrorOutOfMemory ? AZ_ERR_OUT_OF_MEMORY : AZ_ERR_DEVICE; \
        }                                                                   \
    } while (0)

// PCG32 jump table: entry k holds mult^(2^k) and sum_{i<2^k} mult^i.
// PCG32 jump-ahead table: entry 256 k + v advances a stream by v << 8k steps
// as (multiplier, increment factor): state' = mult state + inc factor.  Jumps
// of one generator commute, so make_rng composes one entry per non-zero byte
// of its distance (the same state as stepping one by one).
std::vector<uint64_t> pcgTable() {
    uint64_t bm[64], bg[64];                 // jumps by 2^b steps
    uint64_t m = kPcgMult, g = 1u;
    for (int b = 0; b < 64; ++b) {
        bm[b] = m;
        bg[b] = g;
        g = (m + 1u) * g;
        m = m * m;
    }
    std::vector<uint64_t> t(2 * 8 * 256);
    for (int k = 0; k < 8; ++k)
        for (int v = 0; v < 256; ++v) {
            uint64_t M = 1u, G = 0u;
            for (int b = 0; b < 8; ++b)
                if ((v >> b) & 1) {
                    G = G * bm[8 * k + b] + bg[8 * k + b];
                    M = M * bm[8 * k + b];
                }
            t[2 * (256 * k + v)] = M;
            t[2 * (256 * k + v) + 1] = G;
        }
    return t;
}

// Perlin gradient table drawn from PCG32(0x5045524C, 256), three draws per
// entry, mapped to [-1,1) and normalized.
std::vector<float> perlinTable() {
    uint64_t state = 0u, inc = (256u << 1u) | 1u;
    auto next = [&]() {
        uint64_t old = state;
        state = old * kPcgMult + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    };
    next();
    state += 0x5045524CULL;
    next();
    auto uni = [&]() { return (float)(next() >> 8) * (1.0f / 16777216.0f); };
    std::vector<float> g(3 * 256);
    for (int i = 0; i < 256; ++i) {
        float a = uni(), b = uni(), c = uni();
        float x = (a - 0.5f) * 2.0f, y = (b - 0.5f) * 2.0f, z = (c - 0.5f) * 2.0f;
        float l = sqrtf(x * x + y * y + z * z);
        if (l != 0.0f) {
            x /= l;
            y /= l;
            z /= l;
        } else {
            x = y = z = 0.0f;
        }
        g[3 * i] = x;
        g[3 * i + 1] = y;
        g[3 * i + 2] = z;
    }
    return g;
}

Affine affineOf(const mat4 &m) {
    Affine a;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 4; ++j) a.m[i][j] = m.E[i][j];
    return a;
}

inline uint32_t bitsOf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

DevTexture devTexture(const az_texture &t, std::vector<uint8_t> &texels, std::vector<uint32_t> &perms) {
    DevTexture d;
    memset(&d, 0, sizeof(d));
    d.type = (int32_t)t.Type;
    d.bumpMul = t.BumpMapMultiplier;
    d.bump = t.BumpMap ? 1 : 0;
    d.wrap = (int32_t)t.WrapMode;
    d.blend = (int32_t)t.BlendMode;
    d.filter = (int32_t)t.Filter;
    if (t.Type == TextureType_2D) {
        d.pixelFormat = (int32_t)t.PixelFormat;
        d.width = (int32_t)t.Width;
        d.height = (int32_t)t.Height;
        d.invW = t.InvWidth;
        d.invH = t.InvHeight;
        d.normalizer = t.Normalizer;
        size_t texel = t.PixelFormat == PixelFormat_RGB_F32 ? 12 : 3;
        size_t bytes = (size_t)t.Width * t.Height * texel;
        while (texels.size() % 16) texels.push_back(0);
        d.dataOffset = (uint32_t)texels.size();
        if (t.Pixels && bytes) texels.insert(texels.end(), (const uint8_t *)t.Pixels, (const uint8_t *)t.Pixels + bytes);
        else texels.resize(texels.size() + std::max<size_t>(bytes, 12), 0);
    } else {
        d.perlinAppearance = (int32_t)t.PerlinAppearance;
        d.scaling = t.ScalingFactor;
        d.permOffset = (uint32_t)perms.size();
        for (int i = 0; i < 256; ++i) perms.push_back(t.Perm[i] & 255u);
    }
    return d;
}

template <typename T>
hipError_t uploadVec(DeviceBuffer &b, const std::vector<T> &v, hipStream_t s) {
    size_t n = std::max<size_t>(1, v.size()) * sizeof(T);
    hipError_t e = b.alloc(n);
    if (e != hipSuccess) return e;
    if (!v.empty()) e = hipMemcpyAsync(b.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
    return e;
}

}  // namespace

struct RenderBuffers {
    // Cap on the render queues' device memory (az_gpu_set_render_memory_limit);
    // declared first so that it outlives the buffers it counts.
    AllocLimit queueLimit;
    DeviceBuffer materials, textures, texels, perms, lights, lightTris, prims, itemShade, perlin, pcg;
    DeviceBuffer q0, q1, hits, shadow, acc, counters, pixels;
    DeviceBuffer queries, lobes, recs, batch;   // learned guide
    DeviceBuffer scatLobes, srecs, links, qrec0, qrec1;   // scatter guide: lobes, records, per-slot record of the queues
    DeviceBuffer recCum;                        // prefix of the training records' weights
    DeviceBuffer recKeys, srecKeys;             // per record slot: its path vertex, the draw order
    DeviceBuffer recSort, recSortTmp;           // key-ordered records: keys, slots, radix sort workspace
    DeviceBuffer workCounters;                  // counted renders: extend and connect walk counts (6 x u64)
    DeviceBuffer envCdf;                        // environment luminance distribution
    DeviceBuffer batchArgs, batchTotals;        // batched renders: per-scene arguments, queue totals
    // Levels replayed from a graph: the chunk's parameters and per-level
    // counts in device memory, and two plans (a render's full chunks and its
    // last, smaller one): the level bounds learned on a host-driven chunk,
    // what they were learned on, and the instantiated graph with the launch
    // arguments it was recorded with.
    struct LevelPlan {
        std::vector<uint8_t> boundKey, graphKey;
        std::vector<uint32_t> rayBound, shadowBound;   // per level: rays and shadow rays a level holds
        hipGraphExec_t exec = nullptr;
        uint64_t used = 0;
        void dropGraph() {
            if (exec) (void)hipGraphExecDestroy(exec);
            exec = nullptr;
            graphKey.clear();
        }
    };
    DeviceBuffer chunkDev, levelTotals;
    LevelPlan plans[2];
    uint64_t planClock = 0;
    uint64_t levelCaptures = 0, levelLaunches = 0;
    void dropLevelGraph() {
        for (LevelPlan &p : plans) {
            p.dropGraph();
            p = LevelPlan{};
        }
    }
    // The per-render buffers: what az_gpu_trim frees and the memory cap counts.
    template <typename F>
    void forEachQueue(F f) {
        DeviceBuffer *bufs[] = {&q0,    &q1,        &hits,      &shadow, &acc,   &counters, &pixels,  &queries,
                                &lobes, &recs,      &batch,     &scatLobes, &srecs, &links, &qrec0,  &qrec1,
                                &recCum, &batchArgs, &batchTotals, &chunkDev, &levelTotals, &workCounters,
                                &recKeys, &srecKeys, &recSort, &recSortTmp};
        for (DeviceBuffer *b : bufs) f(*b);
    }
    RenderBuffers() {
        forEachQueue([this](DeviceBuffer &b) { b.lim = &queueLimit; });
    }
    // A chunk that did not fit gives back its sample-sized queues before the
    // render retries with half the samples per pixel, so that the smaller
    // chunk is not squeezed by the larger one's allocations.
    void releaseChunkQueues() {
        DeviceBuffer *bufs[] = {&q0, &q1, &hits, &shadow, &acc, &qrec0, &qrec1};
        for (DeviceBuffer *b : bufs) b->release();
    }
    ~RenderBuffers() { dropLevelGraph(); }
};

// Row and column CDFs of the environment map's luminance over its (u, v)
// cells: cell (c, r) weighs the largest texel luminance among texels
// c-1..c+1, r-1..r+1 (so that it covers what nearest and bilinear lookups
// read there), plus 1e-4 of the mean as a floor (every direction stays
// reachable), times sin(theta) at the row centre.
static std::vector<float> envLuminanceCdf(const az_texture &t) {
    const int W = (int)t.Width, H = (int)t.Height;
    const bool f32 = t.PixelFormat == PixelFormat_RGB_F32;
    std::vector<double> lum((size_t)W * H);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            size_t i = ((size_t)y * W + x) * 3;
            double r, g, b;
            if (f32) {
                const float *p = (const float *)t.Pixels;
                r = p[i], g = p[i + 1], b = p[i + 2];
            } else {
                const uint8_t *p = (const uint8_t *)t.Pixels;
                r = p[i], g = p[i + 1], b = p[i + 2];
            }
            lum[(size_t)y * W + x] = std::fabs(0.2126 * r + 0.7152 * g + 0.0722 * b);
        }
    std::vector<double> cell((size_t)W * H);
    double total = 0.0;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double m = 0.0;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int yy = std::min(std::max(y + dy, 0), H - 1), xx = (x + dx + W) % W;
                    m = std::max(m, lum[(size_t)yy * W + xx]);
                }
            cell[(size_t)y * W + x] = m;
            total += m;
        }
    const double floorv = 1e-4 * std::max(total / ((double)W * H), 1e-30);
    std::vector<float> out((size_t)(H + 1) + (size_t)H * (W + 1));
    std::vector<double> rowMass(H);
    for (int y = 0; y < H; ++y) {
        const double st = std::sin(kPiD * (y + 0.5) / H);
        double acc = 0.0;
        float *col = out.data() + (H + 1) + (size_t)y * (W + 1);
        col[0] = 0.0f;
        std::vector<double> run(W + 1, 0.0);
        for (int x = 0; x < W; ++x) {
            acc += (cell[(size_t)y * W + x] + floorv) * st;
            run[x + 1] = acc;
        }
        for (int x = 1; x <= W; ++x) col[x] = (float)(run[x] / acc);
        col[W] = 1.0f;
        rowMass[y] = acc;
    }
    double acc = 0.0;
    out[0] = 0.0f;
    double rt = 0.0;
    for (int y = 0; y < H; ++y) rt += rowMass[y];
    for (int y = 0; y < H; ++y) {
        acc += rowMass[y];
        out[y + 1] = (float)(acc / rt);
    }
    out[H] = 1.0f;
    return out;
}

static RenderBuffers &buffersOf(az_gpu_ctx *C) {
    // One set of buffers per context, owned by the context's render state.
    static_assert(sizeof(RenderState) >= 1, "");
    if (!C->renderBuffers) C->renderBuffers = new RenderBuffers();
    return *reinterpret_cast<RenderBuffers *>(C->renderBuffers);
}

void az_release_render_buffers(void *p) { delete reinterpret_cast<RenderBuffers *>(p); }

az_status az_upload_shading(az_gpu_ctx *C, const scene *S, const UploadInfo &info) {
    RenderBuffers &B = buffersOf(C);
    DevScene &D = C->devScene;
    memset(&D, 0, sizeof(D));
    // Materials.
    std::vector<DevMaterial> mats(std::max<u32>(1, S->MaterialCount));
    for (u32 i = 0; i < S->MaterialCount; ++i) {
        const material &m = S->Materials[i];
        DevMaterial &d = mats[i];
        memset(&d, 0, sizeof(d));
        const vec3 *src[5] = {&m.DiffuseReflectance, &m.SpecularReflectance, &m.AmbientReflectance,
                              &m.MirrorReflectance, &m.Transparency};
        float *dst[5] = {d.kd, d.ks, d.ka, d.km, d.tr};
        for (int k = 0; k < 5; ++k) {
            dst[k][0] = src[k]->x;
            dst[k][1] = src[k]->y;
            dst[k][2] = src[k]->z;
        }
        d.ior = m.RefractionIndex;
        d.gloss = m.Glossiness;
        d.exponent = m.BRDF.Exponent;
        d.brdfIor = m.BRDF.RefractiveIndex;
        d.brdfType = (int32_t)m.BRDF.Type;
        if (bitsOf(m.Transparency.x) | bitsOf(m.Transparency.y) | bitsOf(m.Transparency.z)) d.flags |= kMatTransparent;
        if (bitsOf(m.MirrorReflectance.x) | bitsOf(m.MirrorReflectance.y) | bitsOf(m.MirrorReflectance.z))
            d.flags |= kMatMirror;
    }
    // Children per shading point: the hemisphere scatter when path tracing,
    // a mirror ray, and a dielectric's refraction and reflection.
    u32 matChildren = 0;
    for (const DevMaterial &d : mats)
        matChildren = std::max<u32>(matChildren, ((d.flags & kMatMirror) ? 1u : 0u) +
                                                     ((d.flags & kMatTransparent) ? 2u : 0u));
    D.maxChildren = std::max<u32>(1u, (S->PathTrace ? 1u : 0u) + matChildren);
    // Lights that shading samples as entity lights (every type but point,
    // area, directional and spot): their shadow rays take a closest hit.
    D.entityLights = 0u;
    for (u32 i = 0; i < S->LightCount; ++i) {
        const int32_t t = (int32_t)S->Lights[i].Type;
        D.entityLights += t != LightType_PointLight && t != LightType_AreaLight && t != LightType_DirectionalLight &&
                                  t != LightType_SpotLight
                              ? 1u
                              : 0u;
    }
    // Textures (index 0 unused: entities use 1-based indices).
    std::vector<uint8_t> texels;
    std::vector<uint32_t> perms;
    std::vector<DevTexture> texs(1 + S->TextureCount);
    memset(texs.data(), 0, sizeof(DevTexture));
    for (u32 i = 0; i < S->TextureCount; ++i) texs[1 + i] = devTexture(S->Textures[i], texels, perms);
    if (S->UseEnvironmentMap) D.env = devTexture(S->EnvironmentMap, texels, perms);
    if (S->UseBackgroundTexture) D.background = devTexture(S->BackgroundTexture, texels, perms);
    if (texels.empty()) texels.resize(16, 0);
    if (perms.empty()) perms.resize(1, 0);
    // Lights.
    std::vector<DevLight> lights(std::max<u32>(1, S->LightCount));
    std::vector<DevLightTri> ltris;
    for (u32 i = 0; i < S->LightCount; ++i) {
        const light &L = S->Lights[i];
        DevLight &d = lights[i];
        memset(&d, 0, sizeof(d));
        d.type = (int32_t)L.Type;
        auto cp3 = [](float *o, const vec3 &v) {
            o[0] = v.x;
            o[1] = v.y;
            o[2] = v.z;
        };
        switch (L.Type) {
        case LightType_PointLight:
            cp3(d.pos, L.Position);
            cp3(d.intensity, L.Intensity);
            break;
        case LightType_AreaLight:
            cp3(d.pos, L.Position);
            cp3(d.intensity, L.Intensity);
            cp3(d.e0, L.EdgeVectors[0]);
            cp3(d.e1, L.EdgeVectors[1]);
            cp3(d.normal, L.Normal);
            break;
        case LightType_DirectionalLight:
            cp3(d.dir, L.Direction);
            cp3(d.intensity, L.Intensity);
            break;
        case LightType_SpotLight:
            cp3(d.pos, L.Position);
            cp3(d.intensity, L.Intensity);
            cp3(d.dir, L.SpotDirection);
            d.cosCov = L.CosHalfCoverage;
            d.cosFall = L.CosHalfFallOff;
            break;
        case LightType_EntityLight: {
            const entity &E = L.Entity;
            d.entityId = E.UniqueEntityID;
            d.inv = affineOf(E.InvTransformationMatrix);
            d.fwd = affineOf(E.TransformationMatrix);
            cp3(d.radiance, E.Radiance);
            if (E.Type == EntityType_Sphere) {
                d.entityKind = (int32_t)kItemSphere;
                cp3(d.center, E.Center);
                d.radius = E.Radius;
            } else if (E.Type == EntityType_MeshInstance) {
                d.entityKind = (int32_t)kItemInstance;
                const mesh *M = E.Mesh;
                d.triOffset = (uint32_t)ltris.size();
                d.triCount = M->TriangleCount;
                d.totalArea = M->TotalArea;
                for (u32 k = 0; k < M->TriangleCount; ++k) {
                    DevLightTri t;
                    cp3(t.p0, M->Triangles[k].P0);
                    cp3(t.p1, M->Triangles[k].P1);
                    cp3(t.p2, M->Triangles[k].P2);
                    cp3(t.n, M->Triangles[k].N);
                    ltris.push_back(t);
                }
            } else {
                C->lastError = "entity light of an unsupported entity type";
                return AZ_ERR_UNSUPPORTED;
            }
        } break;
        }
    }
    if (ltris.empty()) ltris.resize(1);
    // Per-triangle shading attributes in primitive-id order.
    std::vector<DevPrimAttr> prims(std::max<u32>(1, info.triTotal));
    for (u32 m = 0; m < S->MeshCount; ++m) {
        const std::vector<const entity *> &tl = info.meshTriEntities[m];
        parallel_for(tl.size(), [&](size_t k) {
            const entity *e = tl[k];
            DevPrimAttr &p = prims[info.meshPrimBase[m] + k];
            auto cp3 = [](float *o, const vec3 &v) {
                o[0] = v.x;
                o[1] = v.y;
                o[2] = v.z;
            };
            cp3(p.n0, e->N0);
            cp3(p.n1, e->N1);
            cp3(p.n2, e->N2);
            cp3(p.fn, e->FaceNormal);
            p.uv0[0] = e->UV0.x;
            p.uv0[1] = e->UV0.y;
            p.uv1[0] = e->UV1.x;
            p.uv1[1] = e->UV1.y;
            p.uv2[0] = e->UV2.x;
            p.uv2[1] = e->UV2.y;
            cp3(p.p0, e->P0);
            cp3(p.p1, e->P1);
            cp3(p.p2, e->P2);
            p.smooth = e->ShadingMode == ShadingMode_Smooth ? 1 : 0;
        });
    }
    // Per top-level item.
    std::vector<DevItemShade> items(info.topEntities.size());
    for (size_t k = 0; k < items.size(); ++k) {
        const entity *e = info.topEntities[k];
        DevItemShade &d = items[k];
        memset(&d, 0, sizeof(d));
        d.material = e->MaterialIndex < S->MaterialCount ? e->MaterialIndex : 0u;
        d.texture = e->TextureIndex <= S->TextureCount ? e->TextureIndex : 0u;
        d.isEmitter = e->IsEmitter ? 1u : 0u;
        d.shadeClass = item_shade_class(d.isEmitter != 0u, mats[d.material]);
        d.kind = e->Type == EntityType_Sphere ? kItemSphere : kItemInstance;
        d.radiance[0] = e->Radiance.x;
        d.radiance[1] = e->Radiance.y;
        d.radiance[2] = e->Radiance.z;
    }
    std::vector<float> grad = perlinTable();
    std::vector<uint64_t> pcg = pcgTable();
    hipStream_t s = C->stream;
    AZ_HIPR(C, uploadVec(B.materials, mats, s));
    AZ_HIPR(C, uploadVec(B.textures, texs, s));
    AZ_HIPR(C, uploadVec(B.texels, texels, s));
    AZ_HIPR(C, uploadVec(B.perms, perms, s));
    AZ_HIPR(C, uploadVec(B.lights, lights, s));
    AZ_HIPR(C, uploadVec(B.lightTris, ltris, s));
    AZ_HIPR(C, uploadVec(B.prims, prims, s));
    AZ_HIPR(C, uploadVec(B.itemShade, items, s));
    AZ_HIPR(C, uploadVec(B.perlin, grad, s));
    AZ_HIPR(C, uploadVec(B.pcg, pcg, s));
    AZ_HIPR(C, hipStreamSynchronize(s));
    D.materials = (const DevMaterial *)B.materials.ptr;
    D.textures = (const DevTexture *)B.textures.ptr;
    D.texels = (const uint8_t *)B.texels.ptr;
    D.perms = (const uint32_t *)B.perms.ptr;
    D.lights = (const DevLight *)B.lights.ptr;
    D.lightTris = (const DevLightTri *)B.lightTris.ptr;
    D.prims = (const DevPrimAttr *)B.prims.ptr;
    D.itemShade = (const DevItemShade *)B.itemShade.ptr;
    D.perlinGrad = (const float *)B.perlin.ptr;
    D.pcgTable = (const uint64_t *)B.pcg.ptr;
    D.lightCount = S->LightCount;
    D.maxDepth = S->MaxRecursionDepth;
    D.pathTrace = S->PathTrace ? 1u : 0u;
    D.importance = S->ImportanceSampling ? 1u : 0u;
    D.useEnv = S->UseEnvironmentMap ? 1u : 0u;
    D.envCdf = nullptr;
    D.envW = D.envH = 0u;
    if (S->UseEnvironmentMap && S->EnvironmentMap.Type == TextureType_2D && S->EnvironmentMap.Pixels &&
        S->EnvironmentMap.Width > 0 && S->EnvironmentMap.Height > 0) {
        std::vector<float> cdf = envLuminanceCdf(S->EnvironmentMap);
        AZ_HIPR(C, uploadVec(B.envCdf, cdf, C->stream));
        D.envCdf = (const float *)B.envCdf.ptr;
        D.envW = S->EnvironmentMap.Width;
        D.envH = S->EnvironmentMap.Height;
    }
    D.useBackgroundTexture = S->UseBackgroundTexture ? 1u : 0u;
    D.ambient[0] = S->AmbientLight.x;
    D.ambient[1] = S->AmbientLight.y;
    D.ambient[2] = S->AmbientLight.z;
    D.backgroundColor[0] = S->BackgroundColor.x;
    D.backgroundColor[1] = S->BackgroundColor.y;
    D.backgroundColor[2] = S->BackgroundColor.z;
    D.epsIntersection = C->eps.Intersection;
    D.epsShadow = C->eps.ShadowRay;
    D.triTotal = info.triTotal;
    for (int k = 0; k < 3; ++k) {
        float ext = info.boxMax[k] - info.boxMin[k];
        D.boxMin[k] = info.boxMin[k];
        D.boxInv[k] = ext > 0.0f ? 1.0f / ext : 0.0f;
    }
    if (S->MaxRecursionDepth > 200) {
        C->lastError = "MaxRecursionDepth above 200 is not supported";
        return AZ_ERR_UNSUPPORTED;
    }
    C->render.ready = true;
    return AZ_OK;
}

// Camera frame exactly as ProduceImageInLinearSpace sets it up
// (main.cpp:1265-1347), in single precision on the host.
static CameraSetup setupCamera(const camera *Cam, u32 sppOverride) {
    struct H3 {
        float x, y, z;
    };
    auto mk = [](float x, float y, float z) { return H3{x, y, z}; };
    auto sub = [&](H3 a, H3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); };
    auto add = [&](H3 a, H3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); };
    auto mul = [&](H3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); };
    auto smul = [&](float s, H3 a) { return mk(s * a.x, s * a.y, s * a.z); };
    auto crs = [&](H3 a, H3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); };
    auto nrm = [&](H3 a) {
        float l = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
        if (l != 0.0f) return mk(a.x / l, a.y / l, a.z / l);
        return mk(0.0f, 0.0f, 0.0f);
    };
    auto from = [&](const vec3 &v) { return mk(v.x, v.y, v.z); };
    CameraSetup c;
    memset(&c, 0, sizeof(c));   // padding too: the struct is part of a graph's key
    memset(&c, 0, sizeof(c));
    H3 gaze = from(Cam->Gaze);
    if (Cam->GazeIsPoint) gaze = sub(from(Cam->Gaze), from(Cam->Position));
    H3 gn = nrm(gaze);
    H3 az = mk(-gn.x, -gn.y, -gn.z);
    H3 ax, ay;
    if (Cam->IsLeftHanded) {
        ax = nrm(crs(nrm(from(Cam->Up)), gn));
        ay = crs(ax, az);
    } else {
        ax = nrm(crs(gn, nrm(from(Cam->Up))));
        ay = crs(az, ax);
    }
    float fdnd = 1.0f;
    if (Cam->FocusDistance != 0.0f) fdnd = Cam->FocusDistance / Cam->NearDistance;
    c.aperture = Cam->ApertureSize;
    c.width = Cam->Resolution.x;
    c.height = Cam->Resolution.y;
    float r = Cam->NearPlane.x, l = Cam->NearPlane.y, b = Cam->NearPlane.z, t = Cam->NearPlane.w;
    if (Cam->Type == CameraType_Simple) {
        float hh = Cam->NearDistance * tanf((Cam->FovY / 2.0f / 180.0f) * kPiF);
        float aspect = (float)c.width / (float)c.height;
        float wh = aspect * hh;
        t = +hh;
        b = -hh;
        r = +wh;
        l = -wh;
    }
    float psx = (r - l) / (float)c.width, psy = (t - b) / (float)c.height;
    H3 sx = mul(mul(ax, psx), fdnd), sy = mul(mul(ay, psy), fdnd);
    H3 pos = from(Cam->Position);
    H3 tl = add(add(add(pos, smul(Cam->NearDistance, gn)), smul(l, ax)), smul(t, ay));
    H3 base = mul(sub(tl, pos), fdnd);
    float *dsts[6] = {c.base, c.pos, c.ax, c.ay, c.stepX, c.stepY};
    H3 srcs[6] = {base, pos, ax, ay, sx, sy};
    for (int k = 0; k < 6; ++k) {
        dsts[k][0] = srcs[k].x;
        dsts[k][1] = srcs[k].y;
        dsts[k][2] = srcs[k].z;
    }
    u32 spp = sppOverride ? sppOverride : Cam->SampleCount;
    c.sppSqrt = (u32)sqrtf((float)spp);
    if (c.sppSqrt == 0) c.sppSqrt = 1;
    c.invSppSqrt = 1.0f / c.sppSqrt;
    float radius = 1.5f;
    float sigma = radius / 3.0f;
    float twoS2 = 2.0f * (sigma * sigma);
    c.negOneOverTwoSigma2 = -1.0f / twoS2;
    c.oneOverTwoPiSigma2 = 1.0f / (twoS2 * (float)kPiD);
    c.kernel = (int32_t)ceilf(radius - 0.5f);
    return c;
}

// Queue geometry of one level of `count` rays: shade block b appends to
// segment b % nseg, and a segment's slice holds what its blocks can emit (per
// ray the scene's most children -- scatter, mirror, refraction, reflection --
// and one shadow ray per light plus the environment).
struct LevelCaps {
    u32 blocks, nseg, childSegCap, shadowSegCap;
    size_t cap, shadowCap;
};

// Bits of v (0 for 0).
static u32 bit_width(u32 v) {
    u32 b = 0;
    while (b < 32u && (v >> b) != 0u) ++b;
    return b;
}

// Bits of the largest random-stream node of a vertex at depth <= maxDepth:
// a camera ray's node is 1 (child_node(0, 0)) and a child's 5n + k + 1
// (k < 4), so at most 2 * 5^d - 1 at depth d; 32 once that passes the
// 32-bit ids (which then wrap).
static u32 node_key_bits(u32 maxDepth) {
    uint64_t m = 1;
    for (u32 d = 0; d < maxDepth && m <= 0xFFFFFFFFull; ++d) m = 5u * m + 4u;
    return m > 0xFFFFFFFFull ? 32u : std::max<u32>(1u, bit_width((u32)m));
}

// Samples per guide training batch: the size of the render's batch buffer
// and the most a weighted draw gathers into it.
constexpr u32 kGuideBatch = 65536;

static LevelCaps level_caps(u32 count, u32 maxChildren, u32 shadowPerRay) {
    LevelCaps L;
    L.blocks = (count + kRenderBlock - 1) / kRenderBlock;
    L.nseg = std::max<u32>(1u, std::min<u32>(kSegs, L.blocks));
    const u32 perSeg = (L.blocks + L.nseg - 1) / L.nseg * kRenderBlock;
    L.childSegCap = maxChildren * perSeg;
    L.shadowSegCap = std::max<u32>(1u, shadowPerRay * perSeg);
    L.cap = (size_t)L.nseg * L.childSegCap;
    L.shadowCap = (size_t)L.nseg * L.shadowSegCap;
    return L;
}

// True (and the error cleared) when e is a failed allocation, which a render
// answers by halving its chunk.
static bool out_of_memory(hipError_t e) {
    if (e != hipErrorOutOfMemory) return false;
    (void)hipGetLastError();
    return true;
}

// Sizes the path, hit, shadow and accumulator buffers for chunks of up to n
// camera samples, with room for levels of up to 1.5 n rays where a shading
// point can emit more than one child, so that a render seldom allocates (and
// frees) anything past its first chunk.  Skipped when the reservation would
// pass `budget` bytes; a failed allocation is left to the render's own
// allocations (which halve the chunk).
static hipError_t reserve_chunk(RenderBuffers &B, u32 n, u32 maxChildren, u32 shadowPerRay, size_t budget) {
    const u32 m = maxChildren > 1 ? n + n / 2 : n;
    const LevelCaps L = level_caps(m, maxChildren, shadowPerRay);
    const size_t need = 2 * L.cap * sizeof(PathRay) + (size_t)m * (sizeof(PathHit) + 3 * sizeof(AccWord)) +
                        L.shadowCap * sizeof(ShadowRay);
    if (need > budget) return hipSuccess;
    hipError_t e = B.q0.alloc(L.cap * sizeof(PathRay));
    if (e == hipSuccess) e = B.q1.alloc(L.cap * sizeof(PathRay));
    if (e == hipSuccess) e = B.acc.alloc((size_t)n * 3 * sizeof(AccWord));
    if (e == hipSuccess) e = B.hits.alloc((size_t)m * sizeof(PathHit));
    if (e == hipSuccess) e = B.shadow.alloc(L.shadowCap * sizeof(ShadowRay));
    return out_of_memory(e) ? hipSuccess : e;
}

static az_status renderImpl(az_gpu_ctx *C, const camera *Cam, filter_type Filter, b32 Cull, const az_render_opts *Opts,
                            void *d_px, bool f64) {
    RenderBuffers &B = buffersOf(C);
    az_render_opts o;
    memset(&o, 0, sizeof(o));
    if (Opts) o = *Opts;
    CameraSetup cam = setupCamera(Cam, o.SampleCountOverride);
    if (cam.width == 0 || cam.height == 0) return AZ_ERR_INVALID_ARGUMENT;
    u32 row0 = o.RowBegin, row1 = o.RowEnd ? std::min(o.RowEnd, cam.height) : cam.height;
    if (row0 >= row1) return AZ_OK;
    const u32 rows = row1 - row0;
    const u32 nPix = rows * cam.width;
    const u32 sppTotal = cam.sppSqrt * cam.sppSqrt;
    DevScene &S = C->devScene;
    // The learned guide: a fresh network per render and use (environment
    // NEE, scatter), trained on the samples of the first trainSpp samples per
    // pixel.
    if (o.EnvSampling > AZ_ENV_LUMINANCE || o.ScatterSampling > AZ_SCATTER_BRDF || o.GuideTarget > AZ_GUIDE_BOTH)
        return AZ_ERR_INVALID_ARGUMENT;
    const bool envGuide = o.GuideEnable != 0 && o.GuideTarget != AZ_GUIDE_SCATTER && S.useEnv;
    const bool scatGuide = o.GuideEnable != 0 && o.GuideTarget != AZ_GUIDE_ENV && S.pathTrace;
    const bool guided = envGuide || scatGuide;
    const bool envLum = o.EnvSampling == AZ_ENV_LUMINANCE && S.useEnv;
    if (envLum && !S.envCdf) {
        C->lastError = "luminance sampling needs an image environment map";
        return AZ_ERR_UNSUPPORTED;
    }
    // By default a quarter of the render's samples per pixel (4 to 32) train
    // the guide: the image then takes most of its samples from the trained
    // network (at 36 spp on the 1080p interior, training on 8-12 spp gives
    // 40 % less error than on 32, in 12 % less time).
    const u32 trainSpp =
        guided ? (o.GuideTrainSpp ? o.GuideTrainSpp : std::min<u32>(32u, std::max<u32>(4u, sppTotal / 4u))) : 0u;
    // Training chunks hold about 2M camera samples (GuideChunkSpp samples
    // per pixel when given), at most trainSpp / 8 samples per pixel and at
    // most the chunk size: the guide is updated after each, so the first
    // chunk is the only one rendered with the untrained network.  By default
    // the Adam steps are spread over the training chunks so that training
    // takes about kGuideSteps steps whatever the image size.
    const u32 chunkSamples = std::max<u32>(1u, C->renderTuning.chunkSamples);
    const u32 chunkSpp0 = std::max<u32>(1u, std::min<u32>(sppTotal, chunkSamples / std::max<u32>(1u, nPix)));
    // Training records are drawn from in the order of their path vertex,
    // key sample << keyNodeBits | node.
    const u32 keyNodeBits = node_key_bits(S.maxDepth);
    const u32 trainWant = o.GuideChunkSpp ? o.GuideChunkSpp : std::max<u32>(1u, (2u << 20) / std::max<u32>(1u, nPix));
    const u32 trainChunkSpp = std::max<u32>(1u, std::min<u32>(std::min(chunkSpp0, trainWant), trainSpp / 8u));
    const u32 trainChunks = trainSpp ? (std::min(trainSpp, sppTotal) + trainChunkSpp - 1) / trainChunkSpp : 1u;
    const u32 kGuideSteps = 512;
    const u32 stepsPerChunk =
        o.GuideStepsPerChunk ? o.GuideStepsPerChunk : std::max<u32>(1u, (kGuideSteps + trainChunks - 1) / trainChunks);
    const float alpha = o.GuideFraction > 0.0f && o.GuideFraction < 1.0f ? o.GuideFraction : 0.5f;
    // Training records, for each guide: room for one per path vertex of a
    // training chunk (a segment that fills drops its later records, which
    // would be the deep levels'), between 8M and 64M records (48 B each, and
    // a 16-byte link for a scatter record).
    const uint64_t recWant = (uint64_t)nPix * trainChunkSpp * (S.maxDepth + 1u);
    const u32 recCap = (u32)std::min<uint64_t>(64u << 20, std::max<uint64_t>(8u << 20, (recWant + kSegs - 1) / kSegs * kSegs));
    if (C->guide) az_guide_destroy(C->guide);
    if (C->scatterGuide) az_guide_destroy(C->scatterGuide);
    C->guide = C->scatterGuide = nullptr;
    if (envGuide) {
        az_status gs = az_guide_create(&C->guide, C->device, o.Seed ^ 0x67756964ull);
        if (gs != AZ_OK) {
            C->lastError = "guide creation failed";
            return gs;
        }
        AZ_HIPR(C, B.recs.alloc((size_t)recCap * sizeof(GuideSample)));
        AZ_HIPR(C, B.recKeys.alloc((size_t)recCap * sizeof(uint64_t)));
    }
    if (scatGuide) {
        az_status gs = az_guide_create(&C->scatterGuide, C->device, o.Seed ^ 0x7363617474657221ull);
        if (gs != AZ_OK) {
            C->lastError = "guide creation failed";
            return gs;
        }
        AZ_HIPR(C, B.srecs.alloc((size_t)recCap * sizeof(GuideSample)));
        AZ_HIPR(C, B.srecKeys.alloc((size_t)recCap * sizeof(uint64_t)));
        AZ_HIPR(C, B.links.alloc((size_t)recCap * sizeof(ScatterLink)));
    }
    if (guided) AZ_HIPR(C, B.batch.alloc((size_t)kGuideBatch * sizeof(GuideSample)));
    // Samples per pixel per chunk: the tuned chunk size (16M camera samples
    // by default); while the guide trains, at least 8 chunks so that it is
    // updated often.
    u32 chunkSpp = chunkSpp0;
    const u32 shadowPerRay = S.lightCount + (S.useEnv ? 1u : 0u);
    hipStream_t st = C->stream;
    // Counters and offsets of the segmented queues: child, shadow and the two
    // guides' training-record counters (kSegs each, one 128-byte line
    // apart), the totals, and the offsets of the current and next path
    // queues, of the shadow queue and of the training records.
    const size_t cntWords = (size_t)kSegs * kSegStride, offWords = kSegs + 1;
    const size_t ctlWords = 4 * cntWords + 16 + 5 * offWords;
    AZ_HIPR(C, B.counters.alloc(ctlWords * sizeof(uint32_t)));
    uint32_t *ctl = (uint32_t *)B.counters.ptr;
    uint32_t *childCnt = ctl, *shadowCnt = ctl + cntWords, *recCnt = ctl + 2 * cntWords, *srecCnt = ctl + 3 * cntWords;
    uint32_t *dTotal = ctl + 4 * cntWords;
    uint32_t *offQ[2] = {dTotal + 16, dTotal + 16 + offWords};
    uint32_t *offShadow = dTotal + 16 + 2 * offWords, *offRec = dTotal + 16 + 3 * offWords;
    uint32_t *offSrec = dTotal + 16 + 4 * offWords;
    hipLaunchKernelGGL(k_clear_u32, dim3((u32)((ctlWords + 255) / 256)), dim3(256), 0, st, ctl, (u32)ctlWords);
    AZ_HIPR(C, hipGetLastError());
    const u32 recSegCap = recCap / kSegs;
    u32 s0 = 0;
    u32 trainRound = 0;
    // Device memory the queues of one chunk may take before the chunk is
    // split into fewer samples per pixel.
    const size_t budget = (size_t)48 << 30;
    AZ_HIPR(C, reserve_chunk(B, nPix * chunkSpp0, S.maxChildren, shadowPerRay, budget));
    // Levels replayed from a graph: each level of a chunk is sized for a
    // bound learned from a host-driven chunk of the same scene, camera, size
    // and options (its rays and shadow rays plus an eighth and 1,024), the
    // kernels read each level's count on the device from the previous
    // level's segment scan, and the levels (extend, shade, segment scan,
    // connect per bounce) are one graph launch with one read-back of the
    // per-level counts at its end.  A level beyond its bound makes the chunk
    // run again with the host-driven levels, which learn the bounds afresh,
    // so the image never depends on the bounds.  The graph is recorded once
    // per bounce count, bounds and set of buffers, and replayed for every
    // chunk and later render with the same ones.
    const bool counting = C->counting;
    const bool levelGraph = C->levelMode == AZ_LEVELS_AUTO && !guided && !counting;
    memset(&C->renderStats, 0, sizeof(C->renderStats));
    unsigned long long *workCnt = nullptr;
    if (counting) {
        AZ_HIPR(C, B.workCounters.alloc(6 * sizeof(unsigned long long)));
        workCnt = (unsigned long long *)B.workCounters.ptr;
        hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, st, (uint32_t *)workCnt, 12u);
        AZ_HIPR(C, hipGetLastError());
        C->renderStats.Gaussian = Filter == FilterType_Gaussian ? 1u : 0u;
        C->renderStats.Pixels = nPix;
    }
    const u32 levels = S.maxDepth + 1u;
    auto boundKeyOf = [&](u32 n) {
        std::vector<uint8_t> k;
        auto put = [&](const void *p, size_t b) {
            const uint8_t *c = (const uint8_t *)p;
            k.insert(k.end(), c, c + b);
        };
        put(&S, sizeof(S));
        put(&C->dg, sizeof(C->dg));
        put(&cam, sizeof(cam));
        const uint32_t v[] = {n, row0, rows, Cull ? 1u : 0u, (uint32_t)Filter, envLum ? 1u : 0u, o.ScatterSampling};
        put(v, sizeof(v));
        return k;
    };
    RenderBuffers::LevelPlan *plan = nullptr;   // the current chunk's, when its levels replay a graph
    auto levelCapsOf = [&](u32 d) { return level_caps(plan->rayBound[d], S.maxChildren, shadowPerRay); };
    // Level d reads its count from levelTotals[2 d] and its scan writes the
    // next level's count and its own shadow-ray count to [2 d + 2], [2 d + 3].
    auto enqueueLevels = [&](u32 n) -> hipError_t {
        const ChunkParams *cpd = (const ChunkParams *)B.chunkDev.ptr;
        uint32_t *LT = (uint32_t *)B.levelTotals.ptr;
        AccWord *acc = (AccWord *)B.acc.ptr;
        ChunkParams unused;
        memset(&unused, 0, sizeof(unused));
        LevelCaps prev{};
        for (u32 depth = 0; depth < levels; ++depth) {
            const LevelCaps L = levelCapsOf(depth);
            const u32 M = plan->rayBound[depth];
            const int a = (int)(depth & 1u);
            PathRay *qa = (PathRay *)(a ? B.q1.ptr : B.q0.ptr), *qb = (PathRay *)(a ? B.q0.ptr : B.q1.ptr);
            const SegIn qin = depth == 0 ? SegIn{nullptr, 1u, n} : SegIn{offQ[a], prev.nseg, prev.childSegCap};
            const uint32_t *cnt = LT + 2 * depth;
            hipError_t e = launch_extend(C->renderTuning, C->cuCount, C->dg, qa, qin, M, (PathHit *)B.hits.ptr,
                                         S.epsIntersection, Cull ? 1 : 0, st, cnt);
            if (e != hipSuccess) return e;
            GuideIO GI{nullptr, nullptr, recCnt, recSegCap, alpha};
            ShadeOut O{qb, (ShadowRay *)B.shadow.ptr, childCnt, shadowCnt, L.nseg, L.childSegCap, L.shadowSegCap};
            hipLaunchKernelGGL(k_shade<kShadeNoGuide>, dim3(L.blocks), dim3(kRenderBlock), 0, st, S, C->dg, cam, unused, qa,
                               qin, (const PathHit *)B.hits.ptr, M, acc, O, GI, cnt, cpd);
            SegScan sc;
            memset(&sc, 0, sizeof(sc));
            sc.counters[0] = childCnt;
            sc.counters[1] = shadowCnt;
            sc.offsets[0] = offQ[1 - a];
            sc.offsets[1] = offShadow;
            sc.segCap[0] = L.childSegCap;
            sc.segCap[1] = L.shadowSegCap;
            sc.nseg[0] = sc.nseg[1] = L.nseg;
            sc.total = LT + 2 * depth + 2;
            hipLaunchKernelGGL(k_seg_scan, dim3(2), dim3(kSegs), 0, st, sc);
            e = launch_connect(S.entityLights == 0u, C->renderTuning, C->cuCount, C->dg, (const ShadowRay *)B.shadow.ptr,
                               SegIn{offShadow, L.nseg, L.shadowSegCap}, plan->shadowBound[depth], acc,
                               (GuideSample *)B.recs.ptr, (ScatterLink *)B.links.ptr, st, LT + 2 * depth + 3);
            if (e != hipSuccess) return e;
            prev = L;
        }
        return hipGetLastError();
    };
    // Everything the recorded launches hold, compared byte for byte.
    auto levelKey = [&](u32 n) {
        std::vector<uint8_t> k;
        auto put = [&](const void *p, size_t b) {
            const uint8_t *c = (const uint8_t *)p;
            k.insert(k.end(), c, c + b);
        };
        put(&S, sizeof(S));
        put(&C->dg, sizeof(C->dg));
        put(&cam, sizeof(cam));
        put(&C->renderTuning, sizeof(C->renderTuning));
        const void *ptrs[] = {B.q0.ptr,   B.q1.ptr,    B.hits.ptr,     B.shadow.ptr,        B.acc.ptr,
                              B.recs.ptr, B.links.ptr, B.chunkDev.ptr, B.levelTotals.ptr, (void *)ctl};
        put(ptrs, sizeof(ptrs));
        const uint32_t v[] = {n, shadowPerRay, Cull ? 1u : 0u, (uint32_t)C->cuCount, recSegCap};
        put(v, sizeof(v));
        put(&alpha, sizeof(alpha));
        put(plan->rayBound.data(), plan->rayBound.size() * sizeof(uint32_t));
        put(plan->shadowBound.data(), plan->shadowBound.size() * sizeof(uint32_t));
        return k;
    };
    while (s0 < sppTotal) {
        const bool training = guided && s0 < trainSpp;
        u32 spp = std::min(chunkSpp, sppTotal - s0);
        if (training) spp = std::min(spp, std::max<u32>(1u, std::min(trainSpp - s0, trainChunkSpp)));
        u32 n = nPix * spp;
        std::vector<uint8_t> boundKey;
        bool graphChunk = false;
        plan = nullptr;
        if (levelGraph) {
            boundKey = boundKeyOf(n);
            for (RenderBuffers::LevelPlan &p : B.plans)
                if (p.boundKey == boundKey && p.rayBound.size() == levels) plan = &p;
            if (plan) {
                plan->used = ++B.planClock;
                size_t cap = n, hitCap = 0, shadowCap = 1;
                bool fit = true;
                for (u32 d = 0; d < levels; ++d) {
                    const LevelCaps L = levelCapsOf(d);
                    cap = std::max(cap, L.cap);
                    hitCap = std::max<size_t>(hitCap, plan->rayBound[d]);
                    shadowCap = std::max(shadowCap, L.shadowCap);
                    fit = fit && L.cap <= 0xFFFFFFFFull && L.shadowCap <= 0xFFFFFFFFull;
                }
                const size_t need = 2 * cap * sizeof(PathRay) + hitCap * sizeof(PathHit) + shadowCap * sizeof(ShadowRay);
                if (fit && need <= budget) {
                    hipError_t ge = B.q0.alloc(cap * sizeof(PathRay));
                    if (ge == hipSuccess) ge = B.q1.alloc(cap * sizeof(PathRay));
                    if (ge == hipSuccess) ge = B.hits.alloc(hitCap * sizeof(PathHit));
                    if (ge == hipSuccess) ge = B.shadow.alloc(shadowCap * sizeof(ShadowRay));
                    if (ge == hipSuccess) ge = B.chunkDev.alloc(sizeof(ChunkParams));
                    if (ge == hipSuccess) ge = B.levelTotals.alloc((size_t)2 * (levels + 1) * sizeof(uint32_t));
                    if (!out_of_memory(ge)) {
                        AZ_HIPR(C, ge);
                        graphChunk = true;
                    } else {
                        // No room for the replayed levels' queues: the chunk runs
                        // host-driven, sized by its own counts, with what was just
                        // grown given back.
                        AZ_HIPR(C, hipStreamSynchronize(st));
                        B.releaseChunkQueues();
                    }
                }
            }
        }
        hipError_t ae = B.q0.alloc((size_t)n * sizeof(PathRay));
        if (ae == hipSuccess) ae = B.acc.alloc((size_t)n * 3 * sizeof(AccWord));
        if (out_of_memory(ae)) {
            if (chunkSpp == 1) {
                C->lastError = "camera rays of one sample per pixel exceed the free device memory (or the render memory limit)";
                return AZ_ERR_OUT_OF_MEMORY;
            }
            chunkSpp = std::max<u32>(1u, chunkSpp / 2);
            AZ_HIPR(C, hipStreamSynchronize(st));
            B.releaseChunkQueues();
            continue;
        }
        AZ_HIPR(C, ae);
        ChunkParams cp;
        memset(&cp, 0, sizeof(cp));
        cp.seed = o.Seed;
        cp.row0 = row0;
        cp.rows = rows;
        cp.s0 = s0;
        cp.spp = spp;
        cp.capacity = n;
        cp.shadowCapacity = 0;
        cp.cull = Cull ? 1u : 0u;
        cp.gaussian = Filter == FilterType_Gaussian ? 1u : 0u;
        cp.envLuminance = envLum ? 1u : 0u;
        cp.brdfSampling = o.ScatterSampling == AZ_SCATTER_BRDF ? 1u : 0u;
        DeviceBuffer *bufA = &B.q0, *bufB = &B.q1;
        DeviceBuffer *recA = &B.qrec0, *recB = &B.qrec1;   // scatter training: per-slot records of the queues
        const bool scatTraining = scatGuide && training;
        int offA = 0;
        AccWord *acc = (AccWord *)B.acc.ptr;
        if (counting) {
            C->renderStats.CameraSamples += n;
            C->renderStats.Chunks += 1;
        }
        hipLaunchKernelGGL(k_generate, dim3((n + kRenderBlock - 1) / kRenderBlock), dim3(kRenderBlock), 0, st, S, cam,
                           cp, (PathRay *)bufA->ptr, acc, n);
        AZ_HIPR(C, hipGetLastError());
        SegIn qin{nullptr, 1u, n};   // camera rays are dense
        u32 count = n;
        bool overflow = false;
        if (graphChunk) {
            uint32_t *LT = (uint32_t *)B.levelTotals.ptr;
            hipLaunchKernelGGL(k_set_chunk, dim3(1), dim3(64), 0, st, cp, (ChunkParams *)B.chunkDev.ptr, LT, n);
            AZ_HIPR(C, hipGetLastError());
            std::vector<uint8_t> key = levelKey(n);
            if (!plan->exec || key != plan->graphKey) {
                plan->dropGraph();
                hipGraph_t g = nullptr;
                AZ_HIPR(C, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                const hipError_t le = enqueueLevels(n);
                const hipError_t ce = hipStreamEndCapture(st, &g);
                if (le != hipSuccess || ce != hipSuccess) {
                    if (g) (void)hipGraphDestroy(g);
                    AZ_HIPR(C, le != hipSuccess ? le : ce);
                }
                const hipError_t ie = hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ie != hipSuccess) plan->exec = nullptr;
                AZ_HIPR(C, ie);
                plan->graphKey = std::move(key);
                ++B.levelCaptures;
            }
            AZ_HIPR(C, hipGraphLaunch(plan->exec, st));
            ++B.levelLaunches;
            std::vector<uint32_t> lt(2 * (levels + 1));
            AZ_HIPR(C, hipMemcpyAsync(lt.data(), LT, lt.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            AZ_HIPR(C, hipStreamSynchronize(st));
            bool fits = true;
            for (u32 d = 0; d < levels; ++d)
                fits = fits && lt[2 * d] <= plan->rayBound[d] && lt[2 * d + 3] <= plan->shadowBound[d];
            if (!fits) {
                plan->boundKey.clear();
                continue;   // the chunk is redone with the host-driven levels; its accumulators are rebuilt
            }
            count = 0;   // the levels are done
        }
        std::vector<u32> seenRays(levels, 0u), seenShadow(levels, 0u);
        for (u32 depth = 0; depth <= S.maxDepth && count > 0; ++depth) {
            const LevelCaps L = level_caps(count, S.maxChildren, shadowPerRay);
            const u32 blocks = L.blocks, nseg = L.nseg, childSegCap = L.childSegCap, shadowSegCap = L.shadowSegCap;
            const size_t cap = L.cap, shadowCap = L.shadowCap;
            size_t need = cap * sizeof(PathRay) + (size_t)count * (sizeof(PathRay) + sizeof(PathHit)) +
                          shadowCap * sizeof(ShadowRay) +
                          (guided ? (size_t)count * (sizeof(GuideQuery) + 80 * sizeof(float)) : 0) +
                          (scatTraining ? cap * sizeof(uint32_t) : 0);
            if (need > budget || cap > 0xFFFFFFFFull || shadowCap > 0xFFFFFFFFull) {
                overflow = true;
                break;
            }
            ae = bufB->alloc(cap * sizeof(PathRay));
            if (ae == hipSuccess) ae = B.hits.alloc((size_t)count * sizeof(PathHit));
            if (ae == hipSuccess) ae = B.shadow.alloc(shadowCap * sizeof(ShadowRay));
            if (ae == hipSuccess && scatTraining) ae = recB->alloc(cap * sizeof(uint32_t));
            if (out_of_memory(ae)) {
                overflow = true;
                break;
            }
            AZ_HIPR(C, ae);
            PathRay *qa = (PathRay *)bufA->ptr, *qb = (PathRay *)bufB->ptr;
            dim3 grid(blocks);
            AZ_HIPR(C, launch_extend(C->renderTuning, C->cuCount, C->dg, qa, qin, count, (PathHit *)B.hits.ptr,
                                     S.epsIntersection, (int)cp.cull, st, nullptr, workCnt));
            if (counting) C->renderStats.PathRays += count;
            GuideIO GI{nullptr, nullptr, recCnt, recSegCap, alpha};
            ShadeOut O{qb, (ShadowRay *)B.shadow.ptr, childCnt, shadowCnt, nseg, childSegCap, shadowSegCap};
            const bool guideLevel = o.GuideMaxDepth == 0 || depth <= o.GuideMaxDepth;
            GI.keyNodeBits = keyNodeBits;
            if (envGuide && training) {
                GI.recs = (GuideSample *)B.recs.ptr;
                GI.recKeys = (uint64_t *)B.recKeys.ptr;
            }
            if (scatTraining) {
                GI.srecs = (GuideSample *)B.srecs.ptr;
                GI.srecKeys = (uint64_t *)B.srecKeys.ptr;
                GI.links = (ScatterLink *)B.links.ptr;
                GI.srecCounters = srecCnt;
                GI.srecSegCap = recSegCap;
                GI.rayRec = depth > 0 ? (const uint32_t *)recA->ptr : nullptr;
                O.nextRec = (uint32_t *)recB->ptr;
            }
            if (guided && guideLevel && envGuide && !scatGuide) {
                // The environment guide alone: queries and inference fused.
                AZ_HIPR(C, B.lobes.alloc((size_t)count * kGuideLobeBytes));
                const u32 chunks = (count + kInferSamples - 1) / kInferSamples;
                dim3 ig(std::min<u32>((chunks + kInferWaves - 1) / kInferWaves, (u32)C->cuCount));
                hipLaunchKernelGGL(k_guide_query_infer, ig, dim3(64 * kInferWaves), 0, st, S, C->dg,
                                   az_guide_device_weights(C->guide), qa, qin, (const PathHit *)B.hits.ptr, count,
                                   (uint2 *)B.lobes.ptr);
                AZ_HIPR(C, hipGetLastError());
                GI.lobes = (const uint2 *)B.lobes.ptr;
            } else if (guided && guideLevel) {
                // Query every hit, then one batched MFMA inference per guide.
                AZ_HIPR(C, B.queries.alloc((size_t)count * sizeof(GuideQuery)));
                hipLaunchKernelGGL(k_guide_query, grid, dim3(kRenderBlock), 0, st, S, C->dg, qa, qin,
                                   (const PathHit *)B.hits.ptr, count, (GuideQuery *)B.queries.ptr);
                AZ_HIPR(C, hipGetLastError());
                if (envGuide) {
                    AZ_HIPR(C, B.lobes.alloc((size_t)count * kGuideLobeBytes));
                    az_status gs = az_guide_infer_on_stream(C->guide, B.queries.ptr, count, B.lobes.ptr, st);
                    if (gs != AZ_OK) {
                        C->lastError = "guide inference failed";
                        return gs;
                    }
                    GI.lobes = (const uint2 *)B.lobes.ptr;
                }
                if (scatGuide) {
                    AZ_HIPR(C, B.scatLobes.alloc((size_t)count * kGuideLobeBytes));
                    az_status gs = az_guide_infer_on_stream(C->scatterGuide, B.queries.ptr, count, B.scatLobes.ptr, st);
                    if (gs != AZ_OK) {
                        C->lastError = "guide inference failed";
                        return gs;
                    }
                    GI.scatLobes = (const uint2 *)B.scatLobes.ptr;
                }
            }
            // Each level takes the lightest shading kernel that holds the guide
            // code it uses (none past GuideMaxDepth and outside training); with
            // GI's pointers null the kernels compute the same.
            const bool scatCode = GI.scatLobes || GI.srecs || GI.rayRec || O.nextRec;
            const bool envCode = GI.lobes || GI.recs;
            if (scatCode)
                hipLaunchKernelGGL(k_shade<kShadeAllGuides>, grid, dim3(kRenderBlock), 0, st, S, C->dg, cam, cp, qa, qin,
                                   (const PathHit *)B.hits.ptr, count, acc, O, GI, nullptr, nullptr);
            else if (envCode)
                hipLaunchKernelGGL(k_shade<kShadeEnvGuide>, grid, dim3(kRenderBlock), 0, st, S, C->dg, cam, cp, qa, qin,
                                   (const PathHit *)B.hits.ptr, count, acc, O, GI, nullptr, nullptr);
            else
                hipLaunchKernelGGL(k_shade<kShadeNoGuide>, grid, dim3(kRenderBlock), 0, st, S, C->dg, cam, cp, qa, qin,
                                   (const PathHit *)B.hits.ptr, count, acc, O, GI, nullptr, nullptr);
            AZ_HIPR(C, hipGetLastError());
            SegScan sc;
            sc.counters[0] = childCnt;
            sc.counters[1] = shadowCnt;
            sc.offsets[0] = offQ[1 - offA];
            sc.offsets[1] = offShadow;
            sc.segCap[0] = childSegCap;
            sc.segCap[1] = shadowSegCap;
            sc.nseg[0] = sc.nseg[1] = nseg;
            sc.total = dTotal;
            hipLaunchKernelGGL(k_seg_scan, dim3(2), dim3(kSegs), 0, st, sc);
            AZ_HIPR(C, hipGetLastError());
            uint32_t hc[2];
            AZ_HIPR(C, hipMemcpyAsync(hc, dTotal, 8, hipMemcpyDeviceToHost, st));
            AZ_HIPR(C, hipStreamSynchronize(st));
            seenRays[depth] = count;
            seenShadow[depth] = hc[1];
            if (hc[1]) {
                SegIn sin{offShadow, nseg, shadowSegCap};
                AZ_HIPR(C, launch_connect(S.entityLights == 0u, C->renderTuning, C->cuCount, C->dg,
                                          (const ShadowRay *)B.shadow.ptr, sin,
                                          hc[1], acc, (GuideSample *)B.recs.ptr, (ScatterLink *)B.links.ptr, st,
                                          nullptr, counting ? workCnt + 3 : nullptr));
                if (counting) C->renderStats.ShadowRays += hc[1];
            }
            count = hc[0];
            offA = 1 - offA;
            qin = SegIn{offQ[offA], nseg, childSegCap};
            std::swap(bufA, bufB);
            std::swap(recA, recB);
        }
        if (overflow) {
            if (chunkSpp == 1) {
                C->lastError = "path queue exceeds the memory budget (or free device memory) with one sample per chunk";
                return AZ_ERR_OUT_OF_MEMORY;
            }
            chunkSpp = std::max<u32>(1u, chunkSpp / 2);
            AZ_HIPR(C, hipStreamSynchronize(st));
            B.releaseChunkQueues();
            continue;   // the chunk is redone with fewer samples; its accumulators are rebuilt
        }
        if (levelGraph && !graphChunk) {
            // Bounds for the graph-replayed levels of the next chunks of this
            // size, in the plan learned on the same key or the least recently
            // used one.
            RenderBuffers::LevelPlan *p = &B.plans[0];
            for (RenderBuffers::LevelPlan &q : B.plans)
                if (q.boundKey == boundKey) p = &q;
            if (p->boundKey != boundKey && B.plans[1].used < B.plans[0].used) p = &B.plans[1];
            p->dropGraph();
            p->rayBound.assign(levels, 0u);
            p->shadowBound.assign(levels, 0u);
            for (u32 d = 0; d < levels; ++d) {
                p->rayBound[d] = d == 0 ? n : (u32)std::min<uint64_t>(0xFFFFFFFFull, seenRays[d] + seenRays[d] / 8ull + 1024u);
                p->shadowBound[d] = (u32)std::min<uint64_t>(0xFFFFFFFFull, seenShadow[d] + seenShadow[d] / 8ull + 1024u);
            }
            p->boundKey = boundKey;
            p->used = ++B.planClock;
        }
        if (!cp.gaussian) {
            dim3 g((nPix + kRenderBlock - 1) / kRenderBlock);
            if (f64) hipLaunchKernelGGL(k_splat_box<double>, g, dim3(kRenderBlock), 0, st, cam, cp, acc, (double *)d_px);
            else hipLaunchKernelGGL(k_splat_box<float>, g, dim3(kRenderBlock), 0, st, cam, cp, acc, (float *)d_px);
        } else {
            u32 d0 = row0 > (u32)cam.kernel ? row0 - (u32)cam.kernel : 0u;
            u32 d1 = std::min(cam.height, row1 + (u32)cam.kernel);
            dim3 g(splat_gauss_blocks(cam.width, d1 - d0));
            if (f64)
                hipLaunchKernelGGL(k_splat_gauss<double>, g, dim3(kRenderBlock), 0, st, cam, cp, acc, (double *)d_px, d0,
                                   d1 - d0);
            else
                hipLaunchKernelGGL(k_splat_gauss<float>, g, dim3(kRenderBlock), 0, st, cam, cp, acc, (float *)d_px, d0,
                                   d1 - d0);
        }
        AZ_HIPR(C, hipGetLastError());
        if (training) {
            // Adam steps on batches gathered from this chunk's samples: the
            // environment records (queue 0) and the scatter records (queue 1),
            // whose weights wait for their subtrees' sums.
            SegScan sc;
            memset(&sc, 0, sizeof(sc));
            sc.counters[0] = recCnt;
            sc.offsets[0] = offRec;
            sc.counters[1] = srecCnt;
            sc.offsets[1] = offSrec;
            sc.segCap[0] = sc.segCap[1] = recSegCap;
            sc.nseg[0] = sc.nseg[1] = kSegs;
            sc.total = dTotal + 2;
            hipLaunchKernelGGL(k_seg_scan, dim3(2), dim3(kSegs), 0, st, sc);
            AZ_HIPR(C, hipGetLastError());
            uint32_t R[2] = {0u, 0u};
            AZ_HIPR(C, hipMemcpyAsync(R, dTotal + 2, 8, hipMemcpyDeviceToHost, st));
            AZ_HIPR(C, hipStreamSynchronize(st));
            for (int g = 0; g < 2; ++g) {
                az_guide *net = g == 0 ? C->guide : C->scatterGuide;
                const GuideSample *recs = (const GuideSample *)(g == 0 ? B.recs.ptr : B.srecs.ptr);
                if (!net || R[g] < 1024) continue;
                SegIn rin{g == 0 ? offRec : offSrec, kSegs, recSegCap};
                const dim3 rg((R[g] + kRenderBlock - 1) / kRenderBlock);
                if (g == 1) {
                    for (u32 d = S.maxDepth; d-- > 1u;)
                        hipLaunchKernelGGL(k_scatter_propagate, rg, dim3(kRenderBlock), 0, st, (ScatterLink *)B.links.ptr,
                                           rin, R[g], d);
                    hipLaunchKernelGGL(k_scatter_finalize, rg, dim3(kRenderBlock), 0, st, (GuideSample *)B.srecs.ptr,
                                       (const ScatterLink *)B.links.ptr, rin, R[g]);
                    AZ_HIPR(C, hipGetLastError());
                }
                u32 bn = std::min<u32>(kGuideBatch, R[g]) & ~15u;   // within the batch buffer
                const u32 pb = (R[g] + kRenderBlock - 1) / kRenderBlock;
                AZ_HIPR(C, B.recCum.alloc(((size_t)R[g] + pb) * sizeof(float)));
                float *cum = (float *)B.recCum.ptr, *sums = cum + R[g];
                // Records in key order (see k_rec_keys).
                AZ_HIPR(C, B.recSort.alloc((size_t)R[g] * 2 * (sizeof(uint64_t) + sizeof(uint32_t))));
                uint64_t *kin = (uint64_t *)B.recSort.ptr, *kout = kin + R[g];
                uint32_t *vin = (uint32_t *)(kout + R[g]), *order = vin + R[g];
                hipLaunchKernelGGL(k_rec_keys, rg, dim3(kRenderBlock), 0, st,
                                   (const uint64_t *)(g == 0 ? B.recKeys.ptr : B.srecKeys.ptr), rin, R[g], kin, vin);
                AZ_HIPR(C, hipGetLastError());
                size_t tmpBytes = 0;
                // Only the key's bits in use are sorted on.
                const u32 keyBits = std::min<u32>(64u, keyNodeBits + bit_width(n - 1u));
                AZ_HIPR(C, rocprim::radix_sort_pairs(nullptr, tmpBytes, kin, kout, vin, order, R[g], 0, keyBits, st));
                AZ_HIPR(C, B.recSortTmp.alloc(std::max<size_t>(tmpBytes, 16)));
                AZ_HIPR(C, rocprim::radix_sort_pairs(B.recSortTmp.ptr, tmpBytes, kin, kout, vin, order, R[g], 0, keyBits,
                                                     st));
                hipLaunchKernelGGL(k_rec_prefix_blocks, dim3(pb), dim3(kRenderBlock), 0, st, recs,
                                   (const uint32_t *)order, R[g], cum, sums);
                hipLaunchKernelGGL(k_rec_prefix_sums, dim3(1), dim3(kRenderBlock), 0, st, sums, pb);
                hipLaunchKernelGGL(k_rec_prefix_add, dim3(pb), dim3(kRenderBlock), 0, st, cum, R[g], (const float *)sums);
                AZ_HIPR(C, hipGetLastError());
                for (u32 k = 0; k < stepsPerChunk; ++k) {
                    uint64_t seed = (o.Seed + 0x9E3779B97F4A7C15ull * (uint64_t)(g + 1)) * (uint64_t)(trainRound * 131 + k + 1);
                    hipLaunchKernelGGL(k_guide_gather_weighted, dim3((bn + kRenderBlock - 1) / kRenderBlock),
                                       dim3(kRenderBlock), 0, st, recs, (const uint32_t *)order, R[g],
                                       (const float *)cum, (GuideSample *)B.batch.ptr, bn, seed);
                    AZ_HIPR(C, hipGetLastError());
                    az_status gs = az_guide_train_on_stream(net, B.batch.ptr, bn, st);
                    if (gs != AZ_OK) {
                        C->lastError = "guide training failed";
                        return gs;
                    }
                }
            }
            ++trainRound;
        }
        s0 += spp;
    }
    if (counting) {
        unsigned long long w[6];
        AZ_HIPR(C, hipMemcpyAsync(w, workCnt, sizeof(w), hipMemcpyDeviceToHost, st));
        AZ_HIPR(C, hipStreamSynchronize(st));
        C->renderStats.ExtendNodes = w[0];
        C->renderStats.ExtendTriangles = w[1];
        C->renderStats.ExtendItems = w[2];
        C->renderStats.ConnectNodes = w[3];
        C->renderStats.ConnectTriangles = w[4];
        C->renderStats.ConnectItems = w[5];
    }
    return AZ_OK;
}

// ---- batched renders ---------------------------------------------------------------------
// Several scenes, one per context, rendered together: each chunk's stages
// (generate, and per level extend, shade, segment scan, connect, then
// splat) are single launches over the queues of every scene still active,
// with the blocks split into per-scene ranges.  Each scene keeps its own
// queues and counters (its context's buffers); the arguments of a launch
// go to device memory in one copy per stage group.
namespace {

struct BatchScene {
    az_gpu_ctx *C = nullptr;
    RenderBuffers *B = nullptr;
    CameraSetup cam{};
    ChunkParams cp{};
    void *px = nullptr;
    u32 row0 = 0, rows = 0, nPix = 0, sppTotal = 0, chunkSpp = 1, s0 = 0, spp = 0, shadowPerRay = 0;
    bool envLum = false;
    DeviceBuffer *bufA = nullptr, *bufB = nullptr;
    int offA = 0;
    u32 count = 0, depth = 0;
    SegIn qin{};
    uint32_t *childCnt = nullptr, *shadowCnt = nullptr, *offShadow = nullptr;
    uint32_t *offQ[2] = {nullptr, nullptr};
    u32 nseg = 0, childSegCap = 0, shadowSegCap = 0;
};

// Host staging for one group of arguments; the device copy goes to its own
// region of the batch-argument buffer.  The callers synchronize the stream
// before a staging buffer is refilled.
struct ArgPack {
    std::vector<uint8_t> host;
    template <typename T>
    size_t add(const T *items, size_t n) {
        size_t off = (host.size() + 255) & ~(size_t)255;
        host.resize(off + n * sizeof(T));
        memcpy(host.data() + off, items, n * sizeof(T));
        return off;
    }
};

constexpr size_t kBatchRegion = (size_t)1 << 20;   // bytes per argument region (3 regions)

}  // namespace

static az_status renderBatchImpl(az_gpu_ctx *const *ctxs, const camera *const *cams, u32 count, filter_type Filter,
                                 b32 Cull, const az_render_opts *Opts, void *const *pxs) {
    az_gpu_ctx *C0 = ctxs[0];
    RenderBuffers &B0 = buffersOf(C0);
    az_render_opts o;
    memset(&o, 0, sizeof(o));
    if (Opts) o = *Opts;
    if (o.GuideEnable) {
        C0->lastError = "the learned guide is not available in batched renders";
        return AZ_ERR_UNSUPPORTED;
    }
    if (o.EnvSampling > AZ_ENV_LUMINANCE || o.ScatterSampling > AZ_SCATTER_BRDF) return AZ_ERR_INVALID_ARGUMENT;
    const bool f64 = o.AccumulateF64 != 0;
    hipStream_t st = C0->stream;
    const RenderTuning tune = C0->renderTuning;
    std::vector<BatchScene> sc(count);
    const size_t cntWords = (size_t)kSegs * kSegStride, offWords = kSegs + 1;
    const size_t ctlWords = 3 * cntWords + 16 + 4 * offWords;
    for (u32 k = 0; k < count; ++k) {
        BatchScene &b = sc[k];
        b.C = ctxs[k];
        if (!b.C || !cams[k] || !pxs[k]) return AZ_ERR_INVALID_ARGUMENT;
        if (!b.C->hasScene || !b.C->render.ready) return AZ_ERR_NO_SCENE;
        if (b.C->device != C0->device) {
            C0->lastError = "batched contexts must share one device";
            return AZ_ERR_INVALID_ARGUMENT;
        }
        if (b.C->stream != st) AZ_HIPR(C0, hipStreamSynchronize(b.C->stream));
        b.B = &buffersOf(b.C);
        b.px = pxs[k];
        b.cam = setupCamera(cams[k], o.SampleCountOverride);
        if (b.cam.width == 0 || b.cam.height == 0) return AZ_ERR_INVALID_ARGUMENT;
        b.row0 = o.RowBegin;
        const u32 row1 = o.RowEnd ? std::min(o.RowEnd, b.cam.height) : b.cam.height;
        b.rows = row1 > b.row0 ? row1 - b.row0 : 0u;
        b.nPix = b.rows * b.cam.width;
        b.sppTotal = b.rows ? b.cam.sppSqrt * b.cam.sppSqrt : 0u;
        const DevScene &S = b.C->devScene;
        b.envLum = o.EnvSampling == AZ_ENV_LUMINANCE && S.useEnv;
        if (b.envLum && !S.envCdf) {
            C0->lastError = "luminance sampling needs an image environment map";
            return AZ_ERR_UNSUPPORTED;
        }
        b.shadowPerRay = S.lightCount + (S.useEnv ? 1u : 0u);
        // The scenes share one chunk's worth of samples (at least a sample
        // per pixel each), so a batch takes about the memory of one render.
        const u32 share = tune.chunkSamples / count;
        b.chunkSpp = std::max<u32>(1u, std::min<u32>(b.sppTotal, share / std::max<u32>(1u, b.nPix)));
        AZ_HIPR(C0, reserve_chunk(*b.B, b.nPix * b.chunkSpp, S.maxChildren, b.shadowPerRay, (size_t)48 << 30));
        AZ_HIPR(C0, b.B->counters.alloc(ctlWords * sizeof(uint32_t)));
        uint32_t *ctl = (uint32_t *)b.B->counters.ptr;
        b.childCnt = ctl;
        b.shadowCnt = ctl + cntWords;
        uint32_t *dTotal = ctl + 3 * cntWords;
        b.offQ[0] = dTotal + 16;
        b.offQ[1] = dTotal + 16 + offWords;
        b.offShadow = dTotal + 16 + 2 * offWords;
        hipLaunchKernelGGL(k_clear_u32, dim3((u32)((ctlWords + 255) / 256)), dim3(256), 0, st, ctl, (u32)ctlWords);
        AZ_HIPR(C0, hipGetLastError());
    }
    AZ_HIPR(C0, B0.batchTotals.alloc((size_t)2 * count * sizeof(uint32_t)));
    AZ_HIPR(C0, B0.batchArgs.alloc(3 * kBatchRegion));
    uint8_t *region[3] = {(uint8_t *)B0.batchArgs.ptr, (uint8_t *)B0.batchArgs.ptr + kBatchRegion,
                          (uint8_t *)B0.batchArgs.ptr + 2 * kBatchRegion};
    uint32_t *totals = (uint32_t *)B0.batchTotals.ptr;
    const size_t budget = (size_t)48 << 30;
    ArgPack packs[3];
    auto upload = [&](int r) -> hipError_t {
        if (packs[r].host.size() > kBatchRegion) return hipErrorInvalidValue;
        return hipMemcpyAsync(region[r], packs[r].host.data(), packs[r].host.size(), hipMemcpyHostToDevice, st);
    };
    for (;;) {
        std::vector<u32> act;
        for (u32 k = 0; k < count; ++k)
            if (sc[k].s0 < sc[k].sppTotal) act.push_back(k);
        if (act.empty()) break;
        // Chunk set-up and camera rays.
        std::vector<GenArgs> gen;
        BatchMap gm{};
        gm.count = (u32)act.size();
        u32 gblocks = 0;
        for (u32 j = 0; j < act.size(); ++j) {
            BatchScene &b = sc[act[j]];
            b.spp = std::min(b.chunkSpp, b.sppTotal - b.s0);
            const u32 n = b.nPix * b.spp;
            AZ_HIPR(C0, b.B->q0.alloc((size_t)n * sizeof(PathRay)));
            AZ_HIPR(C0, b.B->acc.alloc((size_t)n * 3 * sizeof(AccWord)));
            memset(&b.cp, 0, sizeof(b.cp));
            b.cp.seed = o.Seed;
            b.cp.row0 = b.row0;
            b.cp.rows = b.rows;
            b.cp.s0 = b.s0;
            b.cp.spp = b.spp;
            b.cp.capacity = n;
            b.cp.cull = Cull ? 1u : 0u;
            b.cp.gaussian = Filter == FilterType_Gaussian ? 1u : 0u;
            b.cp.envLuminance = b.envLum ? 1u : 0u;
            b.cp.brdfSampling = o.ScatterSampling == AZ_SCATTER_BRDF ? 1u : 0u;
            b.bufA = &b.B->q0;
            b.bufB = &b.B->q1;
            b.offA = 0;
            b.count = n;
            b.depth = 0;
            b.qin = SegIn{nullptr, 1u, n};
            GenArgs g;
            g.S = b.C->devScene;
            g.cam = b.cam;
            g.cp = b.cp;
            g.q = (PathRay *)b.bufA->ptr;
            g.acc = (AccWord *)b.B->acc.ptr;
            g.n = n;
            gen.push_back(g);
            gm.start[j] = gblocks;
            gblocks += (n + kRenderBlock - 1) / kRenderBlock;
        }
        gm.start[act.size()] = gblocks;
        packs[2].host.clear();
        const size_t genOff = packs[2].add(gen.data(), gen.size());
        AZ_HIPR(C0, upload(2));
        hipLaunchKernelGGL(kb_generate, dim3(gblocks), dim3(kRenderBlock), 0, st,
                           (const GenArgs *)(region[2] + genOff), gm);
        AZ_HIPR(C0, hipGetLastError());
        // Levels.
        for (;;) {
            std::vector<u32> lv;
            for (u32 k : act)
                if (sc[k].count > 0 && sc[k].depth <= sc[k].C->devScene.maxDepth) lv.push_back(k);
            if (lv.empty()) break;
            std::vector<ExtArgs> ext;
            std::vector<ShadeArgs> sh;
            std::vector<SegScan> scans;
            BatchMap em{}, smap{};
            em.count = smap.count = (u32)lv.size();
            u32 eblocks = 0, sblocks = 0;
            const u32 perExt = (u32)tune.traceBlock * kRaysPerLane;
            for (u32 j = 0; j < lv.size(); ++j) {
                BatchScene &b = sc[lv[j]];
                const LevelCaps L = level_caps(b.count, b.C->devScene.maxChildren, b.shadowPerRay);
                const u32 blocks = L.blocks;
                b.nseg = L.nseg;
                b.childSegCap = L.childSegCap;
                b.shadowSegCap = L.shadowSegCap;
                const size_t cap = L.cap, shadowCap = L.shadowCap;
                const size_t need = cap * sizeof(PathRay) + (size_t)b.count * (sizeof(PathRay) + sizeof(PathHit)) +
                                    shadowCap * sizeof(ShadowRay);
                if (need > budget || cap > 0xFFFFFFFFull || shadowCap > 0xFFFFFFFFull) {
                    C0->lastError = "batched render: path queue exceeds the memory budget";
                    return AZ_ERR_OUT_OF_MEMORY;
                }
                AZ_HIPR(C0, b.bufB->alloc(cap * sizeof(PathRay)));
                AZ_HIPR(C0, b.B->hits.alloc((size_t)b.count * sizeof(PathHit)));
                AZ_HIPR(C0, b.B->shadow.alloc(shadowCap * sizeof(ShadowRay)));
                ExtArgs e;
                e.G = b.C->dg;
                e.q = (const PathRay *)b.bufA->ptr;
                e.qin = b.qin;
                e.n = b.count;
                e.hits = (PathHit *)b.B->hits.ptr;
                e.eps = b.C->devScene.epsIntersection;
                e.cull = (int)b.cp.cull;
                ext.push_back(e);
                em.start[j] = eblocks;
                eblocks += (b.count + perExt - 1) / perExt;
                ShadeArgs a;
                a.S = b.C->devScene;
                a.G = b.C->dg;
                a.cam = b.cam;
                a.cp = b.cp;
                a.q = (const PathRay *)b.bufA->ptr;
                a.qin = b.qin;
                a.hits = (const PathHit *)b.B->hits.ptr;
                a.n = b.count;
                a.acc = (AccWord *)b.B->acc.ptr;
                a.O = ShadeOut{(PathRay *)b.bufB->ptr, (ShadowRay *)b.B->shadow.ptr, b.childCnt, b.shadowCnt,
                               b.nseg, b.childSegCap, b.shadowSegCap};
                a.GI = GuideIO{nullptr, nullptr, b.childCnt, 0u, 0.5f};
                sh.push_back(a);
                smap.start[j] = sblocks;
                sblocks += blocks;
                SegScan q;
                q.counters[0] = b.childCnt;
                q.counters[1] = b.shadowCnt;
                q.offsets[0] = b.offQ[1 - b.offA];
                q.offsets[1] = b.offShadow;
                q.segCap[0] = b.childSegCap;
                q.segCap[1] = b.shadowSegCap;
                q.nseg[0] = q.nseg[1] = b.nseg;
                q.total = totals + 2 * j;
                scans.push_back(q);
            }
            em.start[lv.size()] = eblocks;
            smap.start[lv.size()] = sblocks;
            packs[0].host.clear();
            const size_t extOff = packs[0].add(ext.data(), ext.size());
            const size_t shOff = packs[0].add(sh.data(), sh.size());
            const size_t scOff = packs[0].add(scans.data(), scans.size());
            AZ_HIPR(C0, upload(0));
            AZ_HIPR(C0, launch_extend_batch(tune, (const ExtArgs *)(region[0] + extOff), em, eblocks, st));
            hipLaunchKernelGGL(kb_shade, dim3(sblocks), dim3(kRenderBlock), 0, st,
                               (const ShadeArgs *)(region[0] + shOff), smap);
            AZ_HIPR(C0, hipGetLastError());
            hipLaunchKernelGGL(kb_seg_scan, dim3(2 * (u32)lv.size()), dim3(kSegs), 0, st,
                               (const SegScan *)(region[0] + scOff));
            AZ_HIPR(C0, hipGetLastError());
            std::vector<uint32_t> hc(2 * lv.size());
            AZ_HIPR(C0, hipMemcpyAsync(hc.data(), totals, hc.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            AZ_HIPR(C0, hipStreamSynchronize(st));
            std::vector<ConnArgs> con;
            BatchMap cm{};
            u32 cblocks = 0;
            for (u32 j = 0; j < lv.size(); ++j) {
                BatchScene &b = sc[lv[j]];
                const u32 ns = hc[2 * j + 1];
                if (ns) {
                    ConnArgs c;
                    c.G = b.C->dg;
                    c.sq = (const ShadowRay *)b.B->shadow.ptr;
                    c.sin = SegIn{b.offShadow, b.nseg, b.shadowSegCap};
                    c.n = ns;
                    c.acc = (AccWord *)b.B->acc.ptr;
                    c.recs = nullptr;
                    cm.start[con.size()] = cblocks;
                    cblocks += (ns + perExt - 1) / perExt;
                    con.push_back(c);
                }
                b.count = hc[2 * j];
                b.offA = 1 - b.offA;
                b.qin = SegIn{b.offQ[b.offA], b.nseg, b.childSegCap};
                std::swap(b.bufA, b.bufB);
                b.depth++;
            }
            if (!con.empty()) {
                cm.count = (u32)con.size();
                cm.start[con.size()] = cblocks;
                packs[1].host.clear();
                const size_t cOff = packs[1].add(con.data(), con.size());
                AZ_HIPR(C0, upload(1));
                AZ_HIPR(C0, launch_connect_batch(tune, (const ConnArgs *)(region[1] + cOff), cm, cblocks, st));
            }
        }
        // Splat every active scene's chunk into its pixels.
        std::vector<SplatArgs> spl;
        BatchMap pm{};
        pm.count = (u32)act.size();
        u32 pblocks = 0;
        for (u32 j = 0; j < act.size(); ++j) {
            BatchScene &b = sc[act[j]];
            SplatArgs a;
            a.cam = b.cam;
            a.cp = b.cp;
            a.acc = (const AccWord *)b.B->acc.ptr;
            a.px = b.px;
            u32 blocks = (b.nPix + kRenderBlock - 1) / kRenderBlock;
            a.dstRow0 = b.row0;
            a.dstRows = b.rows;
            if (b.cp.gaussian) {
                const u32 row1 = b.row0 + b.rows;
                a.dstRow0 = b.row0 > (u32)b.cam.kernel ? b.row0 - (u32)b.cam.kernel : 0u;
                a.dstRows = std::min(b.cam.height, row1 + (u32)b.cam.kernel) - a.dstRow0;
                blocks = splat_gauss_blocks(b.cam.width, a.dstRows);
            }
            spl.push_back(a);
            pm.start[j] = pblocks;
            pblocks += blocks;
            b.s0 += b.spp;
        }
        pm.start[act.size()] = pblocks;
        AZ_HIPR(C0, hipStreamSynchronize(st));   // the generate arguments' staging is reused
        packs[2].host.clear();
        const size_t pOff = packs[2].add(spl.data(), spl.size());
        AZ_HIPR(C0, upload(2));
        if (f64)
            hipLaunchKernelGGL(kb_splat<double>, dim3(pblocks), dim3(kRenderBlock), 0, st,
                               (const SplatArgs *)(region[2] + pOff), pm);
        else
            hipLaunchKernelGGL(kb_splat<float>, dim3(pblocks), dim3(kRenderBlock), 0, st,
                               (const SplatArgs *)(region[2] + pOff), pm);
        AZ_HIPR(C0, hipGetLastError());
        AZ_HIPR(C0, hipStreamSynchronize(st));   // the splat arguments' staging is reused next chunk
    }
    return AZ_OK;
}

extern "C" {

az_status az_gpu_render_batch(az_gpu_ctx *const *Ctxs, const camera *const *Cameras, u32 Count, filter_type Filter,
                              b32 Cull, const az_render_opts *Opts, void *const *d_Pixels) {
    if (!Ctxs || !Cameras || !d_Pixels || Count == 0 || Count > kMaxBatch || !Ctxs[0]) return AZ_ERR_INVALID_ARGUMENT;
    az_gpu_ctx *C0 = Ctxs[0];
    AZ_HIPR(C0, hipSetDevice(C0->device));
    AZ_HIPR(C0, hipEventRecord(C0->ev0, C0->stream));
    az_status st = renderBatchImpl(Ctxs, Cameras, Count, Filter, Cull, Opts, d_Pixels);
    AZ_HIPR(C0, hipEventRecord(C0->ev1, C0->stream));
    C0->lastStart = C0->ev0;
    C0->lastStop = C0->ev1;
    C0->pendingTiming = true;
    return st;
}

az_status az_gpu_render_device(az_gpu_ctx *C, const camera *Cam, filter_type Filter, b32 Cull,
                               const az_render_opts *Opts, void *d_px) {
    if (!C || !Cam || !d_px) return AZ_ERR_INVALID_ARGUMENT;
    if (!C->hasScene || !C->render.ready) return AZ_ERR_NO_SCENE;
    AZ_HIPR(C, hipSetDevice(C->device));
    AZ_HIPR(C, hipEventRecord(C->ev0, C->stream));
    az_status st = renderImpl(C, Cam, Filter, Cull, Opts, d_px, Opts && Opts->AccumulateF64);
    AZ_HIPR(C, hipEventRecord(C->ev1, C->stream));
    C->lastStart = C->ev0;
    C->lastStop = C->ev1;
    C->pendingTiming = true;
    return st;
}

az_status az_gpu_render(az_gpu_ctx *C, const camera *Cam, filter_type Filter, b32 Cull, const az_render_opts *Opts,
                        pixel *Pixels) {
    if (!C || !Cam || !Pixels) return AZ_ERR_INVALID_ARGUMENT;
    if (!C->hasScene || !C->render.ready) return AZ_ERR_NO_SCENE;
    if (Opts && Opts->AccumulateF64) return AZ_ERR_INVALID_ARGUMENT;   // f64 output is device-only
    RenderBuffers &B = buffersOf(C);
    size_t bytes = (size_t)Cam->Resolution.x * Cam->Resolution.y * sizeof(pixel);
    AZ_HIPR(C, hipSetDevice(C->device));
    AZ_HIPR(C, B.pixels.alloc(bytes));
    AZ_HIPR(C, hipMemcpyAsync(B.pixels.ptr, Pixels, bytes, hipMemcpyHostToDevice, C->stream));
    az_status st = az_gpu_render_device(C, Cam, Filter, Cull, Opts, B.pixels.ptr);
    if (st != AZ_OK) return st;
    AZ_HIPR(C, hipMemcpyAsync(Pixels, B.pixels.ptr, bytes, hipMemcpyDeviceToHost, C->stream));
    AZ_HIPR(C, hipStreamSynchronize(C->stream));
    return AZ_OK;
}

az_status az_gpu_set_render_tuning(az_gpu_ctx *C, u32 chunkSamples, u32 traceBlock, s32 traceWaves, s32 traceLoop) {
    if (!C) return AZ_ERR_INVALID_ARGUMENT;
    RenderTuning t = C->renderTuning;
    if (chunkSamples) t.chunkSamples = chunkSamples;
    if (traceBlock) t.traceBlock = (int)traceBlock;
    if (traceWaves >= 0) t.traceWaves = traceWaves;
    if (traceLoop >= 0) t.traceLoop = traceLoop;
    if (!render_trace_variant_exists(t.traceBlock, t.traceWaves, t.traceLoop)) return AZ_ERR_INVALID_ARGUMENT;
    C->renderTuning = t;
    return AZ_OK;
}

az_status az_gpu_set_level_mode(az_gpu_ctx *C, u32 mode) {
    if (!C || mode > AZ_LEVELS_HOST) return AZ_ERR_INVALID_ARGUMENT;
    C->levelMode = mode;
    return AZ_OK;
}

az_status az_gpu_level_graph_counts(az_gpu_ctx *C, u64 *Recorded, u64 *Launched) {
    if (!C || !Recorded || !Launched) return AZ_ERR_INVALID_ARGUMENT;
    RenderBuffers &B = buffersOf(C);
    *Recorded = B.levelCaptures;
    *Launched = B.levelLaunches;
    return AZ_OK;
}

az_status az_gpu_trim(az_gpu_ctx *C) {
    if (!C) return AZ_ERR_INVALID_ARGUMENT;
    AZ_HIPR(C, hipSetDevice(C->device));
    AZ_HIPR(C, hipStreamSynchronize(C->stream));
    RenderBuffers &B = buffersOf(C);
    B.dropLevelGraph();
    B.forEachQueue([](DeviceBuffer &b) { b.release(); });
    C->outputScratch.release();
    if (C->guide) az_guide_destroy(C->guide);
    if (C->scatterGuide) az_guide_destroy(C->scatterGuide);
    C->guide = C->scatterGuide = nullptr;
    return AZ_OK;
}

static az_status env_probe_check(az_gpu_ctx *C, const void *a, const void *b) {
    if (!C || !a || !b) return AZ_ERR_INVALID_ARGUMENT;
    if (!C->hasScene || !C->render.ready) return AZ_ERR_NO_SCENE;
    if (!C->devScene.useEnv || !C->devScene.envCdf) {
        C->lastError = "luminance sampling needs an image environment map";
        return AZ_ERR_UNSUPPORTED;
    }
    return AZ_OK;
}

az_status az_gpu_env_luminance_sample(az_gpu_ctx *C, const f32 *d_U, u32 N, f32 *d_Out) {
    az_status s = env_probe_check(C, d_U, d_Out);
    if (s != AZ_OK || N == 0) return s;
    AZ_HIPR(C, hipSetDevice(C->device));
    hipLaunchKernelGGL(k_env_lum_sample, dim3((N + kRenderBlock - 1) / kRenderBlock), dim3(kRenderBlock), 0, C->stream,
                       C->devScene, d_U, N, d_Out);
    AZ_HIPR(C, hipGetLastError());
    return AZ_OK;
}

az_status az_gpu_env_luminance_pdf(az_gpu_ctx *C, const f32 *d_Dirs, u32 N, f32 *d_Pdf) {
    az_status s = env_probe_check(C, d_Dirs, d_Pdf);
    if (s != AZ_OK || N == 0) return s;
    AZ_HIPR(C, hipSetDevice(C->device));
    hipLaunchKernelGGL(k_env_lum_pdf, dim3((N + kRenderBlock - 1) / kRenderBlock), dim3(kRenderBlock), 0, C->stream,
                       C->devScene, d_Dirs, N, d_Pdf);
    AZ_HIPR(C, hipGetLastError());
    return AZ_OK;
}

az_status az_gpu_get_render_stats(az_gpu_ctx *C, az_render_stats *Out) {
    if (!C || !Out) return AZ_ERR_INVALID_ARGUMENT;
    *Out = C->renderStats;
    return AZ_OK;
}

az_status az_gpu_set_render_memory_limit(az_gpu_ctx *C, u64 Bytes) {
    if (!C) return AZ_ERR_INVALID_ARGUMENT;
    az_status s = az_gpu_trim(C);
    if (s != AZ_OK) return s;
    buffersOf(C).queueLimit.limit = (size_t)Bytes;
    return AZ_OK;
}

az_status az_gpu_autotune_render(az_gpu_ctx *C, const camera *Cam, filter_type Filter, b32 Cull,
                                 const az_render_opts *Opts, az_tune_result *out) {
    if (!C || !Cam) return AZ_ERR_INVALID_ARGUMENT;
    if (!C->hasScene || !C->render.ready) return AZ_ERR_NO_SCENE;
    AZ_HIPR(C, hipSetDevice(C->device));
    az_render_opts o;
    memset(&o, 0, sizeof(o));
    if (Opts) o = *Opts;
    // Enough samples per pixel that the largest candidate chunk is filled
    // twice (up to the camera's own count), so chunking is part of what is
    // timed.
    const uint32_t chunks[3] = {8u << 20, 16u << 20, 32u << 20};
    const u32 nPix = Cam->Resolution.x * Cam->Resolution.y;
    const u32 camSpp = o.SampleCountOverride ? o.SampleCountOverride : Cam->SampleCount;
    u32 spp = std::max<u32>(1u, std::min<u32>(camSpp, 2u * chunks[2] / std::max<u32>(1u, nPix)));
    u32 sq = (u32)std::sqrt((double)spp);
    o.SampleCountOverride = std::max<u32>(1u, sq * sq);
    o.GuideEnable = 0;
    o.AccumulateF64 = 0;
    RenderBuffers &B = buffersOf(C);
    const size_t bytes = (size_t)Cam->Resolution.x * Cam->Resolution.y * sizeof(pixel);
    AZ_HIPR(C, B.pixels.alloc(bytes));
    AZ_HIPR(C, hipMemsetAsync(B.pixels.ptr, 0, bytes, C->stream));
    RenderTuning variants[16];
    const int nv = render_trace_variants(variants, 16);
    const RenderTuning keep = C->renderTuning;
    // Warm-up with the current setting: kernels loaded, queues allocated.
    az_status st = renderImpl(C, Cam, Filter, Cull, &o, B.pixels.ptr, false);
    if (st != AZ_OK) return st;
    hipEvent_t a = nullptr, b = nullptr;
    AZ_HIPR(C, hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return AZ_ERR_DEVICE;
    }
    az_tune_result res;
    memset(&res, 0, sizeof(res));
    double best = 1e300;
    RenderTuning bestT = keep;
    for (int v = 0; v < nv && st == AZ_OK; ++v)
        for (uint32_t cs : chunks) {
            RenderTuning t = variants[v];
            t.chunkSamples = cs;
            C->renderTuning = t;
            float ms = 0.0f;
            if (hipEventRecord(a, C->stream) != hipSuccess) st = AZ_ERR_DEVICE;
            if (st == AZ_OK) st = renderImpl(C, Cam, Filter, Cull, &o, B.pixels.ptr, false);
            if (st == AZ_OK && (hipEventRecord(b, C->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                                hipEventElapsedTime(&ms, a, b) != hipSuccess))
                st = AZ_ERR_DEVICE;
            if (st != AZ_OK) break;
            if (res.Count < 32) {
                az_tune_entry &e = res.Entries[res.Count];
                e.Block = (u32)t.traceBlock;
                e.LdsStackEntries = (u32)kRenderRing;
                e.MinWaves = t.traceWaves;
                e.ChunkSamples = cs;
                e.LoopForm = (u32)t.traceLoop;
                e.Milliseconds = ms;
                if (ms < best) {
                    best = ms;
                    bestT = t;
                    res.Best = res.Count;
                }
                res.Count++;
            }
        }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    C->renderTuning = st == AZ_OK ? bestT : keep;
    if (st == AZ_OK && out) *out = res;
    return st;
}

b32 az_render_scene(scene *Scene, camera *Camera, epsilon_values *Epsilons, u32 ThreadCount, b32 BackfaceCulling,
                    pixel *Pixels) {
    (void)ThreadCount;
    if (!Scene || !Camera || !Epsilons || !Pixels) return 0;
    az_gpu_ctx *C = nullptr;
    if (az_gpu_create(&C, 0) != AZ_OK) return 0;
    b32 ok = 0;
    if (az_gpu_upload_scene(C, Scene, Epsilons) == AZ_OK) {
        memset(Pixels, 0, sizeof(pixel) * Camera->Resolution.x * Camera->Resolution.y);
        ok = az_gpu_render(C, Camera, FilterType_Box, BackfaceCulling, nullptr, Pixels) == AZ_OK;
    }
    az_gpu_destroy(C);
    return ok;
}

}  // extern "C"
