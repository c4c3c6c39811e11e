// device_trace.h -- ray traversal and primitive tests on gfx950.
//
// Primitive arithmetic is the reference's, operation for operation, so that
// hit ids, t and barycentrics match it bit for bit when this file is compiled
// with -ffp-contract=off and IEEE division/sqrt:
//   triangle : Cramer's rule, main.cpp:147-243
//   sphere   : local quadratic + world distance, main.cpp:64-146
//   instance : motion shift, GetLocalRay (ray.h:24-66) with the renormalized
//              direction, world t as a length ratio, main.cpp:244-281
// Traversal is this project's own: a 4-wide BVH walked front to back with
// distance pruning and an LDS stack (the reference visits every box a ray
// crosses, main.cpp:282-319).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_scene.h"

namespace az {

constexpr float kFltMax = 3.402823466e+38f;   // common.h:39 f32_Max
constexpr float kBaryLower = -0.000001f;      // main.cpp:176

struct Ray3 {
    float ox, oy, oz;
    float dx, dy, dz;
    float time;
};

// Result of a closest-hit walk.
struct HitRec {
    uint32_t prim;      // primitive id or kEmpty
    uint32_t item;      // top-level item index (TLAS leaf order)
    float t;            // world t
    float tl;           // t along the local (renormalized) ray of the item
    float beta, gamma;
};

__device__ inline float len3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

// GetLocalRay (ray.h:24-66) on a ray whose origin is already motion-shifted.
__device__ inline void local_ray(const Affine &m, float ox, float oy, float oz, float dx, float dy,
                                 float dz, float &lox, float &loy, float &loz, float &ldx, float &ldy,
                                 float &ldz, float &dirScale) {
    lox = m.m[0][0] * ox + m.m[0][1] * oy + m.m[0][2] * oz + m.m[0][3];
    loy = m.m[1][0] * ox + m.m[1][1] * oy + m.m[1][2] * oz + m.m[1][3];
    loz = m.m[2][0] * ox + m.m[2][1] * oy + m.m[2][2] * oz + m.m[2][3];
    float x = m.m[0][0] * dx + m.m[0][1] * dy + m.m[0][2] * dz;
    float y = m.m[1][0] * dx + m.m[1][1] * dy + m.m[1][2] * dz;
    float z = m.m[2][0] * dx + m.m[2][1] * dy + m.m[2][2] * dz;
    float l = len3(x, y, z);
    dirScale = l;
    if (l != 0.0f) {
        ldx = x / l;
        ldy = y / l;
        ldz = z / l;
    } else {
        ldx = ldy = ldz = 0.0f;
    }
}

// PointToWorld (math.cpp:1774-1800).
__device__ inline void point_to_world(const Affine &m, float x, float y, float z, float &wx, float &wy,
                                      float &wz) {
    wx = m.m[0][0] * x + m.m[0][1] * y + m.m[0][2] * z + m.m[0][3];
    wy = m.m[1][0] * x + m.m[1][1] * y + m.m[1][2] * z + m.m[1][3];
    wz = m.m[2][0] * x + m.m[2][1] * y + m.m[2][2] * z + m.m[2][3];
}

// World t of a local hit point, main.cpp:258-262 / 100-103:
// Length(WorldPosition - Origin - Motion) / Length(Direction) with
// WorldPosition = PointToWorld(P) + Motion and Origin already shifted.
__device__ inline float world_t(const Affine &fwd, float px, float py, float pz, float mx, float my,
                                float mz, float sox, float soy, float soz, float dirLen) {
    float wx, wy, wz;
    point_to_world(fwd, px, py, pz, wx, wy, wz);
    wx = wx + mx;
    wy = wy + my;
    wz = wz + mz;
    float ex = (wx - sox) - mx, ey = (wy - soy) - my, ez = (wz - soz) - mz;
    return len3(ex, ey, ez) / dirLen;
}

// Triangle test of main.cpp:154-189 against the ray (o, d).  Accepts when the
// barycentrics pass the reference bounds and eps < t; the caller compares t
// with its current best.
__device__ inline bool tri_test(const TriPacked &tp, float ox, float oy, float oz, float dx, float dy,
                                float dz, float eps, float &t, float &beta, float &gamma) {
    const float A = tp.a, B = tp.b, C = dx;
    const float D = tp.d, E = tp.e, F = dy;
    const float G = tp.g, H = tp.h, I = dz;
    float detM = (A * E * I) + (D * H * C) + (G * B * F) - (C * E * G) - (F * H * A) - (I * B * D);
    float K = (tp.p0x - ox), M = (tp.p0y - oy), N = (tp.p0z - oz);
    float detB = (K * E * I) + (M * H * C) + (N * B * F) - (C * E * N) - (F * H * K) - (I * B * M);
    beta = detB / detM;
    if (!(beta > kBaryLower && beta <= 1.0f)) return false;
    float detG = (A * M * I) + (D * N * C) + (G * K * F) - (C * M * G) - (F * N * A) - (I * K * D);
    gamma = detG / detM;
    if (!(gamma > kBaryLower && gamma <= 1.0f)) return false;
    float detT = (A * E * N) + (D * H * K) + (G * B * M) - (K * E * G) - (M * H * A) - (N * B * D);
    float alpha = 1.0f - beta - gamma;
    if (!(alpha > kBaryLower)) return false;
    t = detT / detM;
    return t > eps;
}

// RayAABBIntersection (ray.h:68-125) on a box given as min xyz, max xyz:
// the slabs with 1/d, near and far planes chosen by the sign of 1/d, the
// same comparisons (so NaN slab distances fall where the reference's do),
// hit iff TEnterMax <= TExitMin and TExitMin > 0.
__device__ inline bool ref_box_hit(const float *b, float ox, float oy, float oz, float dx, float dy, float dz) {
    const float ix = 1.0f / dx, iy = 1.0f / dy, iz = 1.0f / dz;
    const float lx = b[0] - ox, ly = b[1] - oy, lz = b[2] - oz;
    const float hx = b[3] - ox, hy = b[4] - oy, hz = b[5] - oz;
    float ex, ey, ez, fx, fy, fz;
    if (ix > 0.0f) { ex = lx * ix; fx = hx * ix; } else { ex = hx * ix; fx = lx * ix; }
    if (iy > 0.0f) { ey = ly * iy; fy = hy * iy; } else { ey = hy * iy; fy = ly * iy; }
    if (iz > 0.0f) { ez = lz * iz; fz = hz * iz; } else { ez = hz * iz; fz = lz * iz; }
    float en = (ex > ey) ? ex : ey;
    en = (ez > en) ? ez : en;
    float fa = (fx < fy) ? fx : fy;
    fa = (fz < fa) ? fz : fa;
    return en <= fa && fa > 0.0f;
}

// Whether the reference's box test on an instance (main.cpp:245) passes for
// a ray whose hit on that instance is at world t w.  A hit point inside the
// box by a margin in every dimension of non-zero extent passes the slab test
// whatever the rounding of the slab distances (its t lies inside every slab
// interval; a dimension of zero extent gives the interval [t, t] at the hit's
// own plane crossing); nearer the boundary the test itself decides.
__device__ inline bool ref_box_accepts(const float *b, float ox, float oy, float oz, float dx, float dy, float dz,
                                       float w) {
    const float p[3] = {ox + dx * w, oy + dy * w, oz + dz * w}, o[3] = {ox, oy, oz};
    bool inside = true;
    for (int k = 0; k < 3; ++k) {
        const float lo = b[k], hi = b[3 + k];
        if (lo == hi) continue;
        const float m = 1e-5f * ((hi - lo) + fabsf(p[k] - o[k]) + fabsf(lo) + fabsf(hi));
        inside = inside && p[k] > lo + m && p[k] < hi - m;
    }
    return inside || ref_box_hit(b, ox, oy, oz, dx, dy, dz);
}

// The reference picks the closest hit by a recursive minimum that keeps the
// right child on equal t (main.cpp:303-307), so among exactly equal t the
// right-most leaf of its BVH wins.  Ranks order the leaves left to right:
// the primitive id inside a mesh, the entity's leaf position at the top level.
__device__ inline bool closer(float t, uint32_t rank, float bestT, uint32_t bestRank) {
    return t < bestT || (t == bestT && rank > bestRank);
}

// Sphere test of main.cpp:64-103 on an item; returns the world t (and the
// local t) when the reference would report a hit.
__device__ inline bool sphere_test(const TopItem &it, const Ray3 &r, float eps, float dirLen, float &tw,
                                   float &tl) {
    float mx = r.time * it.motion[0], my = r.time * it.motion[1], mz = r.time * it.motion[2];
    float sox = r.ox - mx, soy = r.oy - my, soz = r.oz - mz;
    float lox, loy, loz, ldx, ldy, ldz, sc;
    local_ray(it.inv, sox, soy, soz, r.dx, r.dy, r.dz, lox, loy, loz, ldx, ldy, ldz, sc);
    float distx = lox - it.center[0], disty = loy - it.center[1], distz = loz - it.center[2];
    float A = ldx * ldx + ldy * ldy + ldz * ldz;
    float B = distx * ldx + disty * ldy + distz * ldz;
    float C = (distx * distx + disty * disty + distz * distz) - it.radius * it.radius;
    float disc = (B * B) - (A * C);
    if (!(disc > 0.0f)) return false;
    float sd = sqrtf(disc);
    float T1 = (-B + sd) / (A), T2 = (-B - sd) / (A);
    float t = kFltMax;
    if (T1 > eps && T2 > eps) t = T1 < T2 ? T1 : T2;
    else if (T1 > eps) t = T1;
    else if (T2 > eps) t = T2;
    if (!(t > eps && t < kFltMax)) return false;
    float hx = lox + (ldx * t), hy = loy + (ldy * t), hz = loz + (ldz * t);
    float w = world_t(it.fwd, hx, hy, hz, mx, my, mz, sox, soy, soz, dirLen);
    if (!(w > eps && w < kFltMax)) return false;
    tw = w;
    tl = t;
    return true;
}

// Slab test of one node against (o, 1/d), children hit within [0, tclip].
// Near/far planes are chosen by the sign of 1/d as in RayAABBIntersection
// (ray.h:78-102); empty children (lo=+inf, hi=-inf) never hit.
struct NodeHits {
    float key[4];       // the visiting order's key; +inf for a miss
    uint32_t ref[4];
};

// Byte offsets, within a node, of the near-plane float4 of each axis for a
// ray with inverse direction i: the node holds lo x, hi x, lo y, hi y, lo z,
// hi z (16 bytes each), and the far plane of an axis sits at near ^ 16.  A
// ray's signs do not change during a walk, so the near/far choice is made
// once per ray (and instance) as an address instead of per node as selects.
struct NearPlanes {
    uint32_t x, y, z;
};

__device__ inline NearPlanes near_planes(float ix, float iy, float iz) {
    return NearPlanes{ix >= 0.0f ? 0u : 16u, iy >= 0.0f ? 32u : 48u, iz >= 0.0f ? 64u : 80u};
}

// Order in which a node's children hit are visited.  A closest-hit walk
// takes the nearest entry first, which tightens its clip distance soonest.
// An occlusion query has the same answer in any order and stops at its
// first hit, so it takes the child most likely to block the ray first:
// measured on a 1.3M-triangle mesh and on the interior render's shadow rays,
//   kOverlapFirst (longest stretch of the ray inside the box):
//     camera rays +2 %, rays from random points +2 %;
//   kFarExitFirst (furthest box exit; closest-hit lanes nearest first):
//     shadow rays from surface points to lights 9.8 -> 5.7 nodes per ray,
//     where nearest-first spends its steps in the boxes around the origin.
enum ChildOrder { kNearFirst = 0, kOverlapFirst = 1, kFarExitFirst = 2 };

// Node byte offsets are 32-bit: the upload keeps BVHs below 2^25 nodes.
// far (kFarExitFirst only): this lane's ray is an occlusion query.
template <int ORDER = kNearFirst>
__device__ inline int node_test(const Node4 *__restrict__ nodes, uint32_t idx, float ox, float oy,
                                float oz, float ix, float iy, float iz, NearPlanes np, float tclip, NodeHits &h,
                                bool far = false) {
    const char *base = reinterpret_cast<const char *>(nodes);
    const uint32_t o = idx * (uint32_t)sizeof(Node4);
    const float4 nx = *reinterpret_cast<const float4 *>(base + (o + np.x));
    const float4 fx = *reinterpret_cast<const float4 *>(base + (o + (np.x ^ 16u)));
    const float4 ny = *reinterpret_cast<const float4 *>(base + (o + np.y));
    const float4 fy = *reinterpret_cast<const float4 *>(base + (o + (np.y ^ 16u)));
    const float4 nz = *reinterpret_cast<const float4 *>(base + (o + np.z));
    const float4 fz = *reinterpret_cast<const float4 *>(base + (o + (np.z ^ 16u)));
    const uint4 ch = *reinterpret_cast<const uint4 *>(base + (o + 96u));
    float tn[4], tf[4];
#define AZ_SLAB(k, c)                                                                   \
    {                                                                                   \
        float a0 = (nx.c - ox) * ix, a1 = (fx.c - ox) * ix;                             \
        float b0 = (ny.c - oy) * iy, b1 = (fy.c - oy) * iy;                             \
        float c0 = (nz.c - oz) * iz, c1 = (fz.c - oz) * iz;                             \
        tn[k] = fmaxf(fmaxf(a0, b0), fmaxf(c0, 0.0f));                                  \
        tf[k] = fminf(fminf(a1, b1), fminf(c1, tclip));                                 \
    }
    AZ_SLAB(0, x)
    AZ_SLAB(1, y)
    AZ_SLAB(2, z)
    AZ_SLAB(3, w)
#undef AZ_SLAB
    uint32_t cr[4] = {ch.x, ch.y, ch.z, ch.w};
    int n = 0;
    // An empty slot's box (lo = +inf, hi = -inf, as the builder writes it)
    // gives tn = +inf and tf = -inf whatever the ray's signs, so it never
    // passes this test and needs no check of its reference.
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        bool hit = tn[k] <= tf[k];
        // A hit's key stays below +inf, so that no hit sorts with the misses:
        // tn <= tf <= the clip distance, finite but for a batch occlusion
        // query's own tmax, where tn = tf = +inf (nothing reachable) makes
        // the overlap key NaN, taken as -FLT_MAX.
        const float key = ORDER == kNearFirst      ? tn[k]
                          : ORDER == kOverlapFirst ? fmaxf(tn[k] - tf[k], -kFltMax)
                                                   : (far ? -tf[k] : tn[k]);
        h.key[k] = hit ? key : INFINITY;
        h.ref[k] = cr[k];
        n += hit ? 1 : 0;
    }
    // Sort the four (key, ref) pairs ascending; misses sort last.
#define AZ_CSWAP(i, j)                                                              \
    if (h.key[j] < h.key[i]) {                                                      \
        float tt = h.key[i];                                                        \
        h.key[i] = h.key[j];                                                        \
        h.key[j] = tt;                                                              \
        uint32_t rr = h.ref[i];                                                     \
        h.ref[i] = h.ref[j];                                                        \
        h.ref[j] = rr;                                                              \
    }
    AZ_CSWAP(0, 1)
    AZ_CSWAP(2, 3)
    AZ_CSWAP(0, 2)
    AZ_CSWAP(1, 3)
    AZ_CSWAP(1, 2)
#undef AZ_CSWAP
    return n;
}

// Per-lane traversal stack: the top RING entries live in LDS, laid out
// [slot][lane] so that lanes at the same depth touch consecutive words; older
// entries are spilled to a private (scratch) array.  Entry i of the logical
// stack sits in ring slot i % RING while it is among the top RING entries,
// else in spill[i - RING].  Capacity RING + SPILL covers the deepest stack
// the uploaded BVH can need (checked at upload).


template <int RING, int BLOCK>
struct RingStack {
    uint32_t *base;   // &smem[0][lane]
    uint32_t spill[kSpill];
    __device__ inline void push(int &sp, uint32_t v) {
        uint32_t *slot = base + (sp & (RING - 1)) * BLOCK;
        if (sp >= RING) spill[sp - RING] = *slot;
        *slot = v;
        ++sp;
    }
    __device__ inline uint32_t pop(int &sp) {
        --sp;
        uint32_t *slot = base + (sp & (RING - 1)) * BLOCK;
        uint32_t v = *slot;
        if (sp >= RING) *slot = spill[sp - RING];
        return v;
    }
};

struct TraceCounters {
    uint32_t nodes;
    uint32_t tris;
    uint32_t prims;
};

// Two-level walk over a wave's range of rays with refill.  Lanes of a wave
// take rays [begin, end) in order; a lane whose ray finishes takes the next
// unclaimed one at the end of the same step (one ballot, no atomics), so a
// wave keeps its lanes busy until the range runs out instead of waiting for
// its longest ray.  MODE 0: closest hit with eps < t (tMax ignored); MODE 1:
// any hit with eps < t < tMax (world t); MODE 2: per ray, as the pool says;
// MODE 3: any hit, children in MODE 2's order for occlusion queries (a
// renderer's shadow rays when none needs a closest hit).
// Pool provides
//   bool fetch(uint32_t id, Ray3 &r, float &tMax)   -> true: any-hit ray
//   void finish(uint32_t id, bool found, const HitRec &h)
// Every lane of the wave must call it (it ends when the whole range is done).
// ldsLane is the lane's column of a [RING + 4][BLOCK] LDS array: RING stack
// slots, one word for an excluded instance and three for the world ray's
// reciprocal direction.
template <int MODE, bool COUNT, int RING, int BLOCK, bool WW, class Pool>
__device__ inline void trace_pool(const DeviceGeometry &G, Pool &pool, uint32_t begin, uint32_t end, float eps,
                                  bool cull, uint32_t *ldsLane, TraceCounters &cnt) {
    constexpr int kOrder = MODE == 0 ? kNearFirst : MODE == 1 ? kOverlapFirst : kFarExitFirst;
    constexpr bool kAnyOnly = MODE == 1 || MODE == 3;
    RingStack<RING, BLOCK> stack;
    stack.base = ldsLane;
    uint32_t next = begin;                          // wave-uniform: first unclaimed ray
    uint32_t id = 0;
    bool have = false;
    Ray3 r{};
    bool any = kAnyOnly;
    // An instance whose hit the reference's own box test rejected: the walk
    // of the ray is redone once without it.  Kept in the lane's LDS word
    // after its ring (read only at top-level instance leaves), not in a
    // register, so that the walk's register budget is unchanged.
    uint32_t *const excl = ldsLane + RING * BLOCK;
    // The world ray's reciprocal direction, kept in the three LDS words
    // after excl for the returns from instances (three divisions each).
    float *const winv = reinterpret_cast<float *>(ldsLane + (RING + 1) * BLOCK);
    float dirLen = 1.0f, bestW = kFltMax;
    uint32_t bestRank = 0;
    bool found = false;
    HitRec out{};
    // Current walk ray: world ray in the TLAS, instance-local ray in a BLAS.
    float ox = 0.f, oy = 0.f, oz = 0.f, dx = 0.f, dy = 0.f, dz = 0.f, ix = 0.f, iy = 0.f, iz = 0.f;
    NearPlanes np{0u, 32u, 64u};
    // Instance state while inside a BLAS.
    bool inBlas = false;
    uint32_t curItem = 0;
    float sox = 0.f, soy = 0.f, soz = 0.f, mx = 0.f, my = 0.f, mz = 0.f, dscale = 1.0f;
    float localBest = kFltMax;                      // closest local t inside the current BLAS
    uint32_t localPrim = kEmpty;
    float localBeta = 0.f, localGamma = 0.f;
    int sp = 0;
    uint32_t cur = kEmpty;
    for (;;) {
        // Refill the lanes without a ray.
        const uint64_t need = __ballot(!have);
        if (need) {
            const uint32_t rank =
                __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
            if (!have && next + rank < end) {
                id = next + rank;
                have = true;
                float tMax = kFltMax;
                const bool a = pool.fetch(id, r, tMax);
                any = MODE == 2 ? a : kAnyOnly;
                dirLen = len3(r.dx, r.dy, r.dz);
                *excl = kEmpty;
                bestW = any ? tMax : kFltMax;
                bestRank = any ? kEmpty : 0u;       // any: a hit must be strictly below tMax
                found = false;
                out.prim = kEmpty;
                out.item = kEmpty;
                out.t = 0.0f;
                out.tl = 0.0f;
                out.beta = 0.0f;
                out.gamma = 0.0f;
                ox = r.ox;
                oy = r.oy;
                oz = r.oz;
                dx = r.dx;
                dy = r.dy;
                dz = r.dz;
                ix = 1.0f / dx;
                iy = 1.0f / dy;
                iz = 1.0f / dz;
                winv[0] = ix;
                winv[BLOCK] = iy;
                winv[2 * BLOCK] = iz;
                np = near_planes(ix, iy, iz);
                inBlas = false;
                localBest = kFltMax;
                localPrim = kEmpty;
                sp = 0;
                cur = G.tlasRoot;
            }
            next += (uint32_t)__popcll(need);
        }
        if (__ballot(have) == 0) break;
        if (!have) continue;
        // WW ("while-while", Aila and Laine 2009): internal nodes first, a
        // lane descending until it reaches a leaf, a BLAS exit marker or the
        // end of its walk, so that the wave's node tests run together and so
        // do the leaf tests that follow.  Otherwise ("if-if") one node or one
        // leaf per step.  Which is faster depends on the scene.
        bool finished = false, advanced = false;
        // Scheduling of the kinds of step.  Node tests run while the lanes at
        // an internal node are more than 2/3 of the lanes with a ray; the other
        // kinds (a leaf's primitives, entering or leaving an instance, the end
        // of a ray) wait until their lanes are at least 1.5 times the node
        // lanes, or no lane is at a node.  Each kind of step then runs with
        // many lanes at once instead of a few lanes each step: on a
        // 1.3M-triangle mesh this raised the rays traced per second by 10-25 %
        // (if-if) and 20-30 % (while-while).  The walk of each ray is unchanged.
        const int nLanes = __popcll(__ballot(true));
        if (!WW) {
            const bool atNode = cur != kEmpty && !ref_is_leaf(cur);
            const int nNode = __popcll(__ballot(atNode));
            if (!atNode && nNode != 0 && 2 * (nLanes - nNode) < 3 * nNode) continue;
        }
        while (WW && cur != kEmpty && !ref_is_leaf(cur)) {
            {
                const int nNode = __popcll(__ballot(true));     // the lanes still in this loop
                if (2 * (nLanes - nNode) >= 3 * nNode) break;
            }
            if (COUNT) cnt.nodes++;
            float tclip = inBlas ? localBest : bestW;
            NodeHits h;
            int n = node_test<kOrder>(G.nodes, cur, ox, oy, oz, ix, iy, iz, np, tclip, h, any);
            if (n > 0) {
                cur = h.ref[0];
                if (n > 3) stack.push(sp, h.ref[3]);
                if (n > 2) stack.push(sp, h.ref[2]);
                if (n > 1) stack.push(sp, h.ref[1]);
            } else if (sp == 0) {
                finished = true;
                break;
            } else {
                cur = stack.pop(sp);
            }
        }
        if (WW && !finished && cur != kEmpty && !ref_is_leaf(cur)) continue;   // left the loop at a node
        // Then one leaf or exit marker (or, if-if, one node).
        if (finished) {
        } else if (cur == kEmpty) {
            // Leaving a BLAS: convert its closest local hit (main.cpp:256-276).
            if (!any && localPrim != kEmpty) {
                const TopItem &it = G.items[curItem & 0x7FFFFFFFu];
                float px = ox + (dx * localBest), py = oy + (dy * localBest), pz = oz + (dz * localBest);
                float w = world_t(it.fwd, px, py, pz, mx, my, mz, sox, soy, soz, dirLen);
                if (w > eps && closer(w, it.rank, bestW, bestRank)) {
                    bestW = w;
                    bestRank = it.rank;
                    found = true;
                    out.prim = localPrim;
                    out.item = curItem;
                    out.t = w;
                    out.tl = localBest;
                    out.beta = localBeta;
                    out.gamma = localGamma;
                }
            }
            inBlas = false;
            ox = r.ox;
            oy = r.oy;
            oz = r.oz;
            dx = r.dx;
            dy = r.dy;
            dz = r.dz;
            ix = winv[0];
            iy = winv[BLOCK];
            iz = winv[2 * BLOCK];
            np = near_planes(ix, iy, iz);
        } else if (!WW && !ref_is_leaf(cur)) {
            if (COUNT) cnt.nodes++;
            float tclip = inBlas ? localBest : bestW;
            NodeHits h;
            int n = node_test<kOrder>(G.nodes, cur, ox, oy, oz, ix, iy, iz, np, tclip, h, any);
            if (n > 0) {
                cur = h.ref[0];
                if (n > 3) stack.push(sp, h.ref[3]);
                if (n > 2) stack.push(sp, h.ref[2]);
                if (n > 1) stack.push(sp, h.ref[1]);
                advanced = true;
            }
        } else if (!inBlas) {
            // TLAS leaves hold exactly one item.
            uint32_t itemIdx = leaf_first(cur);
            const TopItem &it = G.items[itemIdx];
            if (COUNT) cnt.prims++;
            if (it.kind == kItemSphere) {
                float tw, tl;
                if (sphere_test(it, r, eps, dirLen, tw, tl) && closer(tw, it.rank, bestW, bestRank)) {
                    bestW = tw;
                    bestRank = it.rank;
                    found = true;
                    out.prim = G.triTotal + it.entityId;
                    out.item = itemIdx;
                    out.t = tw;
                    out.tl = tl;
                    out.beta = 0.0f;
                    out.gamma = 0.0f;
                    if (any) finished = true;
                }
            } else if (itemIdx != *excl) {
                mx = r.time * it.motion[0];
                my = r.time * it.motion[1];
                mz = r.time * it.motion[2];
                sox = r.ox - mx;
                soy = r.oy - my;
                soz = r.oz - mz;
                local_ray(it.inv, sox, soy, soz, r.dx, r.dy, r.dz, ox, oy, oz, dx, dy, dz, dscale);
                ix = 1.0f / dx;
                iy = 1.0f / dy;
                iz = 1.0f / dz;
                np = near_planes(ix, iy, iz);
                inBlas = true;
                // Bit 31 marks an instance whose hits take the reference's box test.
                curItem = itemIdx | ((it.kind & kItemBoxTest) ? 0x80000000u : 0u);
                // Local distances scale with |M^-1 d|; the bound is widened so
                // that rounding never prunes a hit that could still win.
                localBest = bestW >= kFltMax ? kFltMax : bestW * dscale * 1.00001f + 1e-30f;
                localPrim = kEmpty;
                stack.push(sp, kEmpty);
                cur = it.blasRoot;
                advanced = true;
            }
        } else {
            const uint32_t first = leaf_first(cur), count = leaf_count(cur);
            for (uint32_t k = 0; k < count; ++k) {
                const TriPacked tp = G.tris[first + k];
                if (COUNT) cnt.tris++;
                if (cull) {
                    const float *fn = G.faceNormals + 3 * (size_t)(first + k);
                    if (!(dx * fn[0] + dy * fn[1] + dz * fn[2] < 0.0f)) continue;
                }
                float t, b, g;
                if (tri_test(tp, ox, oy, oz, dx, dy, dz, eps, t, b, g) &&
                    (any ? t < localBest : closer(t, tp.prim, localBest, localPrim))) {
                    if (any) {
                        const TopItem &it = G.items[curItem & 0x7FFFFFFFu];
                        float px = ox + (dx * t), py = oy + (dy * t), pz = oz + (dz * t);
                        float w = world_t(it.fwd, px, py, pz, mx, my, mz, sox, soy, soz, dirLen);
                        if (w > eps && w < bestW) {
                            found = true;
                            out.prim = tp.prim;
                            out.item = curItem;
                            out.t = w;
                            out.tl = t;
                            out.beta = b;
                            out.gamma = g;
                            finished = true;
                            break;
                        }
                    } else {
                        localBest = t;
                        localPrim = tp.prim;
                        localBeta = b;
                        localGamma = g;
                    }
                }
            }
        }
        if (!finished && !advanced) {
            if (sp == 0) finished = true;
            else cur = stack.pop(sp);
        }
        if (finished && found && (out.item >> 31) && *excl == kEmpty &&
            !ref_box_accepts(G.refBoxes + 6 * (size_t)(out.item & 0x7FFFFFFFu), r.ox, r.oy, r.oz, r.dx, r.dy, r.dz,
                             out.t)) {
            // The reference never enters this instance: walk the ray again
            // without it.
            *excl = out.item & 0x7FFFFFFFu;
            finished = false;
            float tMax = kFltMax;
            (void)pool.fetch(id, r, tMax);
            bestW = any ? tMax : kFltMax;
            bestRank = any ? kEmpty : 0u;
            found = false;
            out.prim = kEmpty;
            out.item = kEmpty;
            ox = r.ox;
            oy = r.oy;
            oz = r.oz;
            dx = r.dx;
            dy = r.dy;
            dz = r.dz;
            ix = 1.0f / dx;
            iy = 1.0f / dy;
            iz = 1.0f / dz;
            np = near_planes(ix, iy, iz);
            inBlas = false;
            localBest = kFltMax;
            localPrim = kEmpty;
            sp = 0;
            cur = G.tlasRoot;
        }
        if (finished) {
            out.item &= 0x7FFFFFFFu;
            pool.finish(id, found, out);
            have = false;
        }
    }
}

}  // namespace az
