// az_oracle.cpp -- TEST INFRASTRUCTURE ONLY: CPU restatement of the azimuth
// render hot path (reference src/azimuth_main.cpp:35-1424, src/azimuth_ray.h,
// src/azimuth_texture.h:106-357, the used parts of src/azimuth_math.cpp).
// Built with g++ -O2 -ffp-contract=off (no FMA contraction, IEEE division and
// square root), which defines the reference's bits for this project.
//
// Defined deviations from the reference (shared with the GPU path):
//   * Random01() (math.cpp:1-8) -> counter-addressed PCG32: the draw at
//     (seed, pixel, sample, path-tree node, dim).  The camera ray is node 0;
//     the child of node n of kind k (0 scatter, 1 mirror, 2 refraction,
//     3 reflection) is 5n+k+1, so the root vertex is node 1.  Dims count the
//     draws of one vertex in the order of the reference's calls, with the
//     two draws of one expression taken left to right.
//   * hit records carry the primitive id and the barycentrics (beta, gamma).
//   * Perlin: modulo 256 on possibly negative lattice indices is a proper
//     modulo; the gradient table is drawn from a fixed PCG32 stream; the
//     permutation is the texture's Perm array as given.
#include "az_oracle.h"

#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

#include "azimuth_platform_posix.h"

namespace {

// The reference reaches threads and the float atomic through its platform
// table (platform.h:26-52); the oracle goes through the POSIX one.
const az_platform_api &platform() {
    static const az_platform_api api = [] {
        az_platform_api a;
        az_platform_posix_api(&a);
        return a;
    }();
    return api;
}

const f64 kPi = 3.14159265359;               // math.h:3
const f32 kPiF = (f32)kPi;                   // math.h:4
const f32 kMaxF = 3.402823466e+38f;          // common.h:39

// ---- vector helpers (math.cpp), same operation order -------------------
struct V { f32 x, y, z; };
inline V mk(f32 x, f32 y, f32 z) { return V{x, y, z}; }
inline V mk(vec3 a) { return V{a.x, a.y, a.z}; }
inline vec3 tov(V a) { vec3 r = {a.x, a.y, a.z}; return r; }
inline V operator+(V a, V b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V operator-(V a, V b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V operator-(V a) { return mk(-a.x, -a.y, -a.z); }
inline V operator*(V a, f32 s) { return mk(a.x * s, a.y * s, a.z * s); }
inline V operator*(f32 s, V a) { return mk(s * a.x, s * a.y, s * a.z); }
inline V had(V a, V b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
inline f32 dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(V a, V b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline f32 length(V a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }   // math.cpp:1289-1296
inline V normalize(V a) {                                                      // math.cpp:1298-1313
    f32 l = length(a);
    if (l != 0.0f) return mk(a.x / l, a.y / l, a.z / l);
    return mk(0.0f, 0.0f, 0.0f);
}
inline V reflect(V v, V r) { return ((2.0f * dot(v, r)) * r) - v; }             // math.cpp:1335-1348
inline f32 sq(f32 v) { return v * v; }
inline f32 Exp(f32 v) { return (f32)exp((f64)v); }                              // math.cpp:149-156
inline f32 Asin(f32 v) { return (f32)asin((f64)v); }                            // math.cpp:221-228
inline f32 Acos(f32 v) { return (f32)acos((f64)v); }                            // math.cpp:239-246
// Sin, Cos, Atan and Power (math.cpp:136-147, 158-219, 248-260) wrap the
// platform's float functions, whose last bit is the C library's own choice;
// this restatement defines each as the double-precision function of the
// float argument rounded to float (the correctly rounded value in all but
// about one case in 2^29); the device computes the same values with its own
// double-precision routines.
inline f32 SinF(f32 v) { return (f32)sin((f64)v); }
inline f32 CosF(f32 v) { return (f32)cos((f64)v); }
inline void SinCosF(f32 v, f32 &sv, f32 &cv) {   // Sin and Cos of one angle (as the device computes them)
    f64 sd, cd;
    sincos((f64)v, &sd, &cd);
    sv = (f32)sd;
    cv = (f32)cd;
}
inline f32 Atan2F(f32 y, f32 x) { return (f32)atan2((f64)y, (f64)x); }
inline f32 PowF(f32 x, f32 y) { return (f32)pow((f64)x, (f64)y); }
inline f32 minf_(f32 a, f32 b) { return a < b ? a : b; }                        // math.h:6
inline f32 axis(V v, u32 d) { return d == 0 ? v.x : (d == 1 ? v.y : v.z); }

V pointToWorld(V v, const mat4 &m) {                                            // math.cpp:1774-1800
    return mk(m.E[0][0] * v.x + m.E[0][1] * v.y + m.E[0][2] * v.z + m.E[0][3],
              m.E[1][0] * v.x + m.E[1][1] * v.y + m.E[1][2] * v.z + m.E[1][3],
              m.E[2][0] * v.x + m.E[2][1] * v.y + m.E[2][2] * v.z + m.E[2][3]);
}
V vectorToWorld(V v, const mat4 &m) {                                           // math.cpp:1802-1825
    return mk(m.E[0][0] * v.x + m.E[0][1] * v.y + m.E[0][2] * v.z,
              m.E[1][0] * v.x + m.E[1][1] * v.y + m.E[1][2] * v.z,
              m.E[2][0] * v.x + m.E[2][1] * v.y + m.E[2][2] * v.z);
}
V normalToWorld(V v, const mat4 &inv) {                                         // math.cpp:1827-1843
    V r = mk(inv.E[0][0] * v.x + inv.E[1][0] * v.y + inv.E[2][0] * v.z,
             inv.E[0][1] * v.x + inv.E[1][1] * v.y + inv.E[2][1] * v.z,
             inv.E[0][2] * v.x + inv.E[1][2] * v.y + inv.E[2][2] * v.z);
    return normalize(r);
}
void orthonormalBasis(V w, V *u, V *v) {                                        // math.cpp:1845-1863
    if (w.z < 0.0) {
        const f32 a = 1.0f / (1.0f - w.z);
        const f32 b = w.x * w.y * a;
        *u = mk(1.0f - w.x * w.x * a, -b, w.x);
        *v = mk(b, w.y * w.y * a - 1.0f, -w.y);
    } else {
        const f32 a = 1.0f / (1.0f + w.z);
        const f32 b = -w.x * w.y * a;
        *u = mk(1.0f - w.x * w.x * a, b, -w.x);
        *v = mk(b, 1.0f - w.y * w.y * a, -w.y);
    }
}
inline b32 bitsNonZero(vec3 v) {                                                // main.cpp:1092-1094
    u32 a, b, c;
    memcpy(&a, &v.x, 4);
    memcpy(&b, &v.y, 4);
    memcpy(&c, &v.z, 4);
    return (a | b | c) != 0;
}

// ---- PCG32 (O'Neill 2014, pcg32_srandom_r / pcg32_random_r; jump-ahead by
// Brown 1994 as in pcg_advance_lcg_64) ------------------------------------
const u64 kMult = 6364136223846793005ULL;
struct Pcg {
    u64 state, inc;
    u32 next() {
        u64 old = state;
        state = old * kMult + inc;
        u32 xs = (u32)(((old >> 18u) ^ old) >> 27u);
        u32 rot = (u32)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    void seed(u64 initstate, u64 initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    void advance(u64 delta) {
        u64 accMult = 1u, accPlus = 0u, curMult = kMult, curPlus = inc;
        while (delta > 0) {
            if (delta & 1u) {
                accMult *= curMult;
                accPlus = accPlus * curMult + curPlus;
            }
            curPlus = (curMult + 1u) * curPlus;
            curMult *= curMult;
            delta >>= 1u;
        }
        state = accMult * state + accPlus;
    }
    f32 uniform() { return (f32)(next() >> 8) * (1.0f / 16777216.0f); }
};

// The draws of one path vertex.
struct VertexRng {
    Pcg p;
    VertexRng(u64 seed, u32 pixel, u32 sample, u32 node) {
        p.seed(seed, pixel);
        p.advance(((u64)sample << 38) + ((u64)node << 6));
    }
    f32 next() { return p.uniform(); }
};

inline u32 childNode(u32 n, u32 kind) { return n * 5u + kind + 1u; }

// ---- Perlin (texture.h:106-195) -----------------------------------------
struct PerlinTable {
    V g[AZ_PERLIN_GRADIENT_SIZE];
    PerlinTable() {
        Pcg p;
        p.seed(0x5045524CULL, 256u);
        for (u32 i = 0; i < AZ_PERLIN_GRADIENT_SIZE; ++i) {
            f32 a = p.uniform(), b = p.uniform(), c = p.uniform();
            V v = mk((a - 0.5f) * 2.0f, (b - 0.5f) * 2.0f, (c - 0.5f) * 2.0f);
            g[i] = normalize(v);
        }
    }
};
const PerlinTable &perlinTable() {
    static PerlinTable t;
    return t;
}
inline u32 mod256(s32 v) { return (u32)(((v % 256) + 256) % 256); }
inline u32 perlinHash(const u32 *P, s32 I, s32 J, s32 K) {                     // texture.h:134-144
    return P[mod256(I + (s32)P[mod256(J + (s32)P[mod256(K)])])];
}
inline f32 perlinWeight(f32 v) {                                                // texture.h:146-154
    v = fabsf(v);
    return 1.0f - ((((6.0f * v) - 15.0f) * v + 10.0f) * v * v * v);
}
f32 samplePerlin(const texture *T, V P) {                                       // texture.h:156-195
    f32 r = 0.0f;
    P = P * T->ScalingFactor;
    s32 I = (s32)floorf(P.x), J = (s32)floorf(P.y), K = (s32)floorf(P.z);
    const PerlinTable &tab = perlinTable();
    for (s32 z = 0; z < 2; ++z)
        for (s32 y = 0; y < 2; ++y)
            for (s32 x = 0; x < 2; ++x) {
                u32 h = perlinHash(T->Perm, I + x, J + y, K + z) & (AZ_PERLIN_GRADIENT_SIZE - 1);
                V e = tab.g[h];
                V v = mk(P.x - (f32)(I + x), P.y - (f32)(J + y), P.z - (f32)(K + z));
                r += dot(e, v) * perlinWeight(v.x) * perlinWeight(v.y) * perlinWeight(v.z);
            }
    if (T->PerlinAppearance == TexturePerlinAppearance_Patch) r = (r + 1.0f) / 2.0f;
    else r = fabsf(r);
    return r;
}

// ---- 2D textures (texture.h:197-357) -------------------------------------
V fetch2D(const texture *T, u32 x, u32 y) {
    u32 idx = (y * T->Width + x) * 3;
    if (T->PixelFormat == PixelFormat_RGB_U8) {
        const u8 *p = (const u8 *)T->Pixels;
        return mk((f32)p[idx], (f32)p[idx + 1], (f32)p[idx + 2]);
    }
    const f32 *p = (const f32 *)T->Pixels;
    return mk(p[idx], p[idx + 1], p[idx + 2]);
}
V sampleTexture2D(const texture *T, f32 u, f32 v, bool norm) {                   // texture.h:239-339
    V r = mk(0.0f, 0.0f, 0.0f);
    if (T->WrapMode == TextureWrapMode_Clamp) {
        f32 mu = u < 1.0f ? u : 1.0f, mv = v < 1.0f ? v : 1.0f;
        u = 0.0f > mu ? 0.0f : mu;
        v = 0.0f > mv ? 0.0f : mv;
    } else {
        u = u - floorf(u);
        v = v - floorf(v);
    }
    if (T->Filter == TextureFilter_Nearest) {
        u32 fx = (u32)floorf(u * ((f32)T->Width - 1.0f) + 0.5f);
        u32 fy = (u32)floorf(v * ((f32)T->Height - 1.0f) + 0.5f);
        r = fetch2D(T, fx, fy);
    } else {
        f32 su = u * (f32)T->Width, sv = v * (f32)T->Height;
        u32 ax = (u32)(su + 0.5f) % T->Width;
        u32 ay = (u32)(sv + 0.5f) % T->Height;
        u32 bx = (ax - 1) % T->Width, by = ay % T->Height;
        u32 cx = (ax - 1) % T->Width, cy = (ay - 1) % T->Height;
        u32 dx = ax % T->Width, dy = (ay - 1) % T->Height;
        V A = fetch2D(T, ax, ay), Bc = fetch2D(T, bx, by), C = fetch2D(T, cx, cy), D = fetch2D(T, dx, dy);
        su = floorf(su + 0.5f);
        sv = floorf(sv + 0.5f);
        if (su >= T->Width) su -= T->Width;
        if (sv >= T->Height) sv -= T->Height;
        f32 fxr = su - (f32)ax, fyr = sv - (f32)ay;
        r = A * (1.0f - fxr) * (1.0f - fyr) + Bc * (fxr) * (1.0f - fyr) + C * (fxr) * (fyr) +
            D * (1.0f - fxr) * (fyr);
    }
    if (norm) {
        r.x /= T->Normalizer;
        r.y /= T->Normalizer;
        r.z /= T->Normalizer;
    }
    return r;
}
V sampleTexture(const texture *T, V P, f32 u, f32 v) {                          // texture.h:341-357
    if (T->Type == TextureType_2D) return sampleTexture2D(T, u, v, true);
    f32 n = samplePerlin(T, P);
    return mk(n, n, n);
}

// ---- scene context ---------------------------------------------------------
struct Ctx {
    const scene *S;
    std::vector<const entity *> triBase;   // per mesh: first triangle entity
    std::vector<u32> primBase;             // per mesh: first primitive id
    u32 triTotal = 0;
    explicit Ctx(const scene *s) : S(s) {
        triBase.resize(s->MeshCount);
        primBase.resize(s->MeshCount);
        for (u32 m = 0; m < s->MeshCount; ++m) {
            const entity *e = s->Meshes[m].BVH;
            while (e && e->Type == EntityType_BVHNode) e = e->Left;
            triBase[m] = e;
            primBase[m] = triTotal;
            triTotal += s->Meshes[m].TriangleCount;
        }
    }
};

struct Ray { V o, d; f32 T; };
struct Hit {
    V P, N;
    f32 u, v;
    const entity *E;
    f32 t;
    u32 prim;
    f32 beta, gamma;
};
inline Hit freshHit() {
    Hit h;
    memset(&h, 0, sizeof(h));
    h.t = kMaxF;
    h.prim = AZ_MISS;
    return h;
}

Ray getLocalRay(const Ray &r, const mat4 &m) {                                  // ray.h:24-66
    Ray out = r;
    out.o = mk(m.E[0][0] * r.o.x + m.E[0][1] * r.o.y + m.E[0][2] * r.o.z + m.E[0][3],
               m.E[1][0] * r.o.x + m.E[1][1] * r.o.y + m.E[1][2] * r.o.z + m.E[1][3],
               m.E[2][0] * r.o.x + m.E[2][1] * r.o.y + m.E[2][2] * r.o.z + m.E[2][3]);
    out.d = mk(m.E[0][0] * r.d.x + m.E[0][1] * r.d.y + m.E[0][2] * r.d.z,
               m.E[1][0] * r.d.x + m.E[1][1] * r.d.y + m.E[1][2] * r.d.z,
               m.E[2][0] * r.d.x + m.E[2][1] * r.d.y + m.E[2][2] * r.d.z);
    out.d = normalize(out.d);
    return out;
}

bool rayAABB(const aabb &b, const Ray &r) {                                     // ray.h:68-125
    V inv = mk(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    V mnd = mk(b.Min) - r.o, mxd = mk(b.Max) - r.o;
    V te, tx;
    if (inv.x > 0.0f) { te.x = mnd.x * inv.x; tx.x = mxd.x * inv.x; } else { te.x = mxd.x * inv.x; tx.x = mnd.x * inv.x; }
    if (inv.y > 0.0f) { te.y = mnd.y * inv.y; tx.y = mxd.y * inv.y; } else { te.y = mxd.y * inv.y; tx.y = mnd.y * inv.y; }
    if (inv.z > 0.0f) { te.z = mnd.z * inv.z; tx.z = mxd.z * inv.z; } else { te.z = mxd.z * inv.z; tx.z = mnd.z * inv.z; }
    f32 tEnter = (te.x > te.y) ? te.x : te.y;
    tEnter = (te.z > tEnter) ? te.z : tEnter;
    f32 tExit = (tx.x < tx.y) ? tx.x : tx.y;
    tExit = (tx.z < tExit) ? tx.z : tExit;
    return tEnter <= tExit && tExit > 0.0f;
}

// RayEntityIntersection (main.cpp:53-385).  `mesh` is the index of the mesh
// whose BVH is being walked (for primitive ids), -1 at the top level.
bool intersect(const Ctx &C, Hit *H, Ray ray, const entity *E, bool cull, f32 eps, s32 meshIdx) {
    bool result = false;
    const texture *tex = nullptr;
    V Pu = mk(0, 0, 0), Pv = mk(0, 0, 0);
    bool deferred = false;
    const scene *S = C.S;
    switch (E->Type) {
    case EntityType_Sphere: {                                                   // main.cpp:64-146
        V motion = ray.T * mk(E->MotionVector);
        ray.o = ray.o - motion;
        Ray lr = getLocalRay(ray, E->InvTransformationMatrix);
        f32 radius = E->Radius;
        V center = mk(E->Center);
        V dist = lr.o - center;
        f32 A = dot(lr.d, lr.d), B = dot(dist, lr.d), Cc = dot(dist, dist) - sq(radius);
        f32 disc = sq(B) - (A * Cc);
        if (disc > 0.0f) {
            f32 sd = sqrtf(disc);
            f32 T1 = (-B + sd) / (A), T2 = (-B - sd) / (A);
            f32 t = kMaxF;
            if (T1 > eps && T2 > eps) t = minf_(T1, T2);
            else if (T1 > eps) t = T1;
            else if (T2 > eps) t = T2;
            if (t > eps && t < H->t) {
                V hp = lr.o + (lr.d * t);
                V lhp = hp - center;
                V hn = normalize(lhp);
                V wp = pointToWorld(hp, E->TransformationMatrix);
                wp = wp + motion;
                f32 wt = length(wp - ray.o - motion) / length(ray.d);
                if (wt > eps && wt < H->t) {
                    f32 theta = Acos(hn.y);
                    f32 phi = Atan2F(hn.z, hn.x);
                    f32 twoPi = kPiF * 2.0f;
                    H->P = wp;
                    if (E->TextureIndex) {
                        tex = S->Textures + (E->TextureIndex - 1);
                        if (tex->BumpMap) {
                            f32 pyPi = kPiF * lhp.y, rPi = kPiF * radius;
                            f32 sph, cph;
                            SinCosF(phi, sph, cph);
                            Pv = mk(cph * pyPi, -SinF(theta) * rPi, sph * pyPi);
                            Pu = mk(twoPi * lhp.z, 0.0f, -twoPi * lhp.x);
                        } else {
                            tex = nullptr;
                        }
                        H->N = hn;
                        deferred = true;
                    } else {
                        H->N = normalToWorld(hn, E->InvTransformationMatrix);
                    }
                    H->u = (-phi + kPiF) / twoPi;
                    H->v = theta / kPiF;
                    H->E = E;
                    H->t = wt;
                    H->prim = C.triTotal + E->UniqueEntityID;
                    H->beta = 0.0f;
                    H->gamma = 0.0f;
                    result = true;
                }
            }
        }
    } break;
    case EntityType_Triangle: {                                                 // main.cpp:147-243
        V P0 = mk(E->P0), P1 = mk(E->P1), P2 = mk(E->P2);
        V fn = mk(E->FaceNormal);
        if (!cull || dot(ray.d, fn) < 0.0f) {
            f32 A = (P0.x - P1.x), B = (P0.x - P2.x), Cc = ray.d.x;
            f32 D = (P0.y - P1.y), Ee = (P0.y - P2.y), F = ray.d.y;
            f32 G = (P0.z - P1.z), Hh = (P0.z - P2.z), I = ray.d.z;
            f32 detM = (A * Ee * I) + (D * Hh * Cc) + (G * B * F) - (Cc * Ee * G) - (F * Hh * A) - (I * B * D);
            f32 K = (P0.x - ray.o.x), M = (P0.y - ray.o.y), N = (P0.z - ray.o.z);
            f32 detB = (K * Ee * I) + (M * Hh * Cc) + (N * B * F) - (Cc * Ee * N) - (F * Hh * K) - (I * B * M);
            const f32 lower = -0.000001f;
            f32 beta = detB / detM;
            if (beta > lower && beta <= 1.0f) {
                f32 detG = (A * M * I) + (D * N * Cc) + (G * K * F) - (Cc * M * G) - (F * N * A) - (I * K * D);
                f32 gamma = detG / detM;
                if (gamma > lower && gamma <= 1.0f) {
                    f32 detT = (A * Ee * N) + (D * Hh * K) + (G * B * M) - (K * Ee * G) - (M * Hh * A) - (N * B * D);
                    f32 alpha = 1.0f - beta - gamma;
                    if (alpha > lower) {
                        f32 t = detT / detM;
                        if (t > eps && t < H->t) {
                            H->P = ray.o + (ray.d * t);
                            if (E->ShadingMode == ShadingMode_Flat) H->N = fn;
                            else H->N = alpha * mk(E->N0) + beta * mk(E->N1) + gamma * mk(E->N2);
                            H->u = (alpha * E->UV0.x) + (beta * E->UV1.x) + (gamma * E->UV2.x);
                            H->v = (alpha * E->UV0.y) + (beta * E->UV1.y) + (gamma * E->UV2.y);
                            H->E = E;
                            H->t = t;
                            H->prim = meshIdx >= 0 ? C.primBase[meshIdx] + (u32)(E - C.triBase[meshIdx]) : AZ_MISS;
                            H->beta = beta;
                            H->gamma = gamma;
                            if (E->TextureIndex) {
                                tex = S->Textures + (E->TextureIndex - 1);
                                if (tex->BumpMap) {
                                    f32 u0 = E->UV0.x, v0 = E->UV0.y, u1 = E->UV1.x, v1 = E->UV1.y;
                                    f32 u2 = E->UV2.x, v2 = E->UV2.y;
                                    f32 div = ((u1 - u0) * (v2 - v0)) - ((v1 - v0) * (u2 - u0));
                                    div = 1.0f / div;
                                    Pu = (v2 - v0) * (P1 - P0) + (v0 - v1) * (P2 - P0);
                                    Pu = Pu * div;
                                    Pv = (u0 - u2) * (P1 - P0) + (u1 - u0) * (P2 - P0);
                                    Pv = Pv * div;
                                } else {
                                    tex = nullptr;
                                }
                            }
                            result = true;
                        }
                    }
                }
            }
        }
    } break;
    case EntityType_MeshInstance: {                                             // main.cpp:244-281
        if (rayAABB(E->AABB, ray)) {
            V motion = ray.T * mk(E->MotionVector);
            ray.o = ray.o - motion;
            Ray lr = getLocalRay(ray, E->InvTransformationMatrix);
            Hit lh = freshHit();
            s32 mi = (s32)(E->Mesh - S->Meshes);
            if (intersect(C, &lh, lr, E->Mesh->BVH, cull, eps, mi)) {
                V wp = pointToWorld(lh.P, E->TransformationMatrix);
                wp = wp + motion;
                f32 wt = length(wp - ray.o - motion) / length(ray.d);
                if (wt > eps && wt < H->t) {
                    *H = lh;
                    H->N = normalToWorld(lh.N, E->InvTransformationMatrix);
                    H->P = wp;
                    H->E = E;
                    H->t = wt;
                    result = true;
                } else {
                    result = false;
                }
            }
        }
    } break;
    case EntityType_BVHNode: {                                                  // main.cpp:282-319
        if (rayAABB(E->AABB, ray)) {
            Hit lh = freshHit();
            bool L = intersect(C, &lh, ray, E->Left, cull, eps, meshIdx);
            Hit rh = freshHit();
            bool R = intersect(C, &rh, ray, E->Right, cull, eps, meshIdx);
            if (L) {
                result = true;
                if (R) *H = (lh.t < rh.t) ? lh : rh;
                else *H = lh;
            } else if (R) {
                result = true;
                *H = rh;
            }
        }
    } break;
    }

    if (tex) {                                                                  // main.cpp:323-377
        if (tex->Type == TextureType_2D) {
            f32 u = H->u, v = H->v;
            V a = sampleTexture2D(tex, u, v, false);
            V b = sampleTexture2D(tex, u + tex->InvWidth, v, false);
            V c = sampleTexture2D(tex, u, v + tex->InvHeight, false);
            V du = b - a, dv = c - a;
            f32 Du = ((du.x + du.y + du.z) / 3.0f) * tex->BumpMapMultiplier;
            f32 Dv = ((dv.x + dv.y + dv.z) / 3.0f) * tex->BumpMapMultiplier;
            V dPu = Pu + Du * H->N;
            V dPv = Pv + Dv * H->N;
            H->N = normalize(cross(dPv, dPu));
        } else {
            const f32 e = 0.001f;
            V P = H->P;
            f32 D = samplePerlin(tex, P);
            f32 Dx = samplePerlin(tex, mk(P.x + e, P.y, P.z));
            f32 Dy = samplePerlin(tex, mk(P.x, P.y + e, P.z));
            f32 Dz = samplePerlin(tex, mk(P.x, P.y, P.z + e));
            V g = mk((Dx - D) / e, (Dy - D) / e, (Dz - D) / e);
            H->N = H->N - g * tex->BumpMapMultiplier;
            H->N = normalize(H->N);
        }
    }
    if (deferred) H->N = normalToWorld(H->N, E->InvTransformationMatrix);
    return result;
}

bool closest(const Ctx &C, Hit *H, const Ray &r, bool cull, f32 eps) {
    *H = freshHit();
    return intersect(C, H, r, C.S->BVHRoot, cull, eps, -1);
}

// ---- lights and BRDFs -------------------------------------------------------
struct LightSample { V dir; f32 invPdf; };

LightSample sampleEntityLight(const entity *LE, V hitP, VertexRng &rng) {        // main.cpp:392-459
    LightSample r = {mk(0, 0, 0), 0.0f};
    V hpl = pointToWorld(hitP, LE->InvTransformationMatrix);
    V dl = mk(0, 0, 0);
    if (LE->Type == EntityType_Sphere) {
        V D = mk(LE->Center) - hpl;
        f32 rod = LE->Radius / length(D);
        f32 cosMax = sqrtf(1.0f - sq(rod));
        f32 cosT = 1.0f + rng.next() * (cosMax - 1.0f);
        f32 sinT = sqrtf(1.0f - sq(cosT));
        f32 phi = 2.0f * kPiF * rng.next();
        f32 cp, sp;
        SinCosF(phi, sp, cp);
        V W = normalize(D), U, Vv;
        orthonormalBasis(W, &U, &Vv);
        dl = normalize(U * sinT * cp + Vv * sinT * sp + W * cosT);
        r.invPdf = 2.0f * kPiF * (1.0f - cosMax);
    } else {
        const mesh *M = LE->Mesh;
        u32 idx = (u32)(rng.next() * (M->TriangleCount));
        idx = idx < M->TriangleCount - 1 ? idx : M->TriangleCount - 1;
        const triangle *T = M->Triangles + idx;
        f32 r1 = sqrtf(rng.next());
        f32 r2 = rng.next();
        V Q = (1.0f - r2) * mk(T->P0) + r2 * mk(T->P1);
        V P = (1.0f - r1) * mk(T->P2) + r1 * Q;
        V wi = P - hpl;
        f32 A = M->TotalArea;
        dl = normalize(wi);
        f32 ca = dot(dl, mk(T->N));
        if (ca < 0.0f) ca *= -1.0f;
        r.invPdf = (A * ca) / sq(length(wi));
    }
    r.dir = vectorToWorld(dl, LE->TransformationMatrix);
    return r;
}

void evaluateBRDF(f32 *Dif, f32 *Spec, const brdf *B, V L, V N, V Vw, f32 cosI) {  // main.cpp:461-569
    switch (B->Type) {
    case BRDF_Phong: {
        V R = reflect(L, N);
        f32 c = dot(-Vw, R);
        *Spec = c > 0.0f ? PowF(c, B->Exponent) / cosI : 0.0f;
    } break;
    case BRDF_ModifiedPhong: {
        V R = reflect(L, N);
        f32 c = dot(-Vw, R);
        *Spec = c > 0.0f ? PowF(c, B->Exponent) : 0.0f;
    } break;
    case BRDF_NormalizedModifiedPhong: {
        *Dif = (f32)(1.0f / kPi);
        V R = reflect(L, N);
        f32 c = dot(-Vw, R);
        if (c > 0.0f) {
            f32 e = B->Exponent;
            *Spec = ((e + 2.0f) / (2.0f * kPiF)) * PowF(c, e);
        } else {
            *Spec = 0.0f;
        }
    } break;
    case BRDF_BlinnPhong: {
        V Hv = normalize(-Vw + L);
        f32 c = dot(N, Hv);
        *Spec = c > 0.0f ? PowF(c, B->Exponent) / cosI : 0.0f;
    } break;
    case BRDF_ModifiedBlinnPhong: {
        V Hv = normalize(-Vw + L);
        f32 c = dot(N, Hv);
        *Spec = c > 0.0f ? PowF(c, B->Exponent) : 0.0f;
    } break;
    case BRDF_NormalizedModifiedBlinnPhong: {
        *Dif = (f32)(1.0f / kPi);
        V Hv = normalize(-Vw + L);
        f32 c = dot(N, Hv);
        if (c > 0.0f) {
            f32 e = B->Exponent;
            *Spec = ((e + 8.0f) / (8.0f * kPiF)) * PowF(c, e);
        } else {
            *Spec = 0.0f;
        }
    } break;
    case BRDF_TorranceSparrow: {
        *Dif = (f32)(1.0f / kPi);
        V Hv = normalize(-Vw + L);
        f32 c = dot(N, Hv);
        if (c > 0.0f) {
            f32 e = B->Exponent;
            f32 ct = PowF(c, e);
            f32 D = ((e + 2.0f) / (2.0f * kPiF)) * ct;
            f32 nw0 = dot(N, -Vw);
            f32 w0wh = dot(-Vw, Hv);
            f32 G = minf_(nw0, cosI);
            G = (2.0f * G * c) / w0wh;
            G = minf_(1.0f, G);
            f32 ri = B->RefractiveIndex;
            f32 R0 = sq(ri - 1.0f) / sq(ri + 1.0f);
            f32 F = R0 + (1.0f - R0) * PowF(1.0f - w0wh, 5.0f);
            *Spec = (D * F * G) / (4.0f * cosI * nw0);
        } else {
            *Spec = 0.0f;
        }
    } break;
    }
}

void lightContribution(V *DC, V *SC, const brdf *B, V Li, V L, V Vw, V N, f32 dist, f32 falloff,
                       f32 invPdf) {                                            // main.cpp:571-603
    f32 d2 = sq(dist);
    V Ln = normalize(L);
    V inc = mk(Li.x / d2, Li.y / d2, Li.z / d2) * falloff;
    f32 cosI = dot(N, Ln);
    if (cosI > 0.0f) {
        f32 dif = 1.0f, spec = 1.0f;
        evaluateBRDF(&dif, &spec, B, Ln, N, Vw, cosI);
        V cm = inc * cosI * invPdf;
        *DC = *DC + cm * dif;
        *SC = *SC + cm * spec;
    }
}

struct Env {
    const Ctx &C;
    const epsilon_values *eps;
    bool cull;
    u64 seed;
    u32 pixel, sample;
};

void calculateLighting(const Env &X, V *DC, V *SC, const Ray &ray, const brdf *B, V hp, V hn,
                       f32 shadowEps, bool inside, VertexRng &rng) {           // main.cpp:605-779
    const scene *S = X.C.S;
    for (u32 li = 0; li < S->LightCount; ++li) {
        const light *Lt = S->Lights + li;
        V L = mk(0, 0, 0);
        V Li = mk(Lt->Intensity);
        bool discardAny = false;
        f32 falloff = 1.0f, ldist = 1.0f, invPdf = 1.0f;
        switch (Lt->Type) {
        case LightType_PointLight: {
            L = mk(Lt->Position) - hp;
            ldist = length(L);
        } break;
        case LightType_AreaLight: {
            f32 r0 = rng.next();
            f32 r1 = rng.next();
            V lp = mk(Lt->Position) + mk(Lt->EdgeVectors[0]) * r0 + mk(Lt->EdgeVectors[1]) * r1;
            falloff = dot(normalize(lp - hp), mk(Lt->Normal));
            if (falloff < 0.0f) falloff *= -1.0f;
            L = lp - hp;
            ldist = length(L);
        } break;
        case LightType_DirectionalLight: {
            L = normalize(-mk(Lt->Direction));
            discardAny = true;
        } break;
        case LightType_SpotLight: {
            L = mk(Lt->Position) - hp;
            ldist = length(L);
            f32 ch = Lt->CosHalfCoverage, cf = Lt->CosHalfFallOff;
            f32 ct = dot(normalize(-L), normalize(mk(Lt->SpotDirection)));
            if (ct < ch) falloff = 0.0f;
            else if (ct > cf) falloff = 1.0f;
            else {
                falloff = (ct - ch) / (cf - ch);
                falloff = PowF(falloff, 4.0f);
                if (falloff < 0.0f) falloff = 0.0f;
                else if (falloff > 1.0f) falloff = 1.0f;
            }
        } break;
        case LightType_EntityLight: {
            LightSample s = sampleEntityLight(&Lt->Entity, hp, rng);
            L = s.dir;
            invPdf = s.invPdf;
            Li = mk(Lt->Entity.Radiance);
        } break;
        }
        if (ldist == 0.0f) ldist = 0.001f;
        V lo = hp + (L * shadowEps);
        bool reaches = true;
        Hit lh;
        Ray sr = {lo, L, ray.T};
        if (closest(X.C, &lh, sr, false, 0.0f)) {
            if (Lt->Type == LightType_EntityLight) {
                if (lh.E->UniqueEntityID != Lt->Entity.UniqueEntityID) reaches = false;
            } else if (discardAny || lh.t < (1.0f - shadowEps)) {
                reaches = false;
            }
        }
        if (reaches && !inside) lightContribution(DC, SC, B, Li, L, ray.d, hn, ldist, falloff, invPdf);
    }
    if (S->UseEnvironmentMap) {
        f32 phi = 2.0f * kPiF * rng.next();
        f32 cp, sp;
        SinCosF(phi, sp, cp);
        f32 theta = Asin(sqrtf(rng.next()));
        f32 ct, st;
        SinCosF(theta, st, ct);
        V U, Vv;
        orthonormalBasis(hn, &U, &Vv);
        V L = normalize(U * st * cp + Vv * st * sp + hn * ct);
        Ray er = {hp + (L * shadowEps), L, ray.T};
        Hit eh;
        if (!closest(X.C, &eh, er, false, 0.0f)) {
            f32 u = (-Atan2F(L.z, L.x) + kPiF) / (2.0f * kPiF);
            f32 v = Acos(L.y) / kPiF;
            V I = sampleTexture(&S->EnvironmentMap, mk(u, v, 1.0f), u, v);
            f32 ca = dot(L, hn);
            if (ca == 0.0f) ca = 0.000001f;
            f32 invPdf = kPiF / ca;
            lightContribution(DC, SC, B, I, L, ray.d, hn, 1.0f, 1.0f, invPdf);
        }
    }
}

V traceRay(const Env &X, Ray ray, u32 depth, bool inside, f32 *travel, u32 node);

V mirrorContribution(const Env &X, u32 depth, const material *Mt, const Ray &ray, V hp, V hn,
                     bool inside, f32 *travel, u32 node, VertexRng &rng) {     // main.cpp:781-854
    V R = reflect(-ray.d, hn);
    f32 gl = Mt->Glossiness;
    if (gl != 0.0f) {
        V rp = mk(fabsf(R.x), fabsf(R.y), fabsf(R.z));
        if (rp.x < rp.y) {
            if (rp.x < rp.z) { rp = R; rp.x = rp.x < 0.0f ? -1.0f : 1.0f; }
            else { rp = R; rp.z = rp.z < 0.0f ? -1.0f : 1.0f; }
        } else {
            if (rp.y < rp.z) { rp = R; rp.y = rp.y < 0.0f ? -1.0f : 1.0f; }
            else { rp = R; rp.z = rp.z < 0.0f ? -1.0f : 1.0f; }
        }
        rp = normalize(rp);
        V u = cross(rp, R);
        V v = cross(u, R);
        f32 x1 = rng.next();
        f32 x2 = rng.next();
        R = R + 2.0f * gl * (u * (x1 - 0.5f) + v * (x2 - 0.5f));
        R = normalize(R);
    }
    const f32 mirrorEps = 0.000825f;
    Ray rr = {hp + (R * mirrorEps), R, ray.T};
    return had(mk(Mt->MirrorReflectance), traceRay(X, rr, depth + 1, inside, travel, childNode(node, 1)));
}

V transparencyContribution(const Env &X, const material *Mt, const Ray &ray, V hp, V hn, u32 depth,
                           bool inside, u32 node) {                            // main.cpp:856-964
    V result = mk(0, 0, 0);
    V tr = mk(Mt->Transparency);
    f32 to = Mt->RefractionIndex, from = 1.0f;
    if (inside) {
        from = Mt->RefractionIndex;
        to = 1.0f;
    }
    V N = hn;
    f32 cosT = -dot(ray.d, N);
    if (cosT < 0.0f) {
        N = -hn;
        cosT = -dot(ray.d, N);
    }
    f32 ratio = from / to;
    f32 delta = 1.0f - sq(ratio) * (1.0f - sq(cosT));
    f32 refl;
    if (delta <= 0.0f) {
        refl = 1.0f;
    } else {
        f32 cosPhi = sqrtf(delta);
        f32 cosAir = inside ? cosPhi : cosT;
        f32 inv = 1.0f - cosAir;
        f32 r0 = sq((from - to) / (from + to));
        refl = r0 + (1.0f - r0) * PowF(inv, 5);
        V T = (ray.d + N * cosT) * ratio - N * cosPhi;
        T = normalize(T);
        const f32 refrEps = 0.000100f;
        Ray rr = {hp + (T * refrEps), T, ray.T};
        f32 td = 0;
        V rc = traceRay(X, rr, depth + 1, !inside, &td, childNode(node, 2));
        V att;
        if (!inside) att = mk(Exp((tr.x - 1.0f) * td), Exp((tr.y - 1.0f) * td), Exp((tr.z - 1.0f) * td));
        else att = mk(1.0f, 1.0f, 1.0f);
        result = result + (1.0f - refl) * had(att, rc);
    }
    const f32 reflEps = 0.000100f;
    V R = reflect(-ray.d, N);
    Ray rr = {hp + (R * reflEps), R, ray.T};
    f32 td = 0;
    V rc = traceRay(X, rr, depth + 1, inside, &td, childNode(node, 3));
    V att = mk(1.0f, 1.0f, 1.0f);
    if (inside) att = mk(Exp((tr.x - 1.0f) * td), Exp((tr.y - 1.0f) * td), Exp((tr.z - 1.0f) * td));
    result = result + refl * had(att, rc);
    return result;
}

V traceRay(const Env &X, Ray ray, u32 depth, bool inside, f32 *travel, u32 node) {  // main.cpp:966-1133
    const scene *S = X.C.S;
    V result = mk(0, 0, 0);
    if (depth <= S->MaxRecursionDepth) {
        Hit h;
        bool hit = closest(X.C, &h, ray, X.cull, X.eps->Intersection);
        if (hit) {
            *travel += h.t;
            const entity *E = h.E;
            if (E->IsEmitter) {
                result = mk(E->Radiance);
            } else {
                VertexRng rng(X.seed, X.pixel, X.sample, node);
                const material *Mt = S->Materials + E->MaterialIndex;
                const brdf *B = &Mt->BRDF;
                V dcol = mk(Mt->DiffuseReflectance), scol = mk(Mt->SpecularReflectance);
                V DC = mk(0, 0, 0), SC = mk(0, 0, 0);
                V hp = h.P, hn = h.N;
                bool ignoreAmbient = false;
                if (E->TextureIndex) {
                    const texture *T = S->Textures + (E->TextureIndex - 1);
                    V ts = sampleTexture(T, hp, h.u, h.v);
                    switch (T->BlendMode) {
                    case TextureBlendMode_ReplaceAll:
                        dcol = ts;
                        DC = mk(1.0f, 1.0f, 1.0f);
                        SC = mk(0.0f, 0.0f, 0.0f);
                        ignoreAmbient = true;
                        break;
                    case TextureBlendMode_ReplaceDiffuse:
                        dcol = ts;
                        break;
                    case TextureBlendMode_BlendDiffuse:
                        dcol = dcol + ts;
                        dcol = dcol * 0.5f;
                        break;
                    }
                }
                calculateLighting(X, &DC, &SC, ray, B, hp, hn, X.eps->ShadowRay, inside, rng);
                if (S->PathTrace) {
                    f32 theta = S->ImportanceSampling ? Asin(sqrtf(rng.next())) : Acos(rng.next());
                    f32 ct, st;
                    SinCosF(theta, st, ct);
                    f32 phi = 2.0f * kPiF * rng.next();
                    f32 cp, sp;
                    SinCosF(phi, sp, cp);
                    V U, Vv;
                    orthonormalBasis(hn, &U, &Vv);
                    V s = normalize(U * st * cp + Vv * st * sp + hn * ct);
                    Ray sr = {hp + (s * X.eps->ShadowRay), s, ray.T};
                    f32 stv = 0.0f;
                    V sc = traceRay(X, sr, depth + 1, inside, &stv, childNode(node, 0));
                    f32 invPdf;
                    if (S->ImportanceSampling) {
                        f32 cs = dot(s, hn);
                        if (cs == 0.0f) cs = 0.000001f;
                        invPdf = kPiF / cs;
                    } else {
                        invPdf = kPiF * 2.0f;
                    }
                    lightContribution(&DC, &SC, B, sc, sr.d, ray.d, hn, 1.0f, 1.0f, invPdf);
                }
                result = result + had(dcol, DC);
                result = result + had(scol, SC);
                bool transparent = bitsNonZero(Mt->Transparency);
                if ((!inside || !transparent) && !ignoreAmbient)
                    result = result + had(mk(Mt->AmbientReflectance), mk(S->AmbientLight));
                if (bitsNonZero(Mt->MirrorReflectance))
                    result = result + mirrorContribution(X, depth, Mt, ray, hp, hn, inside, travel, node, rng);
                if (transparent) result = result + transparencyContribution(X, Mt, ray, hp, hn, depth, inside, node);
            }
        }
    }
    return result;
}

// ---- camera and image loop (main.cpp:1135-1397) ----------------------------
struct CameraSetup {
    V base, pos, ax, ay, stepX, stepY;
    f32 aperture;
    u32 W, H, sppSqrt;
    f32 invSppSqrt;
    s32 kernel;
    f32 oneOverTwoPiSigma2, negOneOverTwoSigma2;
};

CameraSetup setupCamera(const camera *Cam, u32 sppOverride) {                    // main.cpp:1265-1347
    CameraSetup c;
    V gaze = mk(Cam->Gaze);
    if (Cam->GazeIsPoint) gaze = mk(Cam->Gaze) - mk(Cam->Position);
    V gn = normalize(gaze);
    V az = -gn;
    V ax, ay;
    if (Cam->IsLeftHanded) {
        ax = normalize(cross(normalize(mk(Cam->Up)), gn));
        ay = cross(ax, az);
    } else {
        ax = normalize(cross(gn, normalize(mk(Cam->Up))));
        ay = cross(az, ax);
    }
    f32 fdnd = 1.0f;
    if (Cam->FocusDistance != 0.0f) fdnd = Cam->FocusDistance / Cam->NearDistance;
    c.aperture = Cam->ApertureSize;
    c.W = Cam->Resolution.x;
    c.H = Cam->Resolution.y;
    f32 r = Cam->NearPlane.x, l = Cam->NearPlane.y, b = Cam->NearPlane.z, t = Cam->NearPlane.w;
    if (Cam->Type == CameraType_Simple) {
        f32 hh = Cam->NearDistance * tanf((Cam->FovY / 2.0f / 180.0f) * kPiF);
        f32 aspect = (f32)c.W / (f32)c.H;
        f32 wh = aspect * hh;
        t = +hh;
        b = -hh;
        r = +wh;
        l = -wh;
    }
    f32 psx = (r - l) / (f32)c.W, psy = (t - b) / (f32)c.H;
    c.stepX = ax * psx * fdnd;
    c.stepY = ay * psy * fdnd;
    V topLeft = mk(Cam->Position) + Cam->NearDistance * (gn) + l * ax + t * ay;
    c.base = (topLeft - mk(Cam->Position)) * fdnd;
    c.pos = mk(Cam->Position);
    c.ax = ax;
    c.ay = ay;
    u32 spp = sppOverride ? sppOverride : Cam->SampleCount;
    c.sppSqrt = (u32)sqrtf((f32)spp);
    c.invSppSqrt = 1.0f / c.sppSqrt;
    f32 radius = 1.5f;
    f32 sigma = radius / 3.0f;
    f32 twoS2 = 2.0f * sq(sigma);
    c.negOneOverTwoSigma2 = -1.0f / twoS2;
    c.oneOverTwoPiSigma2 = 1.0f / (twoS2 * (f32)kPi);
    c.kernel = (s32)ceilf(radius - 0.5f);
    return c;
}

struct PixelF64 { f64 c[4]; };

inline void atomicAddF32(f32 *p, f32 v) { platform().AtomicAddF32(p, v); }   // platform_win32.cpp:377-395
inline void atomicAddF64(f64 *p, f64 v) {
    std::atomic_ref<f64> a(*p);
    f64 cur = a.load();
    while (!a.compare_exchange_weak(cur, cur + v)) {
    }
}

void splat(pixel *P, u32 idx, V c, f32 w) {
    atomicAddF32(&P[idx].Color.x, c.x);
    atomicAddF32(&P[idx].Color.y, c.y);
    atomicAddF32(&P[idx].Color.z, c.z);
    atomicAddF32(&P[idx].TotalWeight, w);
}
void splat(PixelF64 *P, u32 idx, V c, f32 w) {
    atomicAddF64(&P[idx].c[0], (f64)c.x);
    atomicAddF64(&P[idx].c[1], (f64)c.y);
    atomicAddF64(&P[idx].c[2], (f64)c.z);
    atomicAddF64(&P[idx].c[3], (f64)w);
}

template <typename PixelT>
void produceRows(const Ctx &C, const camera *Cam, const epsilon_values *eps, filter_type filter,
                 bool cull, const az_render_opts &o, PixelT *px, u32 row0, u32 row1) {  // main.cpp:1256-1397
    CameraSetup cs = setupCamera(Cam, o.SampleCountOverride);
    Env X{C, eps, cull, o.Seed, 0, 0};
    s32 wm1 = (s32)cs.W - 1, hm1 = (s32)cs.H - 1;
    const scene *S = C.S;
    for (u32 row = row0; row < row1; ++row) {
        for (u32 col = 0; col < cs.W; ++col) {
            u32 pidx = row * cs.W + col;
            for (u32 sy = 0; sy < cs.sppSqrt; ++sy) {
                for (u32 sx = 0; sx < cs.sppSqrt; ++sx) {
                    f32 offx = 0.5f * cs.invSppSqrt + sx * cs.invSppSqrt;   // main.cpp:1135-1153
                    f32 offy = 0.5f * cs.invSppSqrt + sy * cs.invSppSqrt;
                    f32 scx = col + offx, scy = row + offy;
                    u32 sample = sy * cs.sppSqrt + sx;
                    X.pixel = pidx;
                    X.sample = sample;
                    VertexRng camRng(o.Seed, pidx, sample, 0);
                    f32 lx = camRng.next() - 0.5f;                              // main.cpp:1155-1178
                    f32 ly = camRng.next() - 0.5f;
                    lx *= cs.aperture;
                    ly *= cs.aperture;
                    V lox = cs.ax * lx, loy = cs.ay * ly;
                    V dir = normalize(cs.base + scx * cs.stepX - scy * cs.stepY - lox - loy);
                    Ray ray = {cs.pos + lox + loy, dir, camRng.next()};
                    f32 travel = 0.0f;
                    V col3 = traceRay(X, ray, 0, false, &travel, childNode(0, 0));
                    if (travel == 0.0f) {                                        // main.cpp:1180-1207
                        if (S->UseEnvironmentMap) {
                            f32 th = Acos(ray.d.y);
                            f32 ph = Atan2F(ray.d.z, ray.d.x);
                            f32 u = (-ph + kPiF) / (2.0f * kPiF), v = th / kPiF;
                            col3 = sampleTexture(&S->EnvironmentMap, mk(u, v, 1.0f), u, v);
                        } else if (S->UseBackgroundTexture) {
                            f32 u = (f32)col / (f32)wm1, v = (f32)row / (f32)hm1;
                            col3 = sampleTexture(&S->BackgroundTexture, mk(u, v, 1.0f), u, v);
                        } else {
                            col3 = mk(S->BackgroundColor);
                        }
                    }
                    if (filter == FilterType_Box) {                              // main.cpp:1209-1254
                        splat(px, pidx, col3, 1.0f);
                    } else {
                        s32 gx0 = (s32)col - cs.kernel > 0 ? (s32)col - cs.kernel : 0;
                        s32 gx1 = (s32)col + cs.kernel < wm1 ? (s32)col + cs.kernel : wm1;
                        s32 gy0 = (s32)row - cs.kernel > 0 ? (s32)row - cs.kernel : 0;
                        s32 gy1 = (s32)row + cs.kernel < hm1 ? (s32)row + cs.kernel : hm1;
                        for (s32 gy = gy0; gy <= gy1; ++gy)
                            for (s32 gx = gx0; gx <= gx1; ++gx) {
                                f32 dx = scx - (gx + 0.5f), dy = scy - (gy + 0.5f);
                                f32 d2 = sq(dx) + sq(dy);
                                f32 g = cs.oneOverTwoPiSigma2 * Exp(cs.negOneOverTwoSigma2 * d2);
                                splat(px, (u32)(gy * (s32)cs.W + gx), col3 * g, g);
                            }
                    }
                }
            }
        }
    }
}

// One platform thread per job, then a wait for all of them and the handles
// closed, as RenderScene does with its bands (main.cpp:1823-1880).
PlatformThreadProc(runJob) {
    (*static_cast<std::function<void()> *>(Parameters))();
    return 0;
}

void runThreads(std::vector<std::function<void()>> &jobs) {
    const az_platform_api &P = platform();
    std::vector<thread_handle> h(jobs.size());
    u32 started = 0;
    for (size_t i = 0; i < jobs.size(); ++i) {
        if (P.CreateThread(&h[started], runJob, &jobs[i], nullptr, 0)) {
            ++started;
        } else {
            jobs[i]();   // no thread to be had: run the job here
        }
    }
    if (started) P.WaitForThreads(started, h.data(), 1, 1, 0);
    for (u32 i = 0; i < started; ++i) P.CloseThreadHandle(h[i]);
}

template <typename PixelT>
void renderBands(const scene *S, const camera *Cam, const epsilon_values *eps, filter_type filter,
                 b32 cull, const az_render_opts *opts, PixelT *px, u32 threads) {  // main.cpp:1799-1880
    az_render_opts o;
    memset(&o, 0, sizeof(o));
    if (opts) o = *opts;
    Ctx C(S);
    u32 H = Cam->Resolution.y;
    u32 r0 = o.RowBegin, r1 = o.RowEnd ? o.RowEnd : H;
    if (r1 > H) r1 = H;
    if (r0 >= r1) return;
    if (threads == 0) threads = 1;
    u32 rows = r1 - r0;
    if (threads > rows) threads = rows;
    u32 minPer = rows / threads, resid = rows - minPer * threads, consumed = r0;
    std::vector<std::function<void()>> bands;
    for (u32 t = 0; t < threads; ++t) {
        u32 n = minPer;
        if (resid) {
            ++n;
            --resid;
        }
        u32 a = consumed, b = consumed + n;
        bands.emplace_back([&, a, b] { produceRows(C, Cam, eps, filter, cull != 0, o, px, a, b); });
        consumed += n;
    }
    runThreads(bands);
}

template <typename F>
void parallelFor(u32 n, u32 threads, F f) {
    if (threads == 0) threads = 1;
    std::vector<std::function<void()>> parts;
    u32 chunk = (n + threads - 1) / threads;
    for (u32 t = 0; t < threads; ++t) {
        u32 a = t * chunk, b = a + chunk < n ? a + chunk : n;
        if (a >= b) break;
        parts.emplace_back([=] { for (u32 i = a; i < b; ++i) f(i); });
    }
    runThreads(parts);
}

inline Ray toRay(const az_ray &r) {
    return Ray{mk(r.Origin[0], r.Origin[1], r.Origin[2]), mk(r.Direction[0], r.Direction[1], r.Direction[2]), r.Time};
}
inline az_hit toHit(const Hit &h, bool hit) {
    az_hit o;
    if (!hit) {
        o.Prim = AZ_MISS;
        o.t = 0.0f;
        o.Beta = 0.0f;
        o.Gamma = 0.0f;
        return o;
    }
    o.Prim = h.prim;
    o.t = h.t;
    o.Beta = h.beta;
    o.Gamma = h.gamma;
    return o;
}

// Brute force: the same per-primitive tests over every primitive, with the
// instance transform arithmetic of main.cpp:244-281.  Leaves are visited left
// to right and a later leaf wins an exact tie (`<=`), which is what the
// reference's recursive minimum with ties to the right child yields.
bool bruteClosest(const Ctx &C, Hit *H, const Ray &ray, bool cull, f32 eps) {
    *H = freshHit();
    bool any = false;
    const scene *S = C.S;
    std::vector<const entity *> stack;
    stack.push_back(S->BVHRoot);
    std::vector<const entity *> tops;
    while (!stack.empty()) {
        const entity *e = stack.back();
        stack.pop_back();
        if (e->Type == EntityType_BVHNode) {
            stack.push_back(e->Right);
            stack.push_back(e->Left);
        } else {
            tops.push_back(e);
        }
    }
    for (const entity *E : tops) {
        if (E->Type == EntityType_Sphere) {
            Hit h = freshHit();
            if (intersect(C, &h, ray, E, cull, eps, -1) && h.t <= H->t) {
                *H = h;
                any = true;
            }
        } else if (E->Type == EntityType_MeshInstance) {
            Ray r = ray;
            V motion = r.T * mk(E->MotionVector);
            r.o = r.o - motion;
            Ray lr = getLocalRay(r, E->InvTransformationMatrix);
            s32 mi = (s32)(E->Mesh - S->Meshes);
            Hit best = freshHit();
            bool found = false;
            const entity *base = C.triBase[mi];
            for (u32 k = 0; k < E->Mesh->TriangleCount; ++k) {
                Hit h = freshHit();
                if (intersect(C, &h, lr, base + k, cull, eps, mi) && h.t <= best.t) {
                    best = h;
                    found = true;
                }
            }
            if (found) {
                V wp = pointToWorld(best.P, E->TransformationMatrix);
                wp = wp + motion;
                f32 wt = length(wp - r.o - motion) / length(r.d);
                if (wt > eps && wt <= H->t) {
                    *H = best;
                    H->N = normalToWorld(best.N, E->InvTransformationMatrix);
                    H->P = wp;
                    H->E = E;
                    H->t = wt;
                    any = true;
                }
            }
        }
    }
    return any;
}

}  // namespace

extern "C" {

void azo_trace_closest(const scene *S, const az_ray *rays, u32 n, az_hit *hits, b32 cull, f32 eps,
                       u32 threads) {
    Ctx C(S);
    parallelFor(n, threads, [&](u32 i) {
        Hit h;
        bool hit = closest(C, &h, toRay(rays[i]), cull != 0, eps);
        hits[i] = toHit(h, hit);
    });
}

void azo_trace_any(const scene *S, const az_ray *rays, u32 n, az_hit *hits, u32 threads) {
    Ctx C(S);
    parallelFor(n, threads, [&](u32 i) {
        Hit h;
        bool hit = closest(C, &h, toRay(rays[i]), false, 0.0f);
        hits[i] = toHit(h, hit && h.t < rays[i].TMax);
    });
}

void azo_brute_closest(const scene *S, const az_ray *rays, u32 n, az_hit *hits, b32 cull, f32 eps,
                       u32 threads) {
    Ctx C(S);
    parallelFor(n, threads, [&](u32 i) {
        Hit h;
        bool hit = bruteClosest(C, &h, toRay(rays[i]), cull != 0, eps);
        hits[i] = toHit(h, hit);
    });
}

void azo_render(const scene *S, const camera *Cam, const epsilon_values *eps, filter_type filter,
                b32 cull, const az_render_opts *opts, pixel *px, u32 threads) {
    renderBands(S, Cam, eps, filter, cull, opts, px, threads);
}

void azo_render_f64(const scene *S, const camera *Cam, const epsilon_values *eps, filter_type filter,
                    b32 cull, const az_render_opts *opts, f64 *px, u32 threads) {
    renderBands(S, Cam, eps, filter, cull, opts, reinterpret_cast<PixelF64 *>(px), threads);
}

f32 azo_platform_atomic_sum(u32 threads, u32 addsPerThread, f32 value) {
    f32 total = 0.0f;
    std::vector<std::function<void()>> jobs;
    for (u32 t = 0; t < threads; ++t)
        jobs.emplace_back([&] {
            for (u32 k = 0; k < addsPerThread; ++k) platform().AtomicAddF32(&total, value);
        });
    runThreads(jobs);
    return total;
}

// The float elementary functions as the renderer calls them: 0 Sin, 1 Cos,
// 2 Atan2(x = y-coordinate, y = x-coordinate), 3 Power, 4 Exp, 5 ArcSine,
// 6 ArcCosine, 7 x^4, 8 x^5 (math.cpp:136-260).
int azo_math_eval(u32 fn, const f32 *x, const f32 *y, u32 n, f32 *out) {
    for (u32 i = 0; i < n; ++i) {
        const f32 a = x[i], b = y ? y[i] : 0.0f;
        f32 r;
        switch (fn) {
        case 0: r = SinF(a); break;
        case 1: r = CosF(a); break;
        case 2: r = Atan2F(a, b); break;
        case 3: r = PowF(a, b); break;
        case 4: r = Exp(a); break;
        case 5: r = Asin(a); break;
        case 6: r = Acos(a); break;
        case 7: r = PowF(a, 4.0f); break;
        case 8: r = PowF(a, 5.0f); break;
        default: return 0;
        }
        out[i] = r;
    }
    return 1;
}

f32 azo_rng_draw(u64 seed, u32 pixel, u32 sample, u32 node, u32 dim) {
    VertexRng r(seed, pixel, sample, node);
    f32 v = 0.0f;
    for (u32 i = 0; i <= dim; ++i) v = r.next();
    return v;
}

u32 azo_pcg32_at(u64 initstate, u64 initseq, u64 delta) {
    Pcg p;
    p.seed(initstate, initseq);
    p.advance(delta);
    return p.next();
}

void azo_perlin_gradient(u32 i, f32 *out) {
    V g = perlinTable().g[i & (AZ_PERLIN_GRADIENT_SIZE - 1)];
    out[0] = g.x;
    out[1] = g.y;
    out[2] = g.z;
}

f32 azo_perlin_sample(const texture *T, f32 x, f32 y, f32 z) { return samplePerlin(T, mk(x, y, z)); }

void azo_evaluate_brdf(const brdf *B, const f32 *L, const f32 *N, const f32 *Vw, f32 cosI, f32 *dif,
                       f32 *spec) {
    *dif = 1.0f;
    *spec = 1.0f;
    evaluateBRDF(dif, spec, B, mk(L[0], L[1], L[2]), mk(N[0], N[1], N[2]), mk(Vw[0], Vw[1], Vw[2]), cosI);
}

}  // extern "C"

// ---- output stage -----------------------------------------------------------------------
// OutputPixelsToFile (main.cpp:1620-1797) for the 8-bit formats, then the
// byte conversion of OutputPixelsToPNG/BMP/TGA/JPG (main.cpp:1461-1477).

// Clamp(Val, 0, 1) = MaxOf(0, MinOf(Val, 1)) (math.cpp:11-26, math.h:6-7):
// NaN compares false in MinOf and comes out as 1.
static f32 clamp01Ref(f32 v) {
    const f32 m = v < 1.0f ? v : 1.0f;
    return 0.0f > m ? 0.0f : m;
}

// Luminance (math.cpp:1866-1879, the enabled #if branch).
static f32 luminanceRef(f32 r, f32 g, f32 b) { return r * 0.27f + g * 0.67f + b * 0.06f; }

// CompareF32 (util.h:71-81): the sign of the float difference.
static int compareF32Ref(const void *a, const void *b) {
    const f32 d = *(const f32 *)a - *(const f32 *)b;
    return d > 0.0f ? 1 : (d < 0.0f ? -1 : 0);
}

extern "C" int azo_output_pixels(const pixel *Pixels, u32 Width, u32 Height, const tonemap *Tonemap, f32 InvGamma,
                                 b32 SRGB, u8 *Out) {
    const u32 n = Width * Height;
    if (n == 0) return 0;
    const bool tm = Tonemap && Tonemap->Operator != TonemapOperator_None;
    std::vector<f32> col((size_t)n * 3), lum(tm ? n : 0);
    // Filtering: colour over weight, luminance and the sequential f32 sum
    // of log(L + 1e-5) (main.cpp:1640-1661; LogBaseE = double log,
    // math.cpp:172-179).
    f32 logAvg = 0.0f;
    for (u32 i = 0; i < n; ++i) {
        const f32 w = Pixels[i].TotalWeight;
        const f32 r = Pixels[i].Color.x / w, g = Pixels[i].Color.y / w, b = Pixels[i].Color.z / w;
        col[3 * i + 0] = r;
        col[3 * i + 1] = g;
        col[3 * i + 2] = b;
        if (tm) {
            const f32 L = luminanceRef(r, g, b);
            lum[i] = L;
            logAvg += (f32)log((double)(L + 0.00001f));
        }
    }
    // Photographic operator (main.cpp:1686-1731; Exp = double exp,
    // math.cpp:145-152; RoundAsU32 = (u32)(x + 0.5f), math.cpp:100-107).
    if (tm) {
        logAvg = (f32)exp((double)(logAvg / (f32)n));
        const f32 factor = Tonemap->KeyValue / logAvg;
        const f32 at = (1.0f - Tonemap->BurnoutPercentage * 0.01f) * (n - 1) + 0.5f;
        if (!(at >= 0.0f) || !(at < (f32)n)) return -1;   // the reference would index outside the array
        const u32 li = (u32)at;
        qsort(lum.data(), n, sizeof(f32), compareF32Ref);
        f32 lw = lum[li] * factor;
        lw = lw * lw;
        const f32 sat = Tonemap->Saturation;
        for (u32 i = 0; i < n; ++i) {
            const f32 r = col[3 * i + 0], g = col[3 * i + 1], b = col[3 * i + 2];
            const f32 L = luminanceRef(r, g, b);
            if (L != 0.0f) {
                const f32 Ls = L * factor;
                const f32 Ld = (Ls * (1.0f + (Ls / lw))) / (1.0f + Ls);
                col[3 * i + 0] = clamp01Ref(powf(r / L, sat) * Ld);
                col[3 * i + 1] = clamp01Ref(powf(g / L, sat) * Ld);
                col[3 * i + 2] = clamp01Ref(powf(b / L, sat) * Ld);
            }
        }
    }
    // Gamma (main.cpp:1734-1778) and bytes (main.cpp:1468-1474).
    for (size_t k = 0; k < (size_t)n * 3; ++k) {
        f32 c = col[k];
        if (SRGB) c = c > 0.0031308f ? 1.055f * powf(c, InvGamma) - 0.055f : c * 12.92f;
        else c = powf(c, InvGamma);
        Out[k] = (u8)(clamp01Ref(c) * 255.0f);
    }
    return 0;
}
