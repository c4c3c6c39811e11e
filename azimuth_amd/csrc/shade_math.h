// shade_math.h -- device restatement of the reference shading helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_render.h"
#include "dmath.h"

namespace az {

constexpr double kPiD = 3.14159265359;            // math.h:3
constexpr float kPiF = (float)kPiD;               // math.h:4

struct V3 {
    float x, y, z;
};
__device__ inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ inline V3 v3(const float *p) { return V3{p[0], p[1], p[2]}; }
__device__ inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ inline V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ inline V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__device__ inline V3 had(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline float length(V3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }   // math.cpp:1289-1296
__device__ inline V3 normalize(V3 a) {                                                      // math.cpp:1298-1313
    float l = length(a);
    if (l != 0.0f) return v3(a.x / l, a.y / l, a.z / l);
    return v3(0.0f, 0.0f, 0.0f);
}
__device__ inline V3 reflect(V3 v, V3 r) { return ((2.0f * dot(v, r)) * r) - v; }           // math.cpp:1335-1348
__device__ inline float sq(float v) { return v * v; }

__device__ inline V3 point_to_world(V3 v, const Affine &m) {                               // math.cpp:1774-1800
    return v3(m.m[0][0] * v.x + m.m[0][1] * v.y + m.m[0][2] * v.z + m.m[0][3],
              m.m[1][0] * v.x + m.m[1][1] * v.y + m.m[1][2] * v.z + m.m[1][3],
              m.m[2][0] * v.x + m.m[2][1] * v.y + m.m[2][2] * v.z + m.m[2][3]);
}
__device__ inline V3 vector_to_world(V3 v, const Affine &m) {                              // math.cpp:1802-1825
    return v3(m.m[0][0] * v.x + m.m[0][1] * v.y + m.m[0][2] * v.z,
              m.m[1][0] * v.x + m.m[1][1] * v.y + m.m[1][2] * v.z,
              m.m[2][0] * v.x + m.m[2][1] * v.y + m.m[2][2] * v.z);
}
__device__ inline V3 normal_to_world(V3 v, const Affine &inv) {                            // math.cpp:1827-1843
    V3 r = v3(inv.m[0][0] * v.x + inv.m[1][0] * v.y + inv.m[2][0] * v.z,
              inv.m[0][1] * v.x + inv.m[1][1] * v.y + inv.m[2][1] * v.z,
              inv.m[0][2] * v.x + inv.m[1][2] * v.y + inv.m[2][2] * v.z);
    return normalize(r);
}
__device__ inline void orthonormal_basis(V3 w, V3 &u, V3 &v) {                            // math.cpp:1845-1863
    if (w.z < 0.0) {
        const float a = 1.0f / (1.0f - w.z);
        const float b = w.x * w.y * a;
        u = v3(1.0f - w.x * w.x * a, -b, w.x);
        v = v3(b, w.y * w.y * a - 1.0f, -w.y);
    } else {
        const float a = 1.0f / (1.0f + w.z);
        const float b = -w.x * w.y * a;
        u = v3(1.0f - w.x * w.x * a, b, -w.x);
        v = v3(b, 1.0f - w.y * w.y * a, -w.y);
    }
}

// ---- counter-addressed PCG32 ---------------------------------------------
// State after seeding with (seed, pixel) and jumping to (sample, node); the
// jump composes precomputed 2^k-step maps (mult^(2^k), sum of mult^i).
constexpr uint64_t kPcgMult = 6364136223846793005ULL;

struct Rng {
    uint64_t state, inc;
    __device__ inline uint32_t next() {
        uint64_t old = state;
        state = old * kPcgMult + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    __device__ inline float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

__device__ inline Rng make_rng(const uint64_t *__restrict__ table, uint64_t seed, uint32_t pixel, uint32_t sample,
                               uint32_t node) {
    Rng r;
    r.inc = ((uint64_t)pixel << 1u) | 1u;
    r.state = 0u;
    r.next();
    r.state += seed;
    r.next();
    const uint64_t delta = ((uint64_t)sample << 38) + ((uint64_t)node << 6);
    uint64_t accMult = 1u, accPlus = 0u;
    // One table entry per non-zero byte of the distance (entry 256 k + v
    // advances by v << 8k steps).
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t v = (uint32_t)(delta >> (8 * k)) & 255u;
        if (v == 0u) continue;
        const uint64_t m = table[2 * (256 * k + v)], g = table[2 * (256 * k + v) + 1];
        accMult *= m;
        accPlus = accPlus * m + r.inc * g;
    }
    r.state = accMult * r.state + accPlus;
    return r;
}

__device__ inline uint32_t child_node(uint32_t n, uint32_t kind) { return n * 5u + kind + 1u; }

// ---- textures (texture.h:134-357) ------------------------------------------
__device__ inline uint32_t mod256(int32_t v) { return (uint32_t)(((v % 256) + 256) % 256); }

__device__ inline float perlin_weight(float v) {
    v = fabsf(v);
    return 1.0f - ((((6.0f * v) - 15.0f) * v + 10.0f) * v * v * v);
}

// The lattice gradients of the 8 corners of cell (I, J, K), corner
// (x, y, z) at 4 z + 2 y + x: PerlinHash's permutation chain (texture.h:134-
// 160, with a proper mod 256).
__device__ inline void perlin_cell(const DevScene &S, const DevTexture &T, int32_t I, int32_t J, int32_t K, V3 g[8]) {
    const uint32_t *perm = S.perms + T.permOffset;
    for (int32_t z = 0; z < 2; ++z)
        for (int32_t y = 0; y < 2; ++y)
            for (int32_t x = 0; x < 2; ++x) {
                uint32_t h = perm[mod256(I + x + (int32_t)perm[mod256(J + y + (int32_t)perm[mod256(K + z)])])] & 255u;
                g[4 * z + 2 * y + x] = v3(S.perlinGrad + 3 * h);
            }
}

// The noise at the scaled point P of cell (I, J, K) from its corners'
// gradients, then the texture's appearance (texture.h:161-200).
__device__ inline float perlin_value(const DevTexture &T, const V3 g[8], int32_t I, int32_t J, int32_t K, V3 P) {
    float r = 0.0f;
    for (int32_t z = 0; z < 2; ++z)
        for (int32_t y = 0; y < 2; ++y)
            for (int32_t x = 0; x < 2; ++x) {
                V3 v = v3(P.x - (float)(I + x), P.y - (float)(J + y), P.z - (float)(K + z));
                r += dot(g[4 * z + 2 * y + x], v) * perlin_weight(v.x) * perlin_weight(v.y) * perlin_weight(v.z);
            }
    if (T.perlinAppearance == 0) r = (r + 1.0f) / 2.0f;
    else r = fabsf(r);
    return r;
}

__device__ inline float sample_perlin(const DevScene &S, const DevTexture &T, V3 P) {
    P = P * T.scaling;
    int32_t I = (int32_t)floorf(P.x), J = (int32_t)floorf(P.y), K = (int32_t)floorf(P.z);
    V3 g[8];
    perlin_cell(S, T, I, J, K, g);
    return perlin_value(T, g, I, J, K, P);
}

__device__ inline V3 fetch2d(const DevScene &S, const DevTexture &T, uint32_t x, uint32_t y) {
    uint32_t idx = (y * (uint32_t)T.width + x) * 3u;
    if (T.pixelFormat == 0) {
        const uint8_t *p = S.texels + T.dataOffset;
        return v3((float)p[idx], (float)p[idx + 1], (float)p[idx + 2]);
    }
    const float *p = reinterpret_cast<const float *>(S.texels + T.dataOffset);
    return v3(p[idx], p[idx + 1], p[idx + 2]);
}

__device__ inline V3 sample_texture_2d(const DevScene &S, const DevTexture &T, float u, float v, bool norm) {
    V3 r;
    if (T.wrap == 0) {
        float mu = u < 1.0f ? u : 1.0f, mv = v < 1.0f ? v : 1.0f;
        u = 0.0f > mu ? 0.0f : mu;
        v = 0.0f > mv ? 0.0f : mv;
    } else {
        u = u - floorf(u);
        v = v - floorf(v);
    }
    const uint32_t W = (uint32_t)T.width, H = (uint32_t)T.height;
    if (T.filter == 0) {
        uint32_t fx = (uint32_t)floorf(u * ((float)T.width - 1.0f) + 0.5f);
        uint32_t fy = (uint32_t)floorf(v * ((float)T.height - 1.0f) + 0.5f);
        fx = fx < W ? fx : W - 1u;
        fy = fy < H ? fy : H - 1u;
        r = fetch2d(S, T, fx, fy);
    } else {
        float su = u * (float)T.width, sv = v * (float)T.height;
        uint32_t ax = (uint32_t)(su + 0.5f) % W;
        uint32_t ay = (uint32_t)(sv + 0.5f) % H;
        uint32_t bx = (ax - 1u) % W, by = ay % H;
        uint32_t cx = (ax - 1u) % W, cy = (ay - 1u) % H;
        uint32_t dx = ax % W, dy = (ay - 1u) % H;
        V3 A = fetch2d(S, T, ax, ay), B = fetch2d(S, T, bx, by), C = fetch2d(S, T, cx, cy), D = fetch2d(S, T, dx, dy);
        su = floorf(su + 0.5f);
        sv = floorf(sv + 0.5f);
        if (su >= T.width) su -= T.width;
        if (sv >= T.height) sv -= T.height;
        float fxr = su - (float)ax, fyr = sv - (float)ay;
        r = A * (1.0f - fxr) * (1.0f - fyr) + B * (fxr) * (1.0f - fyr) + C * (fxr) * (fyr) +
            D * (1.0f - fxr) * (fyr);
    }
    if (norm) {
        r.x /= T.normalizer;
        r.y /= T.normalizer;
        r.z /= T.normalizer;
    }
    return r;
}

__device__ inline V3 sample_texture(const DevScene &S, const DevTexture &T, V3 P, float u, float v) {
    if (T.type == 0) return sample_texture_2d(S, T, u, v, true);
    float n = sample_perlin(S, T, P);
    return v3(n, n, n);
}

// ---- BRDFs and light transport (main.cpp:461-603) ---------------------------
__device__ inline void evaluate_brdf(float &dif, float &spec, const DevMaterial &M, V3 L, V3 N, V3 Vw,
                                     float cosI) {
    // Types 0-2 (Phong family) take the cosine to the mirror direction, types
    // 3-6 (Blinn-Phong family, Torrance-Sparrow) the cosine to the half vector;
    // every type raises it to the exponent once.
    const int type = M.brdfType >= 0 && M.brdfType <= 5 ? M.brdfType : 6;   // 6 and unknown: Torrance-Sparrow
    const bool half = type >= 3;
    V3 H = v3(0.0f, 0.0f, 0.0f);
    float c;
    if (half) {
        H = normalize(-Vw + L);
        c = dot(N, H);
    } else {
        V3 R = reflect(L, N);
        c = dot(-Vw, R);
    }
    if (type == 2 || type >= 5) dif = (float)(1.0f / kPiD);
    if (!(c > 0.0f)) {
        spec = 0.0f;
        return;
    }
    const float e = M.exponent;
    const float p = PowF(c, e);
    switch (type) {
    case 0: spec = p / cosI; break;                                  // Phong
    case 1: spec = p; break;                                         // ModifiedPhong
    case 2: spec = ((e + 2.0f) / (2.0f * kPiF)) * p; break;          // NormalizedModifiedPhong
    case 3: spec = p / cosI; break;                                  // BlinnPhong
    case 4: spec = p; break;                                         // ModifiedBlinnPhong
    case 5: spec = ((e + 8.0f) / (8.0f * kPiF)) * p; break;          // NormalizedModifiedBlinnPhong
    default: {                                                       // TorranceSparrow
        float D = ((e + 2.0f) / (2.0f * kPiF)) * p;
        float nw0 = dot(N, -Vw);
        float w0wh = dot(-Vw, H);
        float G = nw0 < cosI ? nw0 : cosI;
        G = (2.0f * G * c) / w0wh;
        G = 1.0f < G ? 1.0f : G;
        float ri = M.brdfIor;
        float R0 = sq(ri - 1.0f) / sq(ri + 1.0f);
        float F = R0 + (1.0f - R0) * PowIF<5>(1.0f - w0wh);
        spec = (D * F * G) / (4.0f * cosI * nw0);
    } break;
    }
}

// CalculateLightContribution (main.cpp:571-603) as the amount a light with
// intensity Li adds to Result = Kd*DC + Ks*SC: returns false when cos <= 0.
__device__ inline bool light_term(const DevMaterial &M, V3 kd, V3 Li, V3 L, V3 Vw, V3 N, float dist,
                                  float falloff, float invPdf, V3 &out) {
    float d2 = sq(dist);
    V3 Ln = normalize(L);
    V3 inc = v3(Li.x / d2, Li.y / d2, Li.z / d2) * falloff;
    float cosI = dot(N, Ln);
    if (!(cosI > 0.0f)) return false;
    float dif = 1.0f, spec = 1.0f;
    evaluate_brdf(dif, spec, M, Ln, N, Vw, cosI);
    V3 cm = inc * cosI * invPdf;
    out = had(kd, cm * dif) + had(v3(M.ks), cm * spec);
    return true;
}

// The factor the scatter child's radiance is multiplied by in the same
// formula (Li = child radiance, dist = falloff = 1).
__device__ inline bool scatter_weight(const DevMaterial &M, V3 kd, V3 S, V3 Vw, V3 N, float invPdf, V3 &out) {
    V3 Sn = normalize(S);
    float cosI = dot(N, Sn);
    if (!(cosI > 0.0f)) return false;
    float dif = 1.0f, spec = 1.0f;
    evaluate_brdf(dif, spec, M, Sn, N, Vw, cosI);
    float c = cosI * invPdf;
    out = had(kd, v3(c * dif, c * dif, c * dif)) + had(v3(M.ks), v3(c * spec, c * spec, c * spec));
    return true;
}

// ---- BRDF lobe sampling of the hemisphere scatter -------------------------------
// An option (the reference draws scatter directions from the cosine or the
// uniform hemisphere only, main.cpp:1029-1083): a one-sample mixture of the
// cosine lobe and the material's specular lobe, the reference's BRDF value
// divided by the mixture's density, so the estimate stays unbiased.
//
// The specular lobe: 0 none (black specular or no exponent), 1 cos^e about
// the viewer's mirror direction (the three Phong models: their lobe is
// (L . reflect(V, N))^e), 2 cos^e(N . H) of the half vector H about the
// normal (the three Blinn models and Torrance-Sparrow's Blinn D).
__device__ inline int spec_lobe(const DevMaterial &M) {
    if (!(M.exponent > 0.0f) || (M.ks[0] == 0.0f && M.ks[1] == 0.0f && M.ks[2] == 0.0f)) return 0;
    return M.brdfType <= 2 ? 1 : 2;
}

// Probability of drawing from the specular lobe: its share of the two
// lobes' approximate albedos (luminance of the reflectance times the lobe's
// cosine-weighted integral: 1 for the normalized models' lobes and
// Torrance-Sparrow's diffuse; pi for the unnormalized diffuse (Kd);
// 2 pi / (e + 2) or 8 pi / (e + 8) for the unnormalized Phong and Blinn
// lobes; Torrance-Sparrow's specular scaled by its Schlick term at the
// viewing angle), kept within [0.1, 0.9] so that neither lobe is starved.
// Only the noise depends on this choice, not the expected value.
__device__ inline float spec_prob(const DevMaterial &M, V3 kd, V3 N, V3 Vo) {
    float ls = 0.2126f * M.ks[0] + 0.7152f * M.ks[1] + 0.0722f * M.ks[2];
    float ld = 0.2126f * kd.x + 0.7152f * kd.y + 0.0722f * kd.z;
    const float e = M.exponent;
    switch (M.brdfType) {
    case 0:
    case 1: ls *= 2.0f * kPiF / (e + 2.0f); ld *= kPiF; break;
    case 3:
    case 4: ls *= 8.0f * kPiF / (e + 8.0f); ld *= kPiF; break;
    case 2:
    case 5: break;
    default: {
        const float r0 = sq(M.brdfIor - 1.0f) / sq(M.brdfIor + 1.0f);
        ls *= r0 + (1.0f - r0) * powf(1.0f - fminf(1.0f, fmaxf(0.0f, dot(N, Vo))), 5.0f);
    } break;
    }
    return fminf(0.9f, fmaxf(0.1f, ls / fmaxf(ls + ld, 1e-20f)));
}

// Density (solid angle) of unit direction L under the specular lobe, Vo the
// unit direction towards the viewer: (e + 1) / 2pi cos^e about the lobe's
// axis; for the half-vector lobe divided by 4 (Vo . H), the Jacobian of
// reflecting Vo about H.
__device__ inline float spec_pdf(int lobe, float e, V3 L, V3 N, V3 Vo) {
    if (lobe == 1) {
        const float c = dot(L, reflect(Vo, N));
        return c > 0.0f ? (e + 1.0f) / (2.0f * kPiF) * powf(c, e) : 0.0f;
    }
    const V3 H = normalize(Vo + L);
    const float c = dot(N, H), vh = dot(Vo, H);
    if (!(c > 0.0f) || !(vh > 0.0f)) return 0.0f;
    return (e + 1.0f) / (2.0f * kPiF) * powf(c, e) / (4.0f * vh);
}

// Density of unit direction L under the mixture (1 - ps) cos/pi + ps specular.
__device__ inline float brdf_mixture_pdf(const DevMaterial &M, V3 kd, V3 N, V3 Vo, V3 L) {
    const int lobe = spec_lobe(M);
    const float ps = lobe ? spec_prob(M, kd, N, Vo) : 0.0f;
    return (1.0f - ps) * fmaxf(dot(N, L), 0.0f) / kPiF + (lobe ? ps * spec_pdf(lobe, M.exponent, L, N, Vo) : 0.0f);
}

// A scatter direction from the mixture (1 - ps) cos/pi + ps specular, from
// three uniforms (lobe choice, then the lobe's two), with the mixture's
// density at it.  Directions below the surface can come from the specular
// lobe; their BRDF term is zero (the caller emits no child), and the density
// still counts them.
__device__ inline V3 sample_brdf_mixture(const DevMaterial &M, V3 kd, V3 N, V3 Vo, float u0, float u1, float u2,
                                         float &pdf) {
    const int lobe = spec_lobe(M);
    const float ps = lobe ? spec_prob(M, kd, N, Vo) : 0.0f;
    const float e = M.exponent;
    const float phi = 2.0f * kPiF * u2;
    V3 L, U, V;
    if (u0 >= ps) {
        const float st = sqrtf(u1), ct = sqrtf(fmaxf(0.0f, 1.0f - u1));
        orthonormal_basis(N, U, V);
        L = normalize(U * (st * cosf(phi)) + V * (st * sinf(phi)) + N * ct);
    } else {
        const float ca = powf(u1, 1.0f / (e + 1.0f)), sa = sqrtf(fmaxf(0.0f, 1.0f - ca * ca));
        const V3 A = lobe == 1 ? reflect(Vo, N) : N;
        orthonormal_basis(A, U, V);
        const V3 X = normalize(U * (sa * cosf(phi)) + V * (sa * sinf(phi)) + A * ca);
        L = lobe == 1 ? X : normalize(reflect(Vo, X));
    }
    pdf = (1.0f - ps) * fmaxf(dot(N, L), 0.0f) / kPiF + (lobe ? ps * spec_pdf(lobe, e, L, N, Vo) : 0.0f);
    return L;
}

}  // namespace az
