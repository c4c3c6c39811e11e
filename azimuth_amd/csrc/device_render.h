// device_render.h -- shading data in HBM and the per-vertex shading
// arithmetic of the wavefront renderer.
//
// Each device function restates one reference function (cited) in the same
// operation order as the CPU oracle; transcendental functions are the
// device's (ocml) and agree with glibc to a few ulp, which the image
// tolerance covers.
#pragma once

#include <stdint.h>

#include "device_scene.h"

namespace az {

// ---- shading records (uploaded once per scene) ----------------------------

enum : uint32_t { kMatTransparent = 1u, kMatMirror = 2u };

struct DevMaterial {                 // material.h:1-19 (+ brdf.h:11-17)
    float kd[3], ks[3], ka[3], km[3], tr[3];
    float ior, gloss, exponent, brdfIor;
    int32_t brdfType;
    uint32_t flags;                  // kMatTransparent / kMatMirror: bit patterns non-zero (main.cpp:1092-1112)
};

struct DevTexture {                  // texture.h:52-86
    int32_t type, pixelFormat, width, height;
    float invW, invH, normalizer, bumpMul;
    int32_t bump, wrap, blend, filter;
    int32_t perlinAppearance;
    float scaling;
    uint32_t dataOffset;             // byte offset into the texel pool (2D)
    uint32_t permOffset;             // word offset into the permutation pool (Perlin)
};

enum : int32_t { kLightPoint = 0, kLightArea = 1, kLightDirectional = 2, kLightSpot = 3, kLightEntity = 4 };

struct DevLight {                    // light.h:9-42
    int32_t type;
    int32_t entityKind;              // entity light: kItemSphere or kItemInstance
    uint32_t entityId;               // entity light: UniqueEntityID of the emitter
    uint32_t triOffset;              // entity mesh light: first entry in the light-triangle pool
    float pos[3], intensity[3], e0[3], e1[3], normal[3], dir[3];
    float cosCov, cosFall;
    float center[3], radius;         // entity sphere
    float radiance[3], totalArea;
    uint32_t triCount, pad0, pad1, pad2;
    Affine inv, fwd;                 // entity transforms
};

struct DevLightTri {                 // mesh triangle array, entity.h:15-20
    float p0[3], p1[3], p2[3], n[3];
};

struct DevPrimAttr {                 // per triangle, in primitive-id order
    float n0[3], n1[3], n2[3], fn[3];
    float uv0[2], uv1[2], uv2[2];
    float p0[3], p1[3], p2[3];
    int32_t smooth;
};

struct DevItemShade {                // per top-level item
    uint32_t material, texture, isEmitter, kind;
    float radiance[3];
    uint32_t shadeClass;             // sort class of its hits in shading (emitter, material flags, BRDF model)
};

struct DevScene {
    const DevMaterial *materials;
    const DevTexture *textures;      // 1-based index in entities; entry 0 unused
    const uint8_t *texels;
    const uint32_t *perms;
    const DevLight *lights;
    const DevLightTri *lightTris;
    const DevPrimAttr *prims;
    const DevItemShade *itemShade;
    const float *perlinGrad;         // 256 x 3
    const uint64_t *pcgTable;        // 8 bytes x 256 values x (multiplier, increment factor)
    DevTexture env;                  // environment map (type 2D) when useEnv
    DevTexture background;
    uint32_t lightCount;
    uint32_t maxDepth;
    uint32_t pathTrace, importance, useEnv, useBackgroundTexture;
    float ambient[3];
    float backgroundColor[3];
    float epsIntersection, epsShadow;
    uint32_t triTotal;
    uint32_t maxChildren;            // most path rays one shading point can emit (1..4)
    uint32_t entityLights;           // entity lights (their shadow rays need a closest hit)
    float boxMin[3], boxInv[3];      // scene box: guide positions are (P - min) * inv
    // Environment luminance distribution over the (u, v) cells of the map:
    // row CDF (envH + 1 floats) then one column CDF per row (envW + 1 each).
    const float *envCdf;
    uint32_t envW, envH;
};

// ---- render queues --------------------------------------------------------

// A ray of the path tree waiting to be traced.  64 bytes.
struct alignas(16) PathRay {
    float ox, oy, oz, time;
    float dx, dy, dz;
    uint32_t node;                   // random-stream node of the vertex this ray will create
    float wx, wy, wz;                // throughput to the sample
    uint32_t sample;                 // chunk-local sample index
    float cx, cy, cz;                // Beer coefficients (Transparency - 1) applied over this ray's distance
    uint32_t flags;                  // bits 0..7 depth, bit 8 inside dielectric
};

struct alignas(16) PathHit {
    uint32_t prim, item;
    float t, tl, beta, gamma;
    float pad0, pad1;
};

// A shadow ray and what it adds to its sample when it reaches the light.
enum : uint32_t { kShadowAny = 0u, kShadowEntity = 1u };

struct alignas(16) ShadowRay {
    float ox, oy, oz, time;
    float dx, dy, dz, tmax;
    float cx, cy, cz;                // contribution when unoccluded
    uint32_t sample;
    uint32_t mode;                   // kShadowAny or kShadowEntity
    uint32_t entityId;               // kShadowEntity: the light's UniqueEntityID
    uint32_t rec;                    // environment guide training record of this ray, or kEmpty
    uint32_t srec;                   // scatter guide record whose path this ray's vertex is on, or kEmpty
};

struct CameraSetup {                 // main.cpp:1265-1347
    float base[3], pos[3], ax[3], ay[3], stepX[3], stepY[3];
    float aperture;
    uint32_t width, height, sppSqrt;
    float invSppSqrt;
    int32_t kernel;
    float oneOverTwoPiSigma2, negOneOverTwoSigma2;
};

struct ChunkParams {
    uint64_t seed;
    uint32_t row0, rows;             // rows of this render
    uint32_t s0, spp;                // samples [s0, s0 + spp) of every pixel
    uint32_t capacity;               // path queue capacity
    uint32_t shadowCapacity;
    uint32_t cull;
    uint32_t gaussian;
    uint32_t envLuminance;           // environment NEE from the luminance distribution
    uint32_t brdfSampling;           // hemisphere scatter from the BRDF lobe mixture
};

struct RenderState {
    bool ready = false;
};

}  // namespace az
