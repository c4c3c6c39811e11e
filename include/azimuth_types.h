/*
 * azimuth_types.h -- the scene-description types of the azimuth renderer,
 * kept field-for-field so that a caller of the reference render call can hand
 * the same structures to this library.
 *
 * Each type names the reference declaration it keeps.  Member names, member
 * order and meaning are the reference's; the only changes are those needed for
 * the declarations to be portable C/C++ on Linux:
 *   - vec2/vec3/vec4 expose one set of member names (x,y,z,w).  The reference
 *     also aliases them as r,g,b / u,v and, for the camera near plane,
 *     r,l,b,t = x,y,z,w (math.h:88-200).  Those aliases were anonymous structs
 *     with repeated member names, which only MSVC accepts.
 *   - entity's MeshInstance and BVHNode members share their leading `AABB`
 *     member through one struct instead of two overlapping ones
 *     (entity.h:51-102); the byte offsets are unchanged.
 *
 * Units are the reference's internal units: light intensities, radiances,
 * background and ambient colours are the XML values divided by 255
 * (xml.cpp:178-180, 980-982, 1132-1134, 1201-1235).
 */
#ifndef AZIMUTH_TYPES_H
#define AZIMUTH_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* common.h:12-28 */
typedef int64_t  s64;
typedef int32_t  s32;
typedef int16_t  s16;
typedef int8_t   s8;
typedef uint64_t u64;
typedef uint32_t u32;
typedef uint16_t u16;
typedef uint8_t  u8;
typedef double   f64;
typedef float    f32;
typedef u32      b32;

/* math.h:45-200, 395-430 */
typedef struct { f32 x, y; } vec2;
typedef struct { u32 x, y; } vec2u;
typedef struct { f32 x, y, z; } vec3;
typedef struct { u32 x, y, z; } vec3u;
typedef struct { f32 x, y, z, w; } vec4;  /* near plane: x=r, y=l, z=b, w=t */
typedef struct { f32 E[4][4]; } mat4;     /* row-major, E[row][col] */
typedef struct { vec3 Min; vec3 Max; vec3 Center; } aabb;

/* io.h:1-7 */
typedef enum {
    FileFormat_PNG, FileFormat_BMP, FileFormat_TGA, FileFormat_JPG, FileFormat_EXR
} file_format;

/* brdf.h:1-17 */
typedef enum {
    BRDF_Phong,
    BRDF_ModifiedPhong,
    BRDF_NormalizedModifiedPhong,
    BRDF_BlinnPhong,
    BRDF_ModifiedBlinnPhong,
    BRDF_NormalizedModifiedBlinnPhong,
    BRDF_TorranceSparrow
} brdf_type;

typedef struct {
    f32 Exponent;
    f32 RefractiveIndex;
    brdf_type Type;
    u32 ID;
} brdf;

/* texture.h:20-86 */
#define AZ_PERLIN_GRADIENT_SIZE 256
typedef enum { TexturePerlinAppearance_Patch, TexturePerlinAppearance_Vein } texture_perlin_appearance;
typedef enum { TextureWrapMode_Clamp, TextureWrapMode_Repeat } texture_wrap_mode;
typedef enum {
    TextureBlendMode_ReplaceAll, TextureBlendMode_ReplaceDiffuse, TextureBlendMode_BlendDiffuse
} texture_blend_mode;
typedef enum { TextureFilter_Nearest, TextureFilter_Bilinear } texture_filter;
typedef enum { TextureType_2D, TextureType_PerlinNoise } texture_type;
typedef enum { PixelFormat_RGB_U8, PixelFormat_RGB_F32 } pixel_format;

typedef struct az_texture {
    char ImagePath[256];
    union {
        struct {
            void *Pixels;
            pixel_format PixelFormat;
            u32 Width;
            u32 Height;
            f32 InvWidth;
            f32 InvHeight;
            f32 Normalizer;
            b32 Loaded;
            file_format ImageFileFormat;
        };
        struct {
            u32 Perm[AZ_PERLIN_GRADIENT_SIZE];
            f32 ScalingFactor;
            texture_perlin_appearance PerlinAppearance;
        };
    };
    f32 BumpMapMultiplier;
    b32 BumpMap;
    b32 Degamma;
    u32 ID;
    texture_wrap_mode WrapMode;
    texture_blend_mode BlendMode;
    texture_filter Filter;
    texture_type Type;
} az_texture;
/* HIP's runtime headers declare a global `texture` template, so translation
 * units compiled as HIP use the tag name az_texture; everyone else also gets
 * the reference's name. */
#if !defined(__HIPCC__)
typedef az_texture texture;
#endif

/* material.h:1-19 */
typedef struct {
    brdf BRDF;
    vec3 AmbientReflectance;
    vec3 DiffuseReflectance;
    vec3 SpecularReflectance;
    vec3 MirrorReflectance;
    vec3 Transparency;
    f32 RefractionIndex;
    f32 Glossiness;
    b32 Degamma;
    u32 BRDFID;
    u32 ID;
} material;

/* entity.h:1-102 */
typedef enum {
    EntityType_Sphere, EntityType_Triangle, EntityType_MeshInstance, EntityType_BVHNode
} entity_type;
typedef enum { ShadingMode_Flat, ShadingMode_Smooth } shading_mode;

typedef struct entity entity;

typedef struct {
    vec3 P0;
    vec3 P1;
    vec3 P2;
    vec3 N;
} triangle;

typedef struct {
    mat4 InvTransformationMatrix;
    mat4 TransformationMatrix;
    vec3 MotionVector;
    vec3 Radiance;
    entity *BVH;
    triangle *Triangles;
    f32 TotalArea;
    s32 UVOffset;
    b32 IsEmitter;
    u32 TriangleCount;
    u32 TextureIndex;
    u32 MaterialIndex;
    u32 ID;
    shading_mode ShadingMode;
} mesh;

struct entity {
    union {
        struct {            /* Sphere */
            vec3 Center;
            f32 Radius;
        };
        struct {            /* Triangle */
            vec3 P0, P1, P2;
            vec3 N0, N1, N2;
            vec3 FaceNormal;
            vec2 UV0, UV1, UV2;
            shading_mode ShadingMode;
        };
        struct {            /* MeshInstance and BVHNode */
            aabb AABB;
            union {
                struct {    /* MeshInstance */
                    mesh *Mesh;
                    u32 BaseMeshID;
                    b32 ResetTransform;
                };
                struct {    /* BVHNode */
                    entity *Right;
                    entity *Left;
                };
            };
        };
    };
    mat4 InvTransformationMatrix;
    mat4 TransformationMatrix;
    vec3 MotionVector;
    vec3 Radiance;
    b32 IsEmitter;
    u32 TextureIndex;   /* 0: no texture, else 1-based index into scene Textures */
    u32 MaterialIndex;  /* index into scene Materials */
    u32 ID;
    u32 UniqueEntityID;
    entity_type Type;
};

/* light.h:1-42 */
typedef enum {
    LightType_PointLight, LightType_AreaLight, LightType_DirectionalLight,
    LightType_SpotLight, LightType_EntityLight
} light_type;

typedef struct {
    union {
        struct { vec3 Position; vec3 Intensity; vec3 EdgeVectors[2]; vec3 Normal; };
        struct { vec3 Direction; vec3 DirIntensity_; };       /* directional: Direction aliases Position */
        struct { vec3 SpotPosition_; vec3 SpotIntensity_; vec3 SpotDirection;
                 f32 CoverageAngle; f32 FallOffAngle; f32 CosHalfCoverage; f32 CosHalfFallOff; };
        struct { entity Entity; };
    };
    u32 ID;
    light_type Type;
} light;

/* camera.h:1-56 */
typedef enum { TonemapOperator_None, TonemapOperator_Photographic } tonemap_operator_type;
typedef struct {
    f32 KeyValue;
    f32 BurnoutPercentage;
    f32 Saturation;
    tonemap_operator_type Operator;
} tonemap;
typedef enum { CameraType_Detailed, CameraType_Simple } camera_type;

typedef struct {
    tonemap Tonemap;
    char ImageName[256];
    file_format SaveFormat;
    union {
        vec4 NearPlane;     /* detailed camera: x=r, y=l, z=b, w=t */
        f32 FovY;           /* simple camera, degrees */
    };
    vec3 Position;
    vec3 Gaze;
    vec3 Up;
    vec3 BackgroundColor;
    vec2u Resolution;
    f32 NearDistance;
    f32 FocusDistance;
    f32 ApertureSize;
    f32 Gamma;
    f32 InvGamma;
    b32 GammaCorrection_sRGB;
    b32 GazeIsPoint;
    b32 IsLeftHanded;
    b32 SaveEXR;
    u32 SampleCount;
    u32 ID;
    camera_type Type;
} camera;

/* scene.h:1-30 */
typedef struct {
    f32 Intersection;
    f32 ShadowRay;
} epsilon_values;

typedef struct {
    vec3 AmbientLight;
    vec3 BackgroundColor;
    az_texture BackgroundTexture;
    entity *BVHRoot;
    mesh *Meshes;
    material *Materials;
    light *Lights;
    az_texture *Textures;
    az_texture EnvironmentMap;
    u32 MeshCount;
    u32 MaterialCount;
    u32 LightCount;
    u32 TextureCount;
    u32 MaxRecursionDepth;
    b32 UseBackgroundTexture;
    b32 UseEnvironmentMap;
    b32 PathTrace;
    b32 ImportanceSampling;
} scene;

/* main.cpp:35-48 */
typedef struct { vec2u Max; vec2u Min; } scissor;
typedef enum { FilterType_Box, FilterType_Gaussian } filter_type;
typedef struct { vec3 Color; f32 TotalWeight; } pixel;

#ifdef __cplusplus
}
#endif

#endif /* AZIMUTH_TYPES_H */
