/*
 * azimuth_scene.h -- building a reference-shaped `scene` from geometry held in
 * memory, in place of the reference's XML importer.
 *
 * The reference fills `scene` in ImportXML (xml.cpp:3021-3930).  Its XML
 * parsing is not part of this library; its mesh file readers (PLY and the
 * binary vertex and face files) and its post-process half (xml.cpp:3632-3918)
 * are restated here:
 *   - per-vertex normals as the normalized sum of unnormalized face cross
 *     products (xml.cpp:3632-3667);
 *   - one triangle entity per face with flat/smooth normals and UVs, the
 *     mesh's `triangle` array and TotalArea (xml.cpp:3669-3749);
 *   - a triangles-only BVH per mesh (xml.cpp:3751, bvh.cpp:223-328);
 *   - one MeshInstance entity per mesh, its transform and world AABB with
 *     motion (xml.cpp:3755-3875);
 *   - UniqueEntityID = position in the entity list (xml.cpp:3880-3887);
 *   - one entity light per emitting sphere/mesh (xml.cpp:3893-3904);
 *   - the top-level BVH over all entities (xml.cpp:3915, bvh.cpp:1-221).
 *
 * The produced scene is exactly what the reference render call reads, so the
 * GPU upload (azimuth_gpu.h) and any reference-style CPU renderer can consume
 * the same object.
 */
#ifndef AZIMUTH_SCENE_H
#define AZIMUTH_SCENE_H

#include "azimuth_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    AZ_OK = 0,
    AZ_ERR_INVALID_ARGUMENT = 1,
    AZ_ERR_OUT_OF_MEMORY = 2,
    AZ_ERR_DEVICE = 3,          /* a HIP runtime call failed */
    AZ_ERR_NO_SCENE = 4,        /* render/trace before a scene was uploaded */
    AZ_ERR_UNSUPPORTED = 5,     /* input outside what the kernels handle */
    AZ_ERR_NOT_BUILT = 6
} az_status;

typedef struct az_scene_builder az_scene_builder;

/* One triangle mesh.  Vertices are xyz triples, Faces are index triples. */
typedef struct {
    const f32 *Vertices;
    u32 VertexCount;
    const u32 *Faces;
    u32 FaceCount;
    const f32 *UVs;             /* optional uv pairs, indexed by vertex + UVOffset */
    u32 UVCount;
    s32 UVOffset;
    mat4 Transformation;        /* object -> world; the inverse is computed here */
    vec3 MotionVector;          /* world-space displacement over the shutter [0,1) */
    vec3 Radiance;
    b32 IsEmitter;              /* an emitting mesh also becomes an entity light */
    u32 TextureIndex;           /* 0: none, else 1-based */
    u32 MaterialIndex;
    shading_mode ShadingMode;
} az_mesh_desc;

typedef struct {
    vec3 Center;
    f32 Radius;
    mat4 Transformation;
    vec3 MotionVector;
    vec3 Radiance;
    b32 IsEmitter;              /* an emitting sphere also becomes an entity light */
    u32 TextureIndex;
    u32 MaterialIndex;
} az_sphere_desc;

typedef struct {
    vec3 AmbientLight;
    vec3 BackgroundColor;
    u32 MaxRecursionDepth;
    b32 PathTrace;
    b32 ImportanceSampling;
} az_scene_params;

az_status az_sb_create(az_scene_builder **Out);
void      az_sb_destroy(az_scene_builder *Builder);
az_status az_sb_set_params(az_scene_builder *Builder, const az_scene_params *Params);
/* Appends a material; Index receives its position in scene->Materials. */
az_status az_sb_add_material(az_scene_builder *Builder, const material *Material, u32 *Index);
/* Appends a texture (pixel data is copied); Index receives the 1-based index.
 * With Texture->Degamma set, the copy gets LoadTexture's degamma step
 * (texture.h:391-413, 441-448: 8-bit channels c -> (u8)(powf(c / 255, 2.2)
 * * 255.001), float channels powf(c, 2.2)), as the reference's loader
 * applies it; so do the environment map and background texture calls. */
az_status az_sb_add_texture(az_scene_builder *Builder, const az_texture *Texture, u32 *Index);
/* Sets the RGB F32 or U8 environment map (pixel data is copied). */
az_status az_sb_set_environment_map(az_scene_builder *Builder, const az_texture *Texture);
/* Sets the RGB F32 or U8 background texture (pixel data is copied): camera
 * rays that miss every object sample it at (col / (W - 1), row / (H - 1))
 * unless an environment map is set (main.cpp:1180-1207, the
 * UseBackgroundTexture branch). */
az_status az_sb_set_background_texture(az_scene_builder *Builder, const az_texture *Texture);
/* Point, area, directional or spot light. */
az_status az_sb_add_light(az_scene_builder *Builder, const light *Light);
az_status az_sb_add_mesh(az_scene_builder *Builder, const az_mesh_desc *Mesh);
az_status az_sb_add_sphere(az_scene_builder *Builder, const az_sphere_desc *Sphere);
/* Runs the post-process and both BVH builds.  The scene is owned by the
 * caller and released with az_scene_free. */
az_status az_sb_build(az_scene_builder *Builder, scene **Out);
void      az_scene_free(scene *Scene);

/* The reference importer's mesh files, read into memory (xml.cpp:2397-2420,
 * 2621-2700, 3311-3630).  Malformed or truncated files give
 * AZ_ERR_INVALID_ARGUMENT, PLY files outside what the reference reads
 * (big-endian, property types other than float and the face list) give
 * AZ_ERR_UNSUPPORTED.
 *
 * PLY, ASCII or binary little-endian: the vertex element's float properties
 * (x y z, and u/s v/t when present: the reference's files carry x y z nx ny
 * nz u v, whose normals it drops), the face element's "list uchar int"
 * indices; quads become (a, b, c) and (a, c, d); faces of other sizes are an
 * error.  With Vertices, Faces and UVs NULL only the counts are returned;
 * with arrays, the counts give their capacity on entry (vertices, faces and
 * uv pairs).  UVCount is 0 when the file has no u v. */
az_status az_ply_read(const char *Path, f32 *Vertices, u32 *VertexCount, u32 *Faces, u32 *FaceCount, f32 *UVs,
                      u32 *UVCount);
/* A mesh whose vertices, faces and uvs come from a PLY file (the plyFile of
 * a <Mesh>, xml.cpp:2621-2660); Mesh gives the rest (transform, material,
 * texture, shading mode, UVOffset added to the vertex index for uv lookup);
 * its Vertices, Faces and UVs are ignored. */
az_status az_sb_add_ply_mesh(az_scene_builder *Builder, const char *Path, const az_mesh_desc *Mesh);
/* <VertexData binaryFile=...>: a u32 count, then count x y z f32 triples.
 * Vertices NULL: the count only; else *Count is the capacity on entry. */
az_status az_read_binary_vertices(const char *Path, f32 *Vertices, u32 *Count);
/* <Faces binaryFile=...>: a u32 count, then count u32 index triples;
 * VertexOffset is added to every index and 1 subtracted unless
 * ZeroBasedIndexing.  Faces NULL: the count only. */
az_status az_read_binary_faces(const char *Path, u32 VertexOffset, b32 ZeroBasedIndexing, u32 *Faces, u32 *Count);

/* 4x4 inverse used for entity transforms (double precision, then rounded). */
az_status az_mat4_inverse(const mat4 *In, mat4 *Out);

/* Number of triangles over all meshes (the primitive id space of triangles;
 * sphere ids follow). */
u32 az_scene_triangle_count(const scene *Scene);

#ifdef __cplusplus
}
#endif

#endif /* AZIMUTH_SCENE_H */
