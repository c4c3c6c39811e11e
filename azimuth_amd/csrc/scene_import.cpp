// scene_import.cpp -- builds a reference-shaped `scene` from in-memory
// geometry: the post-process half of the reference importer
// (xml.cpp:3632-3918) and the reference BVH construction (bvh.cpp:1-328).
//
// The BVH built here is the reference's binary entity tree: midpoint split of
// the node box centre on axis depth%3, the partition loop starting at index 1
// (so the first entity drifts right), half split when one side is empty
// (bvh.cpp:145-190, 279-307).  The partition is run on an index permutation
// and the entities are moved once at the end; the resulting order equals the
// reference's in-place swapping because both apply the same swaps.
#include "azimuth_gpu.h"
#include "azimuth_scene.h"

#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

namespace {

inline vec3 V3(f32 x, f32 y, f32 z) { vec3 r = {x, y, z}; return r; }
inline vec3 Sub(vec3 a, vec3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline vec3 Add(vec3 a, vec3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline vec3 Cross(vec3 a, vec3 b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline f32 Length(vec3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
inline vec3 Normalize(vec3 v) {
    f32 l = Length(v);
    if (l != 0.0f) return V3(v.x / l, v.y / l, v.z / l);
    return V3(0.0f, 0.0f, 0.0f);
}
inline f32 Axis(vec3 v, u32 d) { return d == 0 ? v.x : (d == 1 ? v.y : v.z); }

mat4 Identity() {
    mat4 m;
    memset(&m, 0, sizeof(m));
    m.E[0][0] = m.E[1][1] = m.E[2][2] = m.E[3][3] = 1.0f;
    return m;
}

mat4 Mul(const mat4 &a, const mat4 &b) {
    mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.E[i][j] = a.E[i][0] * b.E[0][j] + a.E[i][1] * b.E[1][j] +
                        a.E[i][2] * b.E[2][j] + a.E[i][3] * b.E[3][j];
    return r;
}

// mat4 * (x,y,z,1), rows 0..2 (math.cpp:1401-1426 with w = 1).
vec3 TransformPoint(const mat4 &m, vec3 v) {
    return V3(m.E[0][0] * v.x + m.E[0][1] * v.y + m.E[0][2] * v.z + m.E[0][3] * 1.0f,
              m.E[1][0] * v.x + m.E[1][1] * v.y + m.E[1][2] * v.z + m.E[1][3] * 1.0f,
              m.E[2][0] * v.x + m.E[2][1] * v.y + m.E[2][2] * v.z + m.E[2][3] * 1.0f);
}

struct Box {
    vec3 mn, mx;
    void reset() {
        mn = V3(FLT_MAX_, FLT_MAX_, FLT_MAX_);
        mx = V3(-FLT_MAX_, -FLT_MAX_, -FLT_MAX_);
    }
    void grow(vec3 p) {
        if (mn.x > p.x) mn.x = p.x;
        if (mn.y > p.y) mn.y = p.y;
        if (mn.z > p.z) mn.z = p.z;
        if (mx.x < p.x) mx.x = p.x;
        if (mx.y < p.y) mx.y = p.y;
        if (mx.z < p.z) mx.z = p.z;
    }
    static constexpr f32 FLT_MAX_ = 3.402823466e+38f;
};

aabb ToAABB(const Box &b) {
    aabb r;
    r.Min = b.mn;
    r.Max = b.mx;
    r.Center = V3(b.mn.x * 0.5f + b.mx.x * 0.5f, b.mn.y * 0.5f + b.mx.y * 0.5f,
                  b.mn.z * 0.5f + b.mx.z * 0.5f);
    return r;
}

// World box of a sphere entity as the reference computes it: the local box,
// grown by the 8 transformed local corners, grown by the box shifted by the
// motion vector (bvh.cpp:22-72).
Box SphereBox(const entity &e) {
    vec3 mn = V3(e.Center.x - e.Radius, e.Center.y - e.Radius, e.Center.z - e.Radius);
    vec3 mx = V3(e.Center.x + e.Radius, e.Center.y + e.Radius, e.Center.z + e.Radius);
    vec3 corners[8] = {V3(mn.x, mn.y, mn.z), V3(mn.x, mn.y, mx.z), V3(mn.x, mx.y, mn.z),
                       V3(mn.x, mx.y, mx.z), V3(mx.x, mn.y, mn.z), V3(mx.x, mn.y, mx.z),
                       V3(mx.x, mx.y, mn.z), V3(mx.x, mx.y, mx.z)};
    Box b;
    b.mn = mn;
    b.mx = mx;
    for (int c = 0; c < 8; ++c) b.grow(TransformPoint(e.TransformationMatrix, corners[c]));
    if (Length(e.MotionVector) > 0.0f) {
        b.grow(Add(b.mn, e.MotionVector));
        b.grow(Add(b.mx, e.MotionVector));
    }
    return b;
}

// The binary tree over a contiguous entity array.  `ItemBox(i)` and
// `ItemPos(i, dim)` describe entity i; the permutation is applied to `Items`
// when the tree is complete and the leaves then point into `Items`.
template <typename BoxFn, typename PosFn>
bool BuildReferenceTree(entity *Items, u32 Count, std::vector<entity *> &NodeStore,
                        entity **Root, BoxFn ItemBox, PosFn ItemPos) {
    if (Count == 0) return false;
    std::vector<u32> perm(Count);
    for (u32 i = 0; i < Count; ++i) perm[i] = i;
    std::vector<Box> boxes(Count);
    for (u32 i = 0; i < Count; ++i) boxes[i] = ItemBox(Items[i]);

    struct Pending { u32 begin, count, depth; entity **slot; };
    struct Link { entity **slot; u32 index; };   // leaf links resolved after permuting
    std::vector<Pending> work;
    std::vector<Link> leaves;
    work.push_back({0, Count, 0, Root});
    while (!work.empty()) {
        Pending p = work.back();
        work.pop_back();
        if (p.count == 1) {
            leaves.push_back({p.slot, p.begin});
            continue;
        }
        entity *node = (entity *)calloc(1, sizeof(entity));
        if (!node) return false;
        NodeStore.push_back(node);
        node->Type = EntityType_BVHNode;
        Box b;
        b.reset();
        for (u32 i = 0; i < p.count; ++i) {
            const Box &ib = boxes[perm[p.begin + i]];
            b.grow(ib.mn);
            b.grow(ib.mx);
        }
        node->AABB = ToAABB(b);
        *p.slot = node;
        u32 dim = p.depth % 3;
        f32 split = Axis(node->AABB.Center, dim);
        u32 left = 0;
        u32 *r = perm.data() + p.begin;
        for (u32 i = 1; i < p.count; ++i) {
            if (ItemPos(Items[r[i]], dim) < split) {
                u32 t = r[left];
                r[left] = r[i];
                r[i] = t;
                ++left;
            }
        }
        if (left == 0 || left == p.count) left = p.count / 2;
        // Right is pushed first so that the left subtree is expanded first;
        // the order of expansion does not change the result.
        work.push_back({p.begin + left, p.count - left, p.depth + 1, &node->Right});
        work.push_back({p.begin, left, p.depth + 1, &node->Left});
    }
    std::vector<entity> moved(Count);
    for (u32 i = 0; i < Count; ++i) moved[i] = Items[perm[i]];
    memcpy(Items, moved.data(), sizeof(entity) * Count);
    for (const Link &l : leaves) *l.slot = Items + l.index;
    return true;
}

}  // namespace

struct az_scene_builder {
    az_scene_params params;
    std::vector<material> materials;
    std::vector<az_texture> textures;
    std::vector<std::vector<u8>> texturePixels;
    az_texture envmap;
    std::vector<u8> envPixels;
    b32 useEnv = 0;
    az_texture background;
    std::vector<u8> backgroundPixels;
    b32 useBackground = 0;
    std::vector<light> lights;
    std::vector<az_mesh_desc> meshes;
    std::vector<std::vector<f32>> meshVerts, meshUVs;
    std::vector<std::vector<u32>> meshFaces;
    std::vector<az_sphere_desc> spheres;
};

// Everything a built scene owns; `s` is first so scene* and storage* coincide.
struct SceneStorage {
    scene s;
    std::vector<mesh> meshes;
    std::vector<material> materials;
    std::vector<light> lights;
    std::vector<az_texture> textures;
    std::vector<std::vector<u8>> texturePixels;
    std::vector<u8> envPixels, backgroundPixels;
    std::vector<entity *> triEntities;   // one calloc'd array per mesh
    std::vector<triangle *> triangles;
    entity *topEntities = nullptr;
    std::vector<entity *> nodes;
    ~SceneStorage() {
        for (entity *e : triEntities) free(e);
        for (triangle *t : triangles) free(t);
        free(topEntities);
        for (entity *n : nodes) free(n);
    }
};

static_assert(offsetof(SceneStorage, s) == 0, "scene must be the first member");

extern "C" {

az_status az_mat4_inverse(const mat4 *In, mat4 *Out) {
    if (!In || !Out) return AZ_ERR_INVALID_ARGUMENT;
    double a[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j)
            a[i][j] = j < 4 ? (double)In->E[i][j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
        if (a[piv][c] == 0.0) return AZ_ERR_INVALID_ARGUMENT;
        if (piv != c)
            for (int j = 0; j < 8; ++j) { double t = a[c][j]; a[c][j] = a[piv][j]; a[piv][j] = t; }
        double inv = 1.0 / a[c][c];
        for (int j = 0; j < 8; ++j) a[c][j] *= inv;
        for (int r = 0; r < 4; ++r) {
            if (r == c || a[r][c] == 0.0) continue;
            double f = a[r][c];
            for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j];
        }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) Out->E[i][j] = (f32)a[i][j + 4];
    return AZ_OK;
}

az_status az_sb_create(az_scene_builder **Out) {
    if (!Out) return AZ_ERR_INVALID_ARGUMENT;
    az_scene_builder *b = new (std::nothrow) az_scene_builder();
    if (!b) return AZ_ERR_OUT_OF_MEMORY;
    memset(&b->params, 0, sizeof(b->params));
    b->params.MaxRecursionDepth = 1;
    memset(&b->envmap, 0, sizeof(b->envmap));
    *Out = b;
    return AZ_OK;
}

void az_sb_destroy(az_scene_builder *Builder) { delete Builder; }

az_status az_sb_set_params(az_scene_builder *B, const az_scene_params *P) {
    if (!B || !P) return AZ_ERR_INVALID_ARGUMENT;
    B->params = *P;
    return AZ_OK;
}

az_status az_sb_add_material(az_scene_builder *B, const material *M, u32 *Index) {
    if (!B || !M) return AZ_ERR_INVALID_ARGUMENT;
    B->materials.push_back(*M);
    if (Index) *Index = (u32)B->materials.size() - 1;
    return AZ_OK;
}

static size_t TexturePixelBytes(const az_texture *T) {
    if (T->Type != TextureType_2D || !T->Pixels) return 0;
    size_t texel = T->PixelFormat == PixelFormat_RGB_F32 ? 12 : 3;
    return (size_t)T->Width * T->Height * texel;
}

// The degamma step of LoadTexture (texture.h:391-413, 441-448) on a copy of
// a texture's texels: 8-bit channels c become (u8)(powf(c / 255, 2.2) *
// 255.001), float channels powf(c, 2.2).  The builder stands in for the
// reference's loader, so it applies the step when the texture asks for it;
// the GPU upload never does (a reference scene's textures were converted
// when they were loaded).
static void DegammaTexels(const az_texture *T, std::vector<u8> &texels) {
    if (T->Type != TextureType_2D || !T->Degamma) return;
    if (T->PixelFormat == PixelFormat_RGB_U8) {
        for (u8 &c : texels) {
            f32 v = powf((f32)c / 255.0f, 2.2f) * 255.001f;
            c = (u8)v;
        }
    } else {
        f32 *p = (f32 *)texels.data();
        for (size_t i = 0; i < texels.size() / sizeof(f32); ++i) p[i] = powf(p[i], 2.2f);
    }
}

az_status az_sb_add_texture(az_scene_builder *B, const az_texture *T, u32 *Index) {
    if (!B || !T) return AZ_ERR_INVALID_ARGUMENT;
    size_t n = TexturePixelBytes(T);
    B->textures.push_back(*T);
    B->texturePixels.emplace_back((const u8 *)T->Pixels, (const u8 *)T->Pixels + n);
    DegammaTexels(T, B->texturePixels.back());
    if (Index) *Index = (u32)B->textures.size();
    return AZ_OK;
}

az_status az_sb_set_environment_map(az_scene_builder *B, const az_texture *T) {
    if (!B || !T || T->Type != TextureType_2D || !T->Pixels || !T->Width || !T->Height)
        return AZ_ERR_INVALID_ARGUMENT;
    size_t n = TexturePixelBytes(T);
    B->envmap = *T;
    B->envPixels.assign((const u8 *)T->Pixels, (const u8 *)T->Pixels + n);
    DegammaTexels(T, B->envPixels);
    B->useEnv = 1;
    return AZ_OK;
}

az_status az_sb_set_background_texture(az_scene_builder *B, const az_texture *T) {
    if (!B || !T || T->Type != TextureType_2D || !T->Pixels || !T->Width || !T->Height)
        return AZ_ERR_INVALID_ARGUMENT;
    size_t n = TexturePixelBytes(T);
    B->background = *T;
    B->backgroundPixels.assign((const u8 *)T->Pixels, (const u8 *)T->Pixels + n);
    DegammaTexels(T, B->backgroundPixels);
    B->useBackground = 1;
    return AZ_OK;
}

az_status az_sb_add_light(az_scene_builder *B, const light *L) {
    if (!B || !L || L->Type == LightType_EntityLight) return AZ_ERR_INVALID_ARGUMENT;
    B->lights.push_back(*L);
    return AZ_OK;
}

az_status az_sb_add_mesh(az_scene_builder *B, const az_mesh_desc *M) {
    if (!B || !M || !M->Vertices || !M->Faces || M->FaceCount == 0) return AZ_ERR_INVALID_ARGUMENT;
    for (u32 f = 0; f < M->FaceCount * 3; ++f)
        if (M->Faces[f] >= M->VertexCount) return AZ_ERR_INVALID_ARGUMENT;
    B->meshes.push_back(*M);
    B->meshVerts.emplace_back(M->Vertices, M->Vertices + 3 * (size_t)M->VertexCount);
    B->meshFaces.emplace_back(M->Faces, M->Faces + 3 * (size_t)M->FaceCount);
    if (M->UVs && M->UVCount)
        B->meshUVs.emplace_back(M->UVs, M->UVs + 2 * (size_t)M->UVCount);
    else
        B->meshUVs.emplace_back();
    return AZ_OK;
}

az_status az_sb_add_sphere(az_scene_builder *B, const az_sphere_desc *S) {
    if (!B || !S || !(S->Radius > 0.0f)) return AZ_ERR_INVALID_ARGUMENT;
    B->spheres.push_back(*S);
    return AZ_OK;
}

az_status az_sb_build(az_scene_builder *B, scene **Out) {
    if (!B || !Out) return AZ_ERR_INVALID_ARGUMENT;
    if (B->meshes.empty() && B->spheres.empty()) return AZ_ERR_INVALID_ARGUMENT;
    SceneStorage *st = new (std::nothrow) SceneStorage();
    if (!st) return AZ_ERR_OUT_OF_MEMORY;
    scene &sc = st->s;
    memset(&sc, 0, sizeof(sc));
    sc.AmbientLight = B->params.AmbientLight;
    sc.BackgroundColor = B->params.BackgroundColor;
    sc.MaxRecursionDepth = B->params.MaxRecursionDepth;
    sc.PathTrace = B->params.PathTrace;
    sc.ImportanceSampling = B->params.ImportanceSampling;

    st->materials = B->materials;
    if (st->materials.empty()) {
        material m;
        memset(&m, 0, sizeof(m));
        m.BRDF.Type = BRDF_BlinnPhong;
        m.DiffuseReflectance = V3(0.5f, 0.5f, 0.5f);
        st->materials.push_back(m);
    }
    st->textures = B->textures;
    st->texturePixels = B->texturePixels;
    for (size_t i = 0; i < st->textures.size(); ++i)
        if (st->textures[i].Type == TextureType_2D)
            st->textures[i].Pixels = st->texturePixels[i].empty() ? nullptr : st->texturePixels[i].data();
    if (B->useEnv) {
        st->envPixels = B->envPixels;
        sc.EnvironmentMap = B->envmap;
        sc.EnvironmentMap.Pixels = st->envPixels.data();
        sc.UseEnvironmentMap = 1;
    }
    if (B->useBackground) {
        st->backgroundPixels = B->backgroundPixels;
        sc.BackgroundTexture = B->background;
        sc.BackgroundTexture.Pixels = st->backgroundPixels.data();
        sc.UseBackgroundTexture = 1;
    }

    // Entities in the reference's order: spheres as parsed, then one instance
    // per mesh appended by the post-process (xml.cpp:3755-3773).
    std::vector<entity> ents;
    std::vector<u32> entityLightIdx;
    for (const az_sphere_desc &s : B->spheres) {
        entity e;
        memset(&e, 0, sizeof(e));
        e.Type = EntityType_Sphere;
        e.Center = s.Center;
        e.Radius = s.Radius;
        e.TransformationMatrix = s.Transformation;
        if (az_mat4_inverse(&s.Transformation, &e.InvTransformationMatrix) != AZ_OK) {
            delete st;
            return AZ_ERR_INVALID_ARGUMENT;
        }
        e.MotionVector = s.MotionVector;
        e.Radiance = s.Radiance;
        e.IsEmitter = s.IsEmitter;
        e.TextureIndex = s.TextureIndex;
        e.MaterialIndex = s.MaterialIndex;
        e.ID = (u32)ents.size() + 1;
        if (s.IsEmitter) entityLightIdx.push_back((u32)ents.size());
        ents.push_back(e);
    }

    const u32 meshCount = (u32)B->meshes.size();
    st->meshes.resize(meshCount);
    for (u32 mi = 0; mi < meshCount; ++mi) {
        const az_mesh_desc &d = B->meshes[mi];
        const std::vector<f32> &vv = B->meshVerts[mi];
        const std::vector<u32> &ff = B->meshFaces[mi];
        const std::vector<f32> &uu = B->meshUVs[mi];
        auto vert = [&](u32 i) { return V3(vv[3 * i], vv[3 * i + 1], vv[3 * i + 2]); };

        // xml.cpp:3632-3667: vertex normals from unnormalized face normals.
        std::vector<vec3> vn(d.VertexCount, V3(0.0f, 0.0f, 0.0f));
        for (u32 f = 0; f < d.FaceCount; ++f) {
            vec3 p0 = vert(ff[3 * f]), p1 = vert(ff[3 * f + 1]), p2 = vert(ff[3 * f + 2]);
            vec3 fn = Cross(Sub(p1, p0), Sub(p2, p0));
            for (int k = 0; k < 3; ++k) vn[ff[3 * f + k]] = Add(vn[ff[3 * f + k]], fn);
        }
        for (vec3 &n : vn) n = Normalize(n);

        mesh &m = st->meshes[mi];
        memset(&m, 0, sizeof(m));
        m.TransformationMatrix = d.Transformation;
        if (az_mat4_inverse(&d.Transformation, &m.InvTransformationMatrix) != AZ_OK) {
            delete st;
            return AZ_ERR_INVALID_ARGUMENT;
        }
        m.MotionVector = d.MotionVector;
        m.Radiance = d.Radiance;
        m.IsEmitter = d.IsEmitter;
        m.TextureIndex = d.TextureIndex;
        m.MaterialIndex = d.MaterialIndex;
        m.ShadingMode = d.ShadingMode;
        m.UVOffset = d.UVOffset;
        m.ID = mi + 1;

        entity *tri = (entity *)calloc(d.FaceCount, sizeof(entity));
        triangle *tris = (triangle *)calloc(d.FaceCount, sizeof(triangle));
        if (!tri || !tris) {
            free(tri);
            free(tris);
            delete st;
            return AZ_ERR_OUT_OF_MEMORY;
        }
        st->triEntities.push_back(tri);
        st->triangles.push_back(tris);
        // xml.cpp:3683-3749: triangle entities, face normals, total area.
        f32 totalArea = 0.0f;
        const u32 uvCount = (u32)(uu.size() / 2);
        for (u32 f = 0; f < d.FaceCount; ++f) {
            entity &t = tri[f];
            t.ShadingMode = d.ShadingMode;
            t.Radiance = d.Radiance;
            t.IsEmitter = d.IsEmitter;
            t.TextureIndex = d.TextureIndex;
            t.MaterialIndex = d.MaterialIndex;
            t.ID = 0;
            t.Type = EntityType_Triangle;
            u32 i0 = ff[3 * f], i1 = ff[3 * f + 1], i2 = ff[3 * f + 2];
            vec3 p0 = vert(i0), p1 = vert(i1), p2 = vert(i2);
            t.P0 = p0;
            t.P1 = p1;
            t.P2 = p2;
            vec3 fn = Cross(Sub(p1, p0), Sub(p2, p0));
            f32 len = Length(fn);
            totalArea += len / 2.0f;
            fn.x /= len;
            fn.y /= len;
            fn.z /= len;
            t.FaceNormal = fn;
            t.N0 = vn[i0];
            t.N1 = vn[i1];
            t.N2 = vn[i2];
            s32 uv0 = (s32)i0 + d.UVOffset, uv1 = (s32)i1 + d.UVOffset, uv2 = (s32)i2 + d.UVOffset;
            if (uv0 >= 0 && (u32)uv0 < uvCount) { t.UV0.x = uu[2 * uv0]; t.UV0.y = uu[2 * uv0 + 1]; }
            if (uv1 >= 0 && (u32)uv1 < uvCount) { t.UV1.x = uu[2 * uv1]; t.UV1.y = uu[2 * uv1 + 1]; }
            if (uv2 >= 0 && (u32)uv2 < uvCount) { t.UV2.x = uu[2 * uv2]; t.UV2.y = uu[2 * uv2 + 1]; }
            tris[f].P0 = p0;
            tris[f].P1 = p1;
            tris[f].P2 = p2;
            tris[f].N = fn;
        }
        m.Triangles = tris;
        m.TotalArea = totalArea;
        m.TriangleCount = d.FaceCount;
        auto triBox = [](const entity &e) {
            Box b;
            b.reset();
            b.grow(e.P0);
            b.grow(e.P1);
            b.grow(e.P2);
            return b;
        };
        auto triPos = [](const entity &e, u32 dim) { return Axis(e.P0, dim); };
        if (!BuildReferenceTree(tri, d.FaceCount, st->nodes, &m.BVH, triBox, triPos)) {
            delete st;
            return AZ_ERR_OUT_OF_MEMORY;
        }
        entity inst;
        memset(&inst, 0, sizeof(inst));
        inst.InvTransformationMatrix = Identity();
        inst.TransformationMatrix = Identity();
        inst.BaseMeshID = mi;
        inst.ResetTransform = 0;
        inst.MotionVector = d.MotionVector;
        inst.Radiance = d.Radiance;
        inst.IsEmitter = d.IsEmitter;
        inst.TextureIndex = d.TextureIndex;
        inst.MaterialIndex = d.MaterialIndex;
        inst.Type = EntityType_MeshInstance;
        if (d.IsEmitter) entityLightIdx.push_back((u32)ents.size());
        ents.push_back(inst);
    }

    // xml.cpp:3791-3875: bind instances to meshes, compose transforms and
    // compute the world box of the mesh BVH root with motion.
    for (entity &e : ents) {
        if (e.Type != EntityType_MeshInstance) continue;
        mesh *m = &st->meshes[e.BaseMeshID];
        e.Mesh = m;
        // A one-triangle mesh has a triangle as its BVH root; the reference
        // then reads the triangle's bytes as a box (xml.cpp:3798).  Here it
        // gets the triangle's own box.
        if (m->BVH->Type == EntityType_BVHNode) {
            e.AABB = m->BVH->AABB;
        } else {
            Box b;
            b.reset();
            b.grow(m->BVH->P0);
            b.grow(m->BVH->P1);
            b.grow(m->BVH->P2);
            e.AABB = ToAABB(b);
        }
        if (!e.ResetTransform) {
            e.TransformationMatrix = Mul(e.TransformationMatrix, m->TransformationMatrix);
            e.InvTransformationMatrix = Mul(m->InvTransformationMatrix, e.InvTransformationMatrix);
        }
        vec3 mn = e.AABB.Min, mx = e.AABB.Max;
        vec3 corners[8] = {V3(mn.x, mn.y, mn.z), V3(mn.x, mn.y, mx.z), V3(mn.x, mx.y, mn.z),
                           V3(mn.x, mx.y, mx.z), V3(mx.x, mn.y, mn.z), V3(mx.x, mn.y, mx.z),
                           V3(mx.x, mx.y, mn.z), V3(mx.x, mx.y, mx.z)};
        Box t;
        t.reset();
        for (int c = 0; c < 8; ++c) t.grow(TransformPoint(e.TransformationMatrix, corners[c]));
        if (Length(e.MotionVector) > 0.0f) {
            Box init = t;
            t.mn = Add(t.mn, e.MotionVector);
            t.mx = Add(t.mx, e.MotionVector);
            if (t.mn.x > init.mn.x) t.mn.x = init.mn.x;
            if (t.mn.y > init.mn.y) t.mn.y = init.mn.y;
            if (t.mn.z > init.mn.z) t.mn.z = init.mn.z;
            if (t.mx.x < init.mx.x) t.mx.x = init.mx.x;
            if (t.mx.y < init.mx.y) t.mx.y = init.mx.y;
            if (t.mx.z < init.mx.z) t.mx.z = init.mx.z;
        }
        e.AABB = ToAABB(t);
    }

    const u32 entityCount = (u32)ents.size();
    st->topEntities = (entity *)calloc(entityCount, sizeof(entity));
    if (!st->topEntities) {
        delete st;
        return AZ_ERR_OUT_OF_MEMORY;
    }
    for (u32 i = 0; i < entityCount; ++i) {
        ents[i].UniqueEntityID = i;
        st->topEntities[i] = ents[i];
    }

    // Lights: the explicit ones, then one entity light per emitter
    // (xml.cpp:3889-3913).  The entity copy is taken before the top-level
    // BVH reorders the entity array, as in the reference.
    st->lights = B->lights;
    for (size_t i = 0; i < st->lights.size(); ++i) st->lights[i].ID = (u32)i + 1;
    for (u32 k = 0; k < (u32)entityLightIdx.size(); ++k) {
        light L;
        memset(&L, 0, sizeof(L));
        L.Entity = st->topEntities[entityLightIdx[k]];
        L.ID = k;
        L.Type = LightType_EntityLight;
        st->lights.push_back(L);
    }

    auto topBox = [](const entity &e) {
        if (e.Type == EntityType_Sphere) return SphereBox(e);
        Box b;
        b.mn = e.AABB.Min;
        b.mx = e.AABB.Max;
        return b;
    };
    auto topPos = [](const entity &e, u32 dim) {
        if (e.Type == EntityType_Sphere) return Axis(TransformPoint(e.TransformationMatrix, e.Center), dim);
        return Axis(e.AABB.Center, dim);
    };
    if (!BuildReferenceTree(st->topEntities, entityCount, st->nodes, &sc.BVHRoot, topBox, topPos)) {
        delete st;
        return AZ_ERR_OUT_OF_MEMORY;
    }

    sc.Meshes = st->meshes.empty() ? nullptr : st->meshes.data();
    sc.MeshCount = meshCount;
    sc.Materials = st->materials.data();
    sc.MaterialCount = (u32)st->materials.size();
    sc.Lights = st->lights.empty() ? nullptr : st->lights.data();
    sc.LightCount = (u32)st->lights.size();
    sc.Textures = st->textures.empty() ? nullptr : st->textures.data();
    sc.TextureCount = (u32)st->textures.size();
    *Out = &st->s;
    return AZ_OK;
}

void az_scene_free(scene *Scene) {
    if (!Scene) return;
    delete reinterpret_cast<SceneStorage *>(Scene);
}

u32 az_scene_triangle_count(const scene *Scene) {
    if (!Scene) return 0;
    u32 n = 0;
    for (u32 i = 0; i < Scene->MeshCount; ++i) n += Scene->Meshes[i].TriangleCount;
    return n;
}

}  // extern "C"

extern "C" {
// Sizes of the ABI structures, for binding layers to check their mirrors:
// entity, mesh, material, light, az_texture, camera, scene, az_mesh_desc,
// az_sphere_desc, az_scene_params, az_render_opts, az_stats, az_ray, az_hit.
u32 az_abi_sizes(u32 *Out, u32 Capacity) {
    const u32 sizes[] = {(u32)sizeof(entity),        (u32)sizeof(mesh),           (u32)sizeof(material),
                         (u32)sizeof(light),         (u32)sizeof(az_texture),     (u32)sizeof(camera),
                         (u32)sizeof(scene),         (u32)sizeof(az_mesh_desc),   (u32)sizeof(az_sphere_desc),
                         (u32)sizeof(az_scene_params), (u32)sizeof(az_render_opts), (u32)sizeof(az_stats),
                         (u32)sizeof(az_ray),        (u32)sizeof(az_hit),         (u32)sizeof(az_tune_result)};
    const u32 n = (u32)(sizeof(sizes) / sizeof(sizes[0]));
    for (u32 i = 0; i < n && i < Capacity; ++i) Out[i] = sizes[i];
    return n;
}
}
