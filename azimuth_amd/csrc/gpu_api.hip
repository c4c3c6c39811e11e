// gpu_api.hip -- the C ABI of azimuth_gpu.h: contexts, scene upload, batch
// trace calls and device-memory helpers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "azimuth_gpu.h"
#include "host_parallel.h"
#include "bvh_build.h"
#include "device_scene.h"
#include "gpu_context.h"
#include "kernels.h"

using namespace az;

namespace {

void setError(az_gpu_ctx *c, const char *what, hipError_t e) {
    if (!c) return;
    char buf[256];
    snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
    c->lastError = buf;
}

#define AZ_HIP(ctx, call)                              \
    do {                                               \
        hipError_t _e = (call);                        \
        if (_e != hipSuccess) {                        \
            setError((ctx), #call, _e);                \
            return AZ_ERR_DEVICE;                      \
        }                                              \
    } while (0)

// Triangle entities of a mesh in the order of its BVH leaves (left first),
// which is the order of the mesh's triangle entity array.
void walkLeaves(const entity *e, std::vector<const entity *> &out) {
    std::vector<const entity *> st;
    st.push_back(e);
    while (!st.empty()) {
        const entity *x = st.back();
        st.pop_back();
        if (!x) continue;
        if (x->Type == EntityType_BVHNode) {
            st.push_back(x->Right);
            st.push_back(x->Left);
        } else {
            out.push_back(x);
        }
    }
}

// The leaves of the entity tree at e, left to right.  A tree expected to
// hold 65,536 leaves or more is split at its top levels (breadth first, in
// order) into up to 64 subtrees, walked on host threads and concatenated in
// order: the walk chases pointers through 288-byte entities.
void collectLeaves(const entity *e, std::vector<const entity *> &out, size_t expected = 0) {
    if (expected < 65536) {
        walkLeaves(e, out);
        return;
    }
    std::vector<const entity *> front{e};
    for (bool grew = true; grew && front.size() < 64;) {
        std::vector<const entity *> next;
        grew = false;
        for (const entity *x : front) {
            if (x && x->Type == EntityType_BVHNode) {
                next.push_back(x->Left);
                next.push_back(x->Right);
                grew = true;
            } else {
                next.push_back(x);
            }
        }
        front.swap(next);
    }
    std::vector<std::vector<const entity *>> parts(front.size());
    parallel_for(front.size(), [&](size_t i) { walkLeaves(front[i], parts[i]); }, 4);
    out.reserve(out.size() + expected);
    for (const std::vector<const entity *> &p : parts) out.insert(out.end(), p.begin(), p.end());
}

Affine toAffine(const mat4 &m) {
    Affine a;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 4; ++j) a.m[i][j] = m.E[i][j];
    return a;
}

BuildBox boxOfPoints(const vec3 *p, int n) {
    BuildBox b;
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = INFINITY;
        b.hi[k] = -INFINITY;
    }
    for (int i = 0; i < n; ++i) {
        const float v[3] = {p[i].x, p[i].y, p[i].z};
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = std::min(b.lo[k], v[k]);
            b.hi[k] = std::max(b.hi[k], v[k]);
        }
    }
    return b;
}

vec3 xform(const mat4 &m, float x, float y, float z) {
    vec3 r;
    r.x = m.E[0][0] * x + m.E[0][1] * y + m.E[0][2] * z + m.E[0][3];
    r.y = m.E[1][0] * x + m.E[1][1] * y + m.E[1][2] * z + m.E[1][3];
    r.z = m.E[2][0] * x + m.E[2][1] * y + m.E[2][2] * z + m.E[2][3];
    return r;
}

// World box of a sphere over the shutter interval: the transformed local box
// grown by the motion vector.
BuildBox sphereWorldBox(const entity &e) {
    vec3 c[16];
    int k = 0;
    for (int i = 0; i < 8; ++i) {
        float x = e.Center.x + ((i & 4) ? e.Radius : -e.Radius);
        float y = e.Center.y + ((i & 2) ? e.Radius : -e.Radius);
        float z = e.Center.z + ((i & 1) ? e.Radius : -e.Radius);
        vec3 w = xform(e.TransformationMatrix, x, y, z);
        c[k++] = w;
        vec3 m = w;
        m.x += e.MotionVector.x;
        m.y += e.MotionVector.y;
        m.z += e.MotionVector.z;
        c[k++] = m;
    }
    return boxOfPoints(c, 16);
}

}  // namespace

// The BVH build of one mesh or of the top level: the binary tree on the
// device (AZ_BUILD_DEVICE, or AZ_BUILD_AUTO from kDeviceBuildMin items on) or
// on the host, then the host's collapse into wide nodes.  Both give the same
// tree.
constexpr size_t kDeviceBuildMin = 65536;

static az_status buildWide(az_gpu_ctx *C, const std::vector<BuildBox> &boxes, const BuildParams &bp, uint32_t itemBase,
                           std::vector<Node4> &nodes, std::vector<uint32_t> &order, BuildStats *bs, uint32_t *root) {
    // The device builder gives the host's tree for its own bin count only.
    const bool device = bp.bins == kDeviceBuildBins &&
                        (C->buildMode == AZ_BUILD_DEVICE ||
                         (C->buildMode == AZ_BUILD_AUTO && boxes.size() >= kDeviceBuildMin));
    if (!device) {
        *root = BuildWideBvh(boxes, bp, itemBase, nodes, order, bs);
        return AZ_OK;
    }
    std::vector<BinNode> b2;
    const int e = BuildBinaryBvhDevice(C->stream, boxes, bp, b2, order);
    if (e != (int)hipSuccess) {
        C->lastError = std::string("device BVH build: ") + hipGetErrorString((hipError_t)e);
        return AZ_ERR_DEVICE;
    }
    *root = CollapseWideBvh(b2, 0, (uint32_t)boxes.size(), bp, itemBase, nodes, bs);
    return AZ_OK;
}

az_status az_upload_geometry(az_gpu_ctx *C, const scene *S, UploadInfo &info) {
    // ---- meshes: leaf-ordered triangles, one BLAS per mesh -----------------
    std::vector<Node4> nodes;
    std::vector<TriPacked> tris;
    std::vector<float> faceN;
    info.meshPrimBase.assign(S->MeshCount, 0);
    info.meshBlasRoot.assign(S->MeshCount, kEmpty);
    info.meshTriBase.assign(S->MeshCount, 0);
    info.meshTriEntities.assign(S->MeshCount, {});
    info.maxBlasStack = 0;
    uint32_t primBase = 0;
    BuildParams bp;
    bp.maxLeafItems = C->maxLeafItems;
    bp.width = C->bvhWidth;
    bp.nodeCost = C->sahNodeCost;
    bp.itemCost = C->sahItemCost;
    for (u32 m = 0; m < S->MeshCount; ++m) {
        const mesh &M = S->Meshes[m];
        std::vector<const entity *> &tl = info.meshTriEntities[m];
        collectLeaves(M.BVH, tl, M.TriangleCount);
        if (tl.size() != M.TriangleCount) {
            C->lastError = "mesh BVH leaf count differs from TriangleCount";
            return AZ_ERR_INVALID_ARGUMENT;
        }
        info.meshPrimBase[m] = primBase;
        std::vector<BuildBox> boxes(tl.size());
        parallel_for(tl.size(), [&](size_t i) {
            vec3 p[3] = {tl[i]->P0, tl[i]->P1, tl[i]->P2};
            boxes[i] = boxOfPoints(p, 3);
            PadBox(boxes[i]);
        });
        std::vector<uint32_t> order;
        BuildStats bs;
        uint32_t triBase = (uint32_t)tris.size();
        uint32_t root = 0;
        az_status bst = buildWide(C, boxes, bp, triBase, nodes, order, &bs, &root);
        if (bst != AZ_OK) return bst;
        info.meshBlasRoot[m] = root;
        info.meshTriBase[m] = triBase;
        info.maxBlasStack = std::max(info.maxBlasStack, bs.maxStack);
        info.leafCount += bs.leaves;
        tris.resize((size_t)triBase + order.size());
        faceN.resize(3 * ((size_t)triBase + order.size()));
        parallel_for(order.size(), [&](size_t i) {
            const entity *e = tl[order[i]];
            TriPacked t;
            t.p0x = e->P0.x;
            t.p0y = e->P0.y;
            t.p0z = e->P0.z;
            t.prim = primBase + order[i];
            t.a = e->P0.x - e->P1.x;
            t.d = e->P0.y - e->P1.y;
            t.g = e->P0.z - e->P1.z;
            t.b = e->P0.x - e->P2.x;
            t.e = e->P0.y - e->P2.y;
            t.h = e->P0.z - e->P2.z;
            t.pad0 = 0.0f;
            t.pad1 = 0.0f;
            tris[triBase + i] = t;
            faceN[3 * (triBase + i) + 0] = e->FaceNormal.x;
            faceN[3 * (triBase + i) + 1] = e->FaceNormal.y;
            faceN[3 * (triBase + i) + 2] = e->FaceNormal.z;
        });
        info.leafOrder.insert(info.leafOrder.end(), order.begin(), order.end());
        primBase += M.TriangleCount;
    }
    info.triTotal = primBase;

    // Meshes with a triangle in a face plane of their own box (all three
    // vertices on the box's min or max in one axis).
    std::vector<uint8_t> meshFaceOnBox(S->MeshCount, 0);
    parallel_for(S->MeshCount, [&](size_t m) {
        const std::vector<const entity *> &tl = info.meshTriEntities[m];
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        auto axis = [](const vec3 &v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); };
        for (const entity *e : tl)
            for (const vec3 *p : {&e->P0, &e->P1, &e->P2})
                for (int k = 0; k < 3; ++k) {
                    lo[k] = std::min(lo[k], axis(*p, k));
                    hi[k] = std::max(hi[k], axis(*p, k));
                }
        for (const entity *e : tl)
            for (int k = 0; k < 3 && !meshFaceOnBox[m]; ++k) {
                const float a = axis(e->P0, k), b = axis(e->P1, k), c = axis(e->P2, k);
                if (a == b && b == c && (a == lo[k] || a == hi[k])) meshFaceOnBox[m] = 1;
            }
    });

    // ---- top level ------------------------------------------------------------
    std::vector<const entity *> tops;
    collectLeaves(S->BVHRoot, tops);
    std::vector<TopItem> items0;
    std::vector<BuildBox> tboxes;
    for (const entity *e : tops) {
        TopItem it;
        memset(&it, 0, sizeof(it));
        it.entityId = e->UniqueEntityID;
        it.rank = (uint32_t)items0.size();   // `tops` lists the reference leaves left to right
        it.inv = toAffine(e->InvTransformationMatrix);
        it.fwd = toAffine(e->TransformationMatrix);
        it.motion[0] = e->MotionVector.x;
        it.motion[1] = e->MotionVector.y;
        it.motion[2] = e->MotionVector.z;
        BuildBox b;
        if (e->Type == EntityType_Sphere) {
            it.kind = kItemSphere;
            it.radius = e->Radius;
            it.center[0] = e->Center.x;
            it.center[1] = e->Center.y;
            it.center[2] = e->Center.z;
            b = sphereWorldBox(*e);
        } else if (e->Type == EntityType_MeshInstance) {
            it.kind = kItemInstance;
            u32 mi = (u32)(e->Mesh - S->Meshes);
            if (mi >= S->MeshCount) {
                C->lastError = "instance refers to a mesh outside scene->Meshes";
                return AZ_ERR_INVALID_ARGUMENT;
            }
            it.index = mi;
            it.blasRoot = info.meshBlasRoot[mi];
            if (meshFaceOnBox[mi]) it.kind |= kItemBoxTest;
            vec3 p[2] = {e->AABB.Min, e->AABB.Max};
            b = boxOfPoints(p, 2);
        } else {
            C->lastError = "unsupported top-level entity (bare triangle)";
            return AZ_ERR_UNSUPPORTED;
        }
        PadBox(b);
        items0.push_back(it);
        tboxes.push_back(b);
        info.topEntities.push_back(e);
    }
    if (items0.empty()) {
        C->lastError = "scene has no entities";
        return AZ_ERR_INVALID_ARGUMENT;
    }
    BuildParams tp;
    tp.maxLeafItems = 1;
    tp.width = C->bvhWidth;
    std::vector<uint32_t> torder;
    BuildStats ts;
    uint32_t tlasRoot = 0;
    az_status tst = buildWide(C, tboxes, tp, 0, nodes, torder, &ts, &tlasRoot);
    if (tst != AZ_OK) return tst;
    std::vector<TopItem> items(items0.size());
    std::vector<const entity *> topOrdered(items0.size());
    for (uint32_t i = 0; i < torder.size(); ++i) {
        items[i] = items0[torder[i]];
        topOrdered[i] = info.topEntities[torder[i]];
    }
    info.topEntities = topOrdered;
    info.maxTlasStack = ts.maxStack;
    for (int k = 0; k < 3; ++k) {
        info.boxMin[k] = INFINITY;
        info.boxMax[k] = -INFINITY;
    }
    for (const BuildBox &b : tboxes)
        for (int k = 0; k < 3; ++k) {
            info.boxMin[k] = std::min(info.boxMin[k], b.lo[k]);
            info.boxMax[k] = std::max(info.boxMax[k], b.hi[k]);
        }
    info.stackNeeded = ts.maxStack + 1 + info.maxBlasStack;

    // ---- upload ---------------------------------------------------------------
    // Traversal addresses nodes with 32-bit byte offsets.
    if (nodes.size() >= ((size_t)1 << 25)) {
        C->lastError = "BVH exceeds 2^25 nodes (4 GB of nodes)";
        return AZ_ERR_UNSUPPORTED;
    }
    C->geom.free();
    AZ_HIP(C, C->geom.nodes.alloc(nodes.size() * sizeof(Node4)));
    AZ_HIP(C, C->geom.tris.alloc(std::max<size_t>(1, tris.size()) * sizeof(TriPacked)));
    AZ_HIP(C, C->geom.faceNormals.alloc(std::max<size_t>(1, faceN.size()) * sizeof(float)));
    AZ_HIP(C, C->geom.items.alloc(items.size() * sizeof(TopItem)));
    // The reference tests an instance's own box before its mesh BVH
    // (RayAABBIntersection(Entity->AABB, Ray), main.cpp:245); the walk repeats
    // that test on these tight boxes, whatever its own (padded) boxes say.
    std::vector<float> refBoxes(6 * items.size());
    for (size_t i = 0; i < items.size(); ++i) {
        const entity *e = topOrdered[i];
        const aabb &bx = e->AABB;
        const float v[6] = {bx.Min.x, bx.Min.y, bx.Min.z, bx.Max.x, bx.Max.y, bx.Max.z};
        for (int k = 0; k < 6; ++k) refBoxes[6 * i + k] = e->Type == EntityType_MeshInstance ? v[k] : 0.0f;
    }
    AZ_HIP(C, C->geom.refBoxes.alloc(refBoxes.size() * sizeof(float)));
    AZ_HIP(C, hipMemcpyAsync(C->geom.refBoxes.ptr, refBoxes.data(), refBoxes.size() * sizeof(float),
                             hipMemcpyHostToDevice, C->stream));
    AZ_HIP(C, hipMemcpyAsync(C->geom.nodes.ptr, nodes.data(), nodes.size() * sizeof(Node4), hipMemcpyHostToDevice,
                             C->stream));
    if (!tris.empty()) {
        AZ_HIP(C, hipMemcpyAsync(C->geom.tris.ptr, tris.data(), tris.size() * sizeof(TriPacked),
                                 hipMemcpyHostToDevice, C->stream));
        AZ_HIP(C, hipMemcpyAsync(C->geom.faceNormals.ptr, faceN.data(), faceN.size() * sizeof(float),
                                 hipMemcpyHostToDevice, C->stream));
    }
    AZ_HIP(C, hipMemcpyAsync(C->geom.items.ptr, items.data(), items.size() * sizeof(TopItem), hipMemcpyHostToDevice,
                             C->stream));
    AZ_HIP(C, hipStreamSynchronize(C->stream));
    DeviceGeometry &G = C->dg;
    G.nodes = (const Node4 *)C->geom.nodes.ptr;
    G.tris = (const TriPacked *)C->geom.tris.ptr;
    G.faceNormals = (const float *)C->geom.faceNormals.ptr;
    G.items = (const TopItem *)C->geom.items.ptr;
    G.refBoxes = (const float *)C->geom.refBoxes.ptr;
    G.tlasRoot = tlasRoot;
    G.triTotal = info.triTotal;
    G.itemCount = (uint32_t)items.size();
    C->hostItems = items;
    C->stats.NodeCount = (u32)nodes.size();
    C->stats.LeafCount = info.leafCount + ts.leaves;
    C->stats.TriangleCount = info.triTotal;
    C->stats.SphereCount = 0;
    C->stats.InstanceCount = 0;
    for (const TopItem &it : items) {
        if ((it.kind & 0xFFu) == kItemSphere) C->stats.SphereCount++;
        else C->stats.InstanceCount++;
    }
    C->stats.NodeBytes = sizeof(Node4);
    C->stats.TriangleBytes = sizeof(TriPacked);
    C->stats.BvhWidth = C->bvhWidth;
    // Stack capacity of the kernels: the LDS ring plus the scratch spill
    // (the renderer's kernels and the default traversal use a 16-entry ring;
    // a shallower tuned ring falls back to the default here).
    if (info.stackNeeded > (uint32_t)(16 + kSpill)) {
        C->lastError = "BVH needs a deeper traversal stack than the kernels hold";
        return AZ_ERR_UNSUPPORTED;
    }
    if ((uint32_t)(C->launch.ring + kSpill) < info.stackNeeded) C->launch = TraceLaunch{};
    C->stats.StackDepth = info.stackNeeded;
    C->stats.BlockSize = (u32)C->launch.block;
    return AZ_OK;
}

extern "C" {

az_status az_gpu_create(az_gpu_ctx **Out, int dev) {
    if (!Out) return AZ_ERR_INVALID_ARGUMENT;
    *Out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || dev < 0 || dev >= count) return AZ_ERR_DEVICE;
    az_gpu_ctx *c = new (std::nothrow) az_gpu_ctx();
    if (!c) return AZ_ERR_OUT_OF_MEMORY;
    c->device = dev;
    if (hipSetDevice(dev) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->tev0) != hipSuccess || hipEventCreate(&c->tev1) != hipSuccess) {
        delete c;
        return AZ_ERR_DEVICE;
    }
    if (hipDeviceGetAttribute(&c->cuCount, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) c->cuCount = 256;
    c->cuCount = std::max(1, c->cuCount);
    memset(&c->stats, 0, sizeof(c->stats));
    void *ws = nullptr;
    (void)az::noalloc_workspace(c, &ws);   // set up here; retried by the entry point if this failed
    *Out = c;
    return AZ_OK;
}

void az_gpu_destroy(az_gpu_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete c;
}

const char *az_gpu_last_error(const az_gpu_ctx *c) { return c ? c->lastError.c_str() : "no context"; }

az_status az_gpu_upload_scene(az_gpu_ctx *C, const scene *S, const epsilon_values *E) {
    if (!C || !S || !E || !S->BVHRoot) return AZ_ERR_INVALID_ARGUMENT;
    AZ_HIP(C, hipSetDevice(C->device));
    C->hasScene = false;
    UploadInfo info;
    az_status st = az_upload_geometry(C, S, info);
    if (st != AZ_OK) return st;
    C->eps = *E;
    st = az_upload_shading(C, S, info);
    if (st != AZ_OK) return st;
    C->hasScene = true;
    return AZ_OK;
}

static az_status traceCommon(az_gpu_ctx *C, const az_ray *rays, u32 n, az_hit *hits, b32 cull, bool any) {
    if (!C) return AZ_ERR_INVALID_ARGUMENT;
    if (!C->hasScene) return AZ_ERR_NO_SCENE;
    if (n && (!rays || !hits)) return AZ_ERR_INVALID_ARGUMENT;
    AZ_HIP(C, hipSetDevice(C->device));
    unsigned long long *cnt = nullptr;
    if (C->counting) {
        AZ_HIP(C, C->counters.alloc(4 * sizeof(unsigned long long)));
        AZ_HIP(C, hipMemsetAsync(C->counters.ptr, 0, 4 * sizeof(unsigned long long), C->stream));
        cnt = (unsigned long long *)C->counters.ptr;
    }
    hipEvent_t e0 = C->ev0, e1 = C->ev1;
    if (C->recordLaunches) {
        if (C->launchEvents.size() < 2 * (C->launchUsed + 1)) {
            hipEvent_t a, b;
            AZ_HIP(C, hipEventCreate(&a));
            AZ_HIP(C, hipEventCreate(&b));
            C->launchEvents.push_back(a);
            C->launchEvents.push_back(b);
        }
        e0 = C->launchEvents[2 * C->launchUsed];
        e1 = C->launchEvents[2 * C->launchUsed + 1];
        C->launchUsed++;
    }
    AZ_HIP(C, hipEventRecord(e0, C->stream));
    AZ_HIP(C, launch_trace(C->launch, any, C->counting, C->dg, rays, n, hits, C->eps.Intersection, cull ? 1 : 0, cnt,
                           C->cuCount, C->stream));
    AZ_HIP(C, hipEventRecord(e1, C->stream));
    C->lastStart = e0;   // LastKernelMs reads the latest pair
    C->lastStop = e1;
    C->pendingTiming = true;
    if (C->counting) {
        unsigned long long h[4] = {0, 0, 0, 0};
        AZ_HIP(C, hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, C->stream));
        AZ_HIP(C, hipStreamSynchronize(C->stream));
        C->stats.NodesVisited = h[0];
        C->stats.TrianglesTested = h[1];
        C->stats.PrimitivesTested = h[2];
    }
    return AZ_OK;
}

az_status az_gpu_trace_closest(az_gpu_ctx *C, const az_ray *rays, u32 n, az_hit *hits, b32 cull) {
    return traceCommon(C, rays, n, hits, cull, false);
}

az_status az_gpu_trace_any(az_gpu_ctx *C, const az_ray *rays, u32 n, az_hit *hits) {
    return traceCommon(C, rays, n, hits, 0, true);
}

az_status az_gpu_get_stats(az_gpu_ctx *C, az_stats *S) {
    if (!C || !S) return AZ_ERR_INVALID_ARGUMENT;
    if (C->pendingTiming) {
        AZ_HIP(C, hipEventSynchronize(C->lastStop));
        float ms = 0.0f;
        AZ_HIP(C, hipEventElapsedTime(&ms, C->lastStart, C->lastStop));
        C->stats.LastKernelMs = ms;
        C->pendingTiming = false;
    }
    *S = C->stats;
    return AZ_OK;
}

az_status az_gpu_set_counting(az_gpu_ctx *C, b32 on) {
    if (!C) return AZ_ERR_INVALID_ARGUMENT;
    C->counting = on != 0;
    return AZ_OK;
}

az_status az_gpu_set_launch(az_gpu_ctx *C, u32 block, u32 ring, s32 waves, s32 loop) {
    if (!C) return AZ_ERR_INVALID_ARGUMENT;
    int b = block ? (int)block : C->launch.block;
    int r = ring ? (int)ring : C->launch.ring;
    int w = waves >= 0 ? (int)waves : C->launch.waves;
    int f = loop >= 0 ? (int)loop : C->launch.loop;
    if (!trace_variant_exists(r, b, w, f)) return AZ_ERR_INVALID_ARGUMENT;
    if (C->hasScene && (uint32_t)(r + kSpill) < C->stats.StackDepth) {
        C->lastError = "LDS stack ring too shallow for the uploaded BVH";
        return AZ_ERR_INVALID_ARGUMENT;
    }
    C->launch.block = b;
    C->launch.ring = r;
    C->launch.waves = w;
    C->launch.loop = f;
    C->stats.BlockSize = (u32)b;
    return AZ_OK;
}

az_status az_gpu_set_build(az_gpu_ctx *C, u32 maxLeafItems, f32 nodeCost, f32 itemCost) {
    if (!C || maxLeafItems < 1 || maxLeafItems > kMaxLeafItems || !(nodeCost > 0.0f) || !(itemCost > 0.0f))
        return AZ_ERR_INVALID_ARGUMENT;
    C->maxLeafItems = maxLeafItems;
    C->sahNodeCost = nodeCost;
    C->sahItemCost = itemCost;
    return AZ_OK;
}

az_status az_gpu_set_build_mode(az_gpu_ctx *C, u32 mode) {
    if (!C || mode > AZ_BUILD_DEVICE) return AZ_ERR_INVALID_ARGUMENT;
    C->buildMode = mode;
    return AZ_OK;
}

az_status az_gpu_set_bvh_width(az_gpu_ctx *C, u32 width) {
    if (!C || width < 2 || width > 4) return AZ_ERR_INVALID_ARGUMENT;
    C->bvhWidth = width;
    return AZ_OK;
}

az_status az_gpu_autotune_build(az_gpu_ctx *C, const scene *S, const epsilon_values *E, const az_ray *rays, u32 n,
                                u32 repeats, az_tune_result *out) {
    if (!C || !S || !E || !rays || n == 0) return AZ_ERR_INVALID_ARGUMENT;
    AZ_HIP(C, hipSetDevice(C->device));
    repeats = std::max<u32>(1u, repeats);
    const u32 widths[3] = {2u, 3u, 4u}, leaves[3] = {2u, 4u, 8u};
    const u32 keepWidth = C->bvhWidth, keepLeaf = C->maxLeafItems;
    DeviceBuffer hits;
    AZ_HIP(C, hits.alloc((size_t)n * sizeof(az_hit)));
    hipEvent_t a = nullptr, b = nullptr;
    AZ_HIP(C, hipEventCreate(&a));
    az_status st = AZ_OK;
    if (hipEventCreate(&b) != hipSuccess) st = AZ_ERR_DEVICE;
    az_tune_result res;
    memset(&res, 0, sizeof(res));
    double best = 1e300;
    u32 bestW = keepWidth, bestL = keepLeaf;
    for (u32 wi = 0; wi < 3 && st == AZ_OK; ++wi)
        for (u32 li = 0; li < 3 && st == AZ_OK; ++li) {
            C->bvhWidth = widths[wi];
            C->maxLeafItems = leaves[li];
            st = az_gpu_upload_scene(C, S, E);
            if (st != AZ_OK) break;
            const TraceLaunch L = C->launch;
            hipError_t e = launch_trace(L, false, false, C->dg, rays, n, (az_hit *)hits.ptr, C->eps.Intersection, 0,
                                        nullptr, C->cuCount, C->stream);
            if (e == hipSuccess) e = hipEventRecord(a, C->stream);
            for (u32 r = 0; r < repeats && e == hipSuccess; ++r)
                e = launch_trace(L, false, false, C->dg, rays, n, (az_hit *)hits.ptr, C->eps.Intersection, 0, nullptr,
                                 C->cuCount, C->stream);
            if (e == hipSuccess) e = hipEventRecord(b, C->stream);
            if (e == hipSuccess) e = hipEventSynchronize(b);
            float ms = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
            if (e != hipSuccess) {
                C->lastError = std::string("build autotune: ") + hipGetErrorString(e);
                st = AZ_ERR_DEVICE;
                break;
            }
            az_tune_entry &t = res.Entries[res.Count++];
            t.Block = (u32)L.block;
            t.LdsStackEntries = (u32)L.ring;
            t.MinWaves = L.waves;
            t.LoopForm = (u32)L.loop;
            t.BvhWidth = widths[wi];
            t.MaxLeafItems = leaves[li];
            t.Milliseconds = (double)ms / repeats;
            if (t.Milliseconds < best) {
                best = t.Milliseconds;
                bestW = widths[wi];
                bestL = leaves[li];
                res.Best = res.Count - 1;
            }
        }
    (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    // Leave the scene uploaded with the fastest build (or, after a failure,
    // with the settings the caller had).
    C->bvhWidth = st == AZ_OK ? bestW : keepWidth;
    C->maxLeafItems = st == AZ_OK ? bestL : keepLeaf;
    az_status up = az_gpu_upload_scene(C, S, E);
    if (st != AZ_OK) return st;
    if (up != AZ_OK) return up;
    if (out) *out = res;
    return AZ_OK;
}

az_status az_gpu_autotune_trace(az_gpu_ctx *C, const az_ray *rays, u32 n, u32 repeats, az_tune_result *out) {
    if (!C || !rays || n == 0) return AZ_ERR_INVALID_ARGUMENT;
    if (!C->hasScene) return AZ_ERR_NO_SCENE;
    AZ_HIP(C, hipSetDevice(C->device));
    repeats = std::max<u32>(1u, repeats);
    TraceLaunch cands[32];
    const int k = trace_variants(cands, 32);
    DeviceBuffer hits;
    AZ_HIP(C, hits.alloc((size_t)n * sizeof(az_hit)));
    hipEvent_t a = nullptr, b = nullptr;
    AZ_HIP(C, hipEventCreate(&a));
    az_status st = AZ_OK;
    if (hipEventCreate(&b) != hipSuccess) st = AZ_ERR_DEVICE;
    double best = 1e300;
    int bestIdx = -1;
    az_tune_result res;
    memset(&res, 0, sizeof(res));
    for (int i = 0; i < k && st == AZ_OK; ++i) {
        // A ring whose stack (ring plus scratch spill) is too shallow for
        // this scene's BVH is not a candidate.
        if ((uint32_t)(cands[i].ring + kSpill) < C->stats.StackDepth) continue;
        hipError_t e = launch_trace(cands[i], false, false, C->dg, rays, n, (az_hit *)hits.ptr, C->eps.Intersection, 0,
                                    nullptr, C->cuCount, C->stream);
        if (e == hipSuccess) e = hipEventRecord(a, C->stream);
        for (u32 r = 0; r < repeats && e == hipSuccess; ++r)
            e = launch_trace(cands[i], false, false, C->dg, rays, n, (az_hit *)hits.ptr, C->eps.Intersection, 0,
                             nullptr, C->cuCount, C->stream);
        if (e == hipSuccess) e = hipEventRecord(b, C->stream);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        if (e != hipSuccess) {
            C->lastError = std::string("autotune: ") + hipGetErrorString(e);
            st = AZ_ERR_DEVICE;
            break;
        }
        const double per = (double)ms / repeats;
        if (res.Count >= 32) continue;
        az_tune_entry &t = res.Entries[res.Count++];
        t.Block = (u32)cands[i].block;
        t.LdsStackEntries = (u32)cands[i].ring;
        t.MinWaves = cands[i].waves;
        t.ChunkSamples = 0;
        t.LoopForm = (u32)cands[i].loop;
        t.BvhWidth = C->bvhWidth;
        t.MaxLeafItems = C->maxLeafItems;
        t.Milliseconds = per;
        if (per < best) {
            best = per;
            bestIdx = i;
            res.Best = res.Count - 1;
        }
    }
    (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (st != AZ_OK) return st;
    if (bestIdx < 0) return AZ_ERR_UNSUPPORTED;
    C->launch = cands[bestIdx];
    C->stats.BlockSize = (u32)cands[bestIdx].block;
    if (out) *out = res;
    return AZ_OK;
}

az_status az_gpu_malloc(az_gpu_ctx *C, void **p, u64 bytes) {
    if (!C || !p) return AZ_ERR_INVALID_ARGUMENT;
    AZ_HIP(C, hipSetDevice(C->device));
    AZ_HIP(C, hipMalloc(p, bytes ? bytes : 1));
    return AZ_OK;
}

az_status az_gpu_free(az_gpu_ctx *C, void *p) {
    if (!C) return AZ_ERR_INVALID_ARGUMENT;
    if (!p) return AZ_OK;
    AZ_HIP(C, hipStreamSynchronize(C->stream));
    AZ_HIP(C, hipFree(p));
    return AZ_OK;
}

az_status az_gpu_upload(az_gpu_ctx *C, void *dst, const void *src, u64 bytes) {
    if (!C || (bytes && (!dst || !src))) return AZ_ERR_INVALID_ARGUMENT;
    AZ_HIP(C, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, C->stream));
    AZ_HIP(C, hipStreamSynchronize(C->stream));
    return AZ_OK;
}

az_status az_gpu_download(az_gpu_ctx *C, void *dst, const void *src, u64 bytes) {
    if (!C || (bytes && (!dst || !src))) return AZ_ERR_INVALID_ARGUMENT;
    AZ_HIP(C, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, C->stream));
    AZ_HIP(C, hipStreamSynchronize(C->stream));
    return AZ_OK;
}

az_status az_gpu_memset(az_gpu_ctx *C, void *dst, int v, u64 bytes) {
    if (!C || (bytes && !dst)) return AZ_ERR_INVALID_ARGUMENT;
    AZ_HIP(C, hipMemsetAsync(dst, v, bytes, C->stream));
    return AZ_OK;
}

az_status az_gpu_synchronize(az_gpu_ctx *C) {
    if (!C) return AZ_ERR_INVALID_ARGUMENT;
    AZ_HIP(C, hipStreamSynchronize(C->stream));
    return AZ_OK;
}

az_status az_gpu_device_copy(az_gpu_ctx *C, void *dst, const void *src, u64 bytes) {
    if (!C || (bytes && (!dst || !src)) || (bytes % 16) != 0) return AZ_ERR_INVALID_ARGUMENT;
    AZ_HIP(C, launch_copy16(src, dst, bytes, C->stream));
    return AZ_OK;
}

az_status az_gpu_record_launch_times(az_gpu_ctx *C, b32 on) {
    if (!C) return AZ_ERR_INVALID_ARGUMENT;
    C->recordLaunches = on != 0;
    C->launchUsed = 0;
    return AZ_OK;
}

az_status az_gpu_collect_launch_times(az_gpu_ctx *C, f64 *ms, u32 cap, u32 *count) {
    if (!C || !count || (cap && !ms)) return AZ_ERR_INVALID_ARGUMENT;
    AZ_HIP(C, hipStreamSynchronize(C->stream));
    u32 n = 0;
    for (size_t i = 0; i < C->launchUsed; ++i) {
        float f = 0.0f;
        AZ_HIP(C, hipEventElapsedTime(&f, C->launchEvents[2 * i], C->launchEvents[2 * i + 1]));
        if (n < cap) ms[n] = f;
        ++n;
    }
    *count = n;
    C->launchUsed = 0;
    return AZ_OK;
}

az_status az_gpu_timer_start(az_gpu_ctx *C) {
    if (!C) return AZ_ERR_INVALID_ARGUMENT;
    AZ_HIP(C, hipEventRecord(C->tev0, C->stream));
    return AZ_OK;
}

az_status az_gpu_timer_stop(az_gpu_ctx *C, f64 *ms) {
    if (!C || !ms) return AZ_ERR_INVALID_ARGUMENT;
    AZ_HIP(C, hipEventRecord(C->tev1, C->stream));
    AZ_HIP(C, hipEventSynchronize(C->tev1));
    float f = 0.0f;
    AZ_HIP(C, hipEventElapsedTime(&f, C->tev0, C->tev1));
    *ms = f;
    return AZ_OK;
}

}  // extern "C"
