// device_scene.h -- layout of the scene in HBM, shared by the host upload code
// and the HIP kernels.
//
// Geometry is two-level, as in the reference (a top-level BVH over spheres and
// mesh instances, one BVH per mesh: xml.cpp:3751,3915).  Each level is a
// 4-wide BVH of 128-byte nodes (one L2 line); triangles are stored in leaf
// order, 48 bytes each.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define AZ_HD __host__ __device__
#else
#define AZ_HD
#endif

namespace az {

// Child reference inside a node or on the traversal stack.
//   bit 31 clear : internal node index
//   bit 31 set   : leaf; bits 0..26 first item, bits 27..30 item count - 1
//   0xFFFFFFFF   : empty slot / instance-exit marker on the stack
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kMaxLeafItems = 16;
constexpr uint32_t kMaxLeafFirst = (1u << 27) - 1;
// Traversal stack entries per lane kept in scratch below the LDS ring.
constexpr int kSpill = 80;
// Rays a lane of the renderer's traversal walks per launch on average: a
// wave owns 64 x this many consecutive rays of a level's queue and refills
// its lanes from them (launches too small to fill the device twice take
// fewer).  8 against 4: fewer waves, each staging its queue's segment
// offsets once for twice the rays; the 1080p interior render 498 -> 488 ms,
// 16 gave 490.  Batch queries keep 4 (kTraceRaysPerLane, kernels.h).
constexpr uint32_t kRaysPerLane = 8;

inline AZ_HD bool ref_is_leaf(uint32_t r) { return (r & kLeafBit) != 0u; }
inline AZ_HD uint32_t leaf_first(uint32_t r) { return r & 0x07FFFFFFu; }
inline AZ_HD uint32_t leaf_count(uint32_t r) { return ((r >> 27) & 0xFu) + 1u; }
inline AZ_HD uint32_t make_leaf(uint32_t first, uint32_t count) {
    return kLeafBit | ((count - 1u) << 27) | first;
}

// 128-byte 4-wide node.  Child boxes are stored per axis so that one 16-byte
// load brings the same bound of all four children.  Empty children have
// lo = +inf, hi = -inf and never hit.
struct alignas(128) Node4 {
    float lox[4], hix[4];
    float loy[4], hiy[4];
    float loz[4], hiz[4];
    uint32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "Node4 must be one 128-byte line");

// Triangle as fetched by the intersection test: P0 and the two edge
// differences P0-P1, P0-P2 that the reference's Cramer's rule uses
// (main.cpp:154-162), plus the primitive id.  48 bytes.
struct alignas(16) TriPacked {
    float p0x, p0y, p0z;
    uint32_t prim;
    float a, d, g, pad0;   // P0 - P1 (x, y, z)
    float b, e, h, pad1;   // P0 - P2 (x, y, z)
};
static_assert(sizeof(TriPacked) == 48, "TriPacked is 48 bytes");

// Row-major 3x4 affine part of a mat4.
struct Affine {
    float m[3][4];
};

// Top-level item: a mesh instance (BLAS reference) or a sphere.
enum : uint32_t { kItemInstance = 0u, kItemSphere = 1u };
// Set in an instance's kind when its mesh has triangles lying in a face plane
// of its own box (quads, walls, boxes): hits there lie on the box boundary,
// where the reference's box test on the instance (main.cpp:245) depends on
// rounding, so the walk repeats that test on such an instance's hits.
constexpr uint32_t kItemBoxTest = 0x100u;

struct alignas(16) TopItem {
    uint32_t kind;
    uint32_t blasRoot;     // instance: root node index of the mesh BVH
    uint32_t entityId;     // UniqueEntityID of the reference entity
    uint32_t index;        // instance: mesh index; sphere: sphere index
    Affine inv;            // InvTransformationMatrix
    Affine fwd;            // TransformationMatrix
    float motion[3];
    float radius;          // sphere only
    float center[3];       // sphere only (object space)
    uint32_t rank;         // position of the entity among the reference BVH's leaves (left to right)
};

struct DeviceGeometry {
    const Node4 *nodes;         // all BVH levels in one array
    const TriPacked *tris;      // leaf-ordered triangles of all meshes
    const float *faceNormals;   // per leaf-ordered triangle, xyz (for culling)
    const TopItem *items;       // top-level items in TLAS leaf order
    const float *refBoxes;      // per item: the reference entity's AABB (min xyz, max xyz), instances only
    uint32_t tlasRoot;
    uint32_t triTotal;          // sphere primitive ids start here
    uint32_t itemCount;
    uint32_t pad;
};

}  // namespace az
