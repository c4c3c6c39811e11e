// bvh_build.h -- host builder of the device BVH: binned-SAH binary tree,
// collapsed into 4-wide 128-byte nodes.
#pragma once

#include <stdint.h>
#include <vector>

#include "device_scene.h"

namespace az {

struct BuildBox {
    float lo[3], hi[3];
};

struct BuildParams {
    uint32_t maxLeafItems = 4;     // leaves never hold more items than this
    uint32_t width = 4;            // children per node (2..4; unused slots stay empty)
    uint32_t bins = 32;            // SAH bins per axis
    float nodeCost = 1.0f;         // SAH cost of visiting one 4-wide node
    float itemCost = 1.0f;         // SAH cost of testing one item
    bool serialCollapse = false;   // collapse into wide nodes on one thread (for checks)
};

struct BuildStats {
    uint32_t nodes = 0;
    uint32_t leaves = 0;
    uint32_t maxDepth = 0;         // 4-wide levels on the deepest path
    uint32_t maxStack = 0;         // stack entries a front-to-back walk can need
};

// A node of the binary binned-SAH tree: its box, its children (internal) or
// its item range [first, first + count) of the order (leaf).
struct BinNode {
    float lo[3], hi[3];
    uint32_t left = 0, right = 0;
    uint32_t first = 0, count = 0;
    uint32_t leaf = 0;
};

// The binary tree on the host: node 0 is the root; `order` receives the item
// permutation.  Each split partitions its range stably (items keep their
// relative order on each side), so the device builder (bvh_gpu.hip) makes
// the same tree.
void BuildBinaryBvh(const std::vector<BuildBox> &boxes, const BuildParams &params, std::vector<BinNode> &b2,
                    std::vector<uint32_t> &order);

// Collapses a binary tree (root `root`) into `width`-wide nodes appended to
// `nodes`; leaf references point at itemBase + first.  Returns the root node
// index.
uint32_t CollapseWideBvh(const std::vector<BinNode> &b2, uint32_t root, uint32_t n, const BuildParams &params,
                         uint32_t itemBase, std::vector<Node4> &nodes, BuildStats *stats);

// The device builder bins with this many bins per axis only; BuildParams
// with other bins are built on the host.
constexpr uint32_t kDeviceBuildBins = 32;

// Builds the binary tree on the device (hipSuccess, or the error of the
// first failing call, hipErrorInvalidValue when params.bins is not
// kDeviceBuildBins; all device memory is freed either way).
int BuildBinaryBvhDevice(void *stream, const std::vector<BuildBox> &boxes, const BuildParams &params,
                         std::vector<BinNode> &b2, std::vector<uint32_t> &order);

// Builds a 4-wide BVH over `boxes` and appends its nodes to `nodes`.  `order`
// receives the item permutation: leaf references index into it, so item
// order[i] must be stored at position itemBase + i.  Returns the root node
// index.  The root is always an internal node (one leaf child when there are
// few items).  Boxes are used as given; callers pad them.
uint32_t BuildWideBvh(const std::vector<BuildBox> &boxes, const BuildParams &params,
                      uint32_t itemBase, std::vector<Node4> &nodes, std::vector<uint32_t> &order,
                      BuildStats *stats);

// Pads a box outward so that a point the reference triangle test accepts
// (barycentrics down to -1e-6) and the rounding of the slab test stay inside.
void PadBox(BuildBox &b);

}  // namespace az
