// bvh_build.cpp -- binned-SAH binary BVH collapsed into 4-wide nodes.
//
// The device BVH is this project's own structure and is not the reference's
// entity tree: the reference BVH (bvh.cpp:1-328) splits at the box midpoint
// and is walked without pruning (main.cpp:282-319).  Any tree whose boxes
// contain every point the reference triangle and sphere tests can accept gives
// the same closest hit when walked front to back with distance pruning; the
// padding in PadBox keeps that containment under float rounding.
#include "bvh_build.h"
#include "host_parallel.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

namespace az {

namespace {

struct Aabb {
    float lo[3], hi[3];
    void reset() {
        for (int k = 0; k < 3; ++k) {
            lo[k] = FLT_MAX;
            hi[k] = -FLT_MAX;
        }
    }
    void grow(const Aabb &b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void growPoint(const float *p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    float area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0.0f || dy < 0.0f || dz < 0.0f) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Bin {
    Aabb box;
    uint32_t count;
};

}  // namespace

void PadBox(BuildBox &b) {
    float ext = 0.0f, mag = 0.0f;
    for (int k = 0; k < 3; ++k) {
        ext = std::max(ext, b.hi[k] - b.lo[k]);
        mag = std::max(mag, std::max(std::fabs(b.lo[k]), std::fabs(b.hi[k])));
    }
    float pad = 4e-6f * ext + 2e-6f * mag + 1e-30f;
    for (int k = 0; k < 3; ++k) {
        b.lo[k] -= pad;
        b.hi[k] += pad;
    }
}

static float binArea(const BinNode &b) {
    float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    if (dx < 0.0f || dy < 0.0f || dz < 0.0f) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

void BuildBinaryBvh(const std::vector<BuildBox> &boxes, const BuildParams &P, std::vector<BinNode> &b2,
                    std::vector<uint32_t> &order) {
    const uint32_t n = (uint32_t)boxes.size();
    const uint32_t maxLeaf = std::max(1u, std::min(P.maxLeafItems, kMaxLeafItems));
    order.resize(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::vector<Aabb> ib(n);
    std::vector<float> cen(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            ib[i].lo[k] = boxes[i].lo[k];
            ib[i].hi[k] = boxes[i].hi[k];
            cen[3 * (size_t)i + k] = 0.5f * (boxes[i].lo[k] + boxes[i].hi[k]);
        }
    }
    b2.clear();
    // ---- binary build -------------------------------------------------------
    b2.reserve(n > 0 ? 2 * (size_t)n : 1);
    struct Work { uint32_t node, begin, end; };
    std::vector<Work> stack;
    b2.emplace_back();
    if (n > 0) stack.push_back({0, 0, n});
    const uint32_t nb = std::max(4u, std::min(P.bins, 64u));
    std::vector<Bin> bins(nb);
    std::vector<Aabb> rightBox(nb);
    std::vector<uint32_t> rightCount(nb);
    while (!stack.empty()) {
        Work w = stack.back();
        stack.pop_back();
        Aabb box, cb;
        box.reset();
        cb.reset();
        for (uint32_t i = w.begin; i < w.end; ++i) {
            box.grow(ib[order[i]]);
            cb.growPoint(&cen[3 * (size_t)order[i]]);
        }
        for (int k = 0; k < 3; ++k) {
            b2[w.node].lo[k] = box.lo[k];
            b2[w.node].hi[k] = box.hi[k];
        }
        const uint32_t cnt = w.end - w.begin;
        float bestCost = FLT_MAX;
        int bestAxis = -1;
        uint32_t bestSplit = 0;
        if (cnt > 1) {
            for (int ax = 0; ax < 3; ++ax) {
                float ext = cb.hi[ax] - cb.lo[ax];
                if (!(ext > 0.0f)) continue;
                float scale = (float)nb / ext;
                for (uint32_t b = 0; b < nb; ++b) {
                    bins[b].box.reset();
                    bins[b].count = 0;
                }
                for (uint32_t i = w.begin; i < w.end; ++i) {
                    uint32_t it = order[i];
                    int b = (int)((cen[3 * (size_t)it + ax] - cb.lo[ax]) * scale);
                    b = std::min(std::max(b, 0), (int)nb - 1);
                    bins[b].box.grow(ib[it]);
                    bins[b].count++;
                }
                Aabb acc;
                acc.reset();
                uint32_t c = 0;
                for (int b = (int)nb - 1; b > 0; --b) {
                    acc.grow(bins[b].box);
                    c += bins[b].count;
                    rightBox[b] = acc;
                    rightCount[b] = c;
                }
                acc.reset();
                c = 0;
                for (uint32_t b = 0; b + 1 < nb; ++b) {
                    acc.grow(bins[b].box);
                    c += bins[b].count;
                    uint32_t rc = rightCount[b + 1];
                    if (c == 0 || rc == 0) continue;
                    float cost = acc.area() * (float)c + rightBox[b + 1].area() * (float)rc;
                    if (cost < bestCost) {
                        bestCost = cost;
                        bestAxis = ax;
                        bestSplit = b + 1;
                    }
                }
            }
        }
        float area = box.area();
        float splitCost = bestAxis >= 0 && area > 0.0f
                              ? 0.5f * P.nodeCost + P.itemCost * bestCost / area
                              : FLT_MAX;
        float leafCost = P.itemCost * (float)cnt;
        bool makeLeaf = cnt <= 1 || (cnt <= maxLeaf && leafCost <= splitCost);
        if (makeLeaf) {
            b2[w.node].leaf = 1u;
            b2[w.node].first = w.begin;
            b2[w.node].count = cnt;
            continue;
        }
        uint32_t mid;
        if (bestAxis >= 0) {
            float ext = cb.hi[bestAxis] - cb.lo[bestAxis];
            float scale = (float)nb / ext;
            uint32_t *b0 = order.data() + w.begin, *b1 = order.data() + w.end;
            uint32_t *m = std::stable_partition(b0, b1, [&](uint32_t it) {
                int b = (int)((cen[3 * (size_t)it + bestAxis] - cb.lo[bestAxis]) * scale);
                b = std::min(std::max(b, 0), (int)nb - 1);
                return (uint32_t)b < bestSplit;
            });
            mid = (uint32_t)(m - order.data());
            if (mid == w.begin || mid == w.end) mid = w.begin + cnt / 2;
        } else {
            mid = w.begin + cnt / 2;   // all centroids coincide
        }
        uint32_t l = (uint32_t)b2.size();
        b2.emplace_back();
        b2.emplace_back();
        b2[w.node].left = l;
        b2[w.node].right = l + 1;
        stack.push_back({l + 1, mid, w.end});
        stack.push_back({l, w.begin, mid});
    }
}

namespace {

Node4 emptyNode() {
    Node4 z;
    for (int k = 0; k < 4; ++k) {
        z.lox[k] = z.loy[k] = z.loz[k] = INFINITY;
        z.hix[k] = z.hiy[k] = z.hiz[k] = -INFINITY;
        z.child[k] = kEmpty;
        z.pad[k] = 0;
    }
    return z;
}

// The children of the wide node made from binary node b: a leaf's node
// holds the leaf itself; otherwise its two children, then repeatedly the
// two children of the internal child with the largest area in its place,
// up to `width`.
uint32_t gatherKids(const std::vector<BinNode> &b2, uint32_t b, uint32_t width, uint32_t kids[4]) {
    const BinNode &src = b2[b];
    if (src.leaf) {
        kids[0] = b;
        return 1;
    }
    uint32_t k = 0;
    kids[k++] = src.left;
    kids[k++] = src.right;
    while (k < width) {
        int best = -1;
        float bestArea = -1.0f;
        for (uint32_t c = 0; c < k; ++c) {
            const BinNode &cn = b2[kids[c]];
            if (cn.leaf) continue;
            const float a = binArea(cn);
            if (a > bestArea) {
                bestArea = a;
                best = (int)c;
            }
        }
        if (best < 0) break;
        const uint32_t open = kids[best];
        kids[best] = b2[open].left;
        kids[k++] = b2[open].right;
    }
    return k;
}

// Boxes and leaf references of a wide node's children (internal children
// are linked by the caller).
void fillNode(const std::vector<BinNode> &b2, const uint32_t *kids, uint32_t k, uint32_t itemBase, Node4 &dst,
              BuildStats &st) {
    for (uint32_t c = 0; c < k; ++c) {
        const BinNode &cn = b2[kids[c]];
        dst.lox[c] = cn.lo[0];
        dst.loy[c] = cn.lo[1];
        dst.loz[c] = cn.lo[2];
        dst.hix[c] = cn.hi[0];
        dst.hiy[c] = cn.hi[1];
        dst.hiz[c] = cn.hi[2];
        if (cn.leaf) {
            dst.child[c] = make_leaf(itemBase + cn.first, cn.count);
            st.leaves++;
        }
    }
}

// The wide nodes of the subtree at binary node `b` (depth `depth`), in the
// order of a depth-first walk that takes the last child first: node 0 is
// the subtree's root and child references are local indices.
struct SubCollapse {
    std::vector<Node4> nodes;
    std::vector<uint32_t> kidCount;
    BuildStats st;
    void run(const std::vector<BinNode> &b2, uint32_t b, uint32_t depth, uint32_t width, uint32_t itemBase) {
        struct Pend {
            uint32_t b2node, parent, slot, depth;
        };
        std::vector<Pend> pend{{b, kEmpty, 0, depth}};
        while (!pend.empty()) {
            const Pend p = pend.back();
            pend.pop_back();
            const uint32_t me = (uint32_t)nodes.size();
            nodes.push_back(emptyNode());
            if (p.parent != kEmpty) nodes[p.parent].child[p.slot] = me;
            st.maxDepth = std::max(st.maxDepth, p.depth);
            uint32_t kids[4];
            const uint32_t k = gatherKids(b2, p.b2node, width, kids);
            kidCount.push_back(k);
            fillNode(b2, kids, k, itemBase, nodes[me], st);
            for (uint32_t c = 0; c < k; ++c)
                if (!b2[kids[c]].leaf) pend.push_back({kids[c], me, c, p.depth + 1});
        }
    }
};

}  // namespace

uint32_t CollapseWideBvh(const std::vector<BinNode> &b2, uint32_t root, uint32_t n, const BuildParams &P,
                         uint32_t itemBase, std::vector<Node4> &nodes, BuildStats *stats) {
    const uint32_t width = std::max(2u, std::min(P.width, 4u));
    BuildStats st;
    const uint32_t rootIndex = (uint32_t)nodes.size();
    if (n == 0) {
        nodes.push_back(emptyNode());
        if (stats) *stats = st;
        return rootIndex;
    }
    // Each wide node is made from one binary node and takes its children
    // from gatherKids; nodes are numbered in the order of a depth-first walk
    // from the root that takes the last child first, so every subtree is one
    // contiguous run.  The nodes above depth kSplit are made here; the
    // subtrees at depth kSplit are collapsed on host threads and their runs
    // copied into place, which numbers every node as the serial walk does.
    constexpr uint32_t kSplit = 4;
    struct Top {
        uint32_t b2node, depth, k;
        uint32_t kids[4];
        int32_t sub[4];   // child c: a top node (index into tops), a subtree (-1 - index into subs), or none
    };
    std::vector<Top> tops;
    std::vector<uint32_t> subRoots, subDepth;
    std::vector<uint32_t> open{0};
    tops.push_back(Top{root, 1, 0, {}, {0, 0, 0, 0}});
    const bool parallel = n >= 65536 && !P.serialCollapse;
    while (!open.empty()) {
        const uint32_t t = open.back();
        open.pop_back();
        Top cur = tops[t];
        cur.k = gatherKids(b2, cur.b2node, width, cur.kids);
        for (uint32_t c = 0; c < cur.k; ++c) {
            cur.sub[c] = 0;
            if (b2[cur.kids[c]].leaf) continue;
            if (!parallel || cur.depth + 1 < kSplit) {
                cur.sub[c] = (int32_t)tops.size();
                tops.push_back(Top{cur.kids[c], cur.depth + 1, 0, {}, {0, 0, 0, 0}});
                open.push_back((uint32_t)tops.size() - 1);
            } else {
                cur.sub[c] = -1 - (int32_t)subRoots.size();
                subRoots.push_back(cur.kids[c]);
                subDepth.push_back(cur.depth + 1);
            }
        }
        tops[t] = cur;
    }
    std::vector<SubCollapse> subs(subRoots.size());
    parallel_for(subs.size(), [&](size_t i) { subs[i].run(b2, subRoots[i], subDepth[i], width, itemBase); }, 2);
    // Numbering: the serial walk's order.  A node's internal children are
    // all top nodes (above depth kSplit - 1) or all subtrees (at kSplit - 1):
    // top children are pushed in slot order, so the last is entered first;
    // subtree runs follow their parent, last child first.
    std::vector<uint32_t> topIndex(tops.size()), subBase(subs.size());
    uint32_t next = 0;
    std::vector<uint32_t> walk{0};
    while (!walk.empty()) {
        const uint32_t t = walk.back();
        walk.pop_back();
        topIndex[t] = next++;
        for (uint32_t c = 0; c < tops[t].k; ++c) {
            const int32_t s = tops[t].sub[c];
            if (s > 0) walk.push_back((uint32_t)s);
        }
        for (uint32_t c = tops[t].k; c-- > 0;) {
            const int32_t s = tops[t].sub[c];
            if (s < 0) {
                subBase[-1 - s] = next;
                next += (uint32_t)subs[-1 - s].nodes.size();
            }
        }
    }
    const uint32_t total = next;
    nodes.resize((size_t)rootIndex + total);
    std::vector<uint32_t> kidCount(total);
    for (uint32_t t = 0; t < tops.size(); ++t) {
        Node4 nd = emptyNode();
        fillNode(b2, tops[t].kids, tops[t].k, itemBase, nd, st);
        for (uint32_t c = 0; c < tops[t].k; ++c) {
            const int32_t s = tops[t].sub[c];
            if (s > 0) nd.child[c] = rootIndex + topIndex[s];
            else if (s < 0) nd.child[c] = rootIndex + subBase[-1 - s];
        }
        nodes[rootIndex + topIndex[t]] = nd;
        kidCount[topIndex[t]] = tops[t].k;
        st.maxDepth = std::max(st.maxDepth, tops[t].depth);
    }
    parallel_for(subs.size(), [&](size_t i) {
        const SubCollapse &sc = subs[i];
        const uint32_t base = subBase[i];
        for (size_t j = 0; j < sc.nodes.size(); ++j) {
            Node4 nd = sc.nodes[j];
            for (int c = 0; c < 4; ++c)
                if (nd.child[c] != kEmpty && !ref_is_leaf(nd.child[c])) nd.child[c] += rootIndex + base;
            nodes[(size_t)rootIndex + base + j] = nd;
            kidCount[base + j] = sc.kidCount[j];
        }
    }, 2);
    for (const SubCollapse &sc : subs) {
        st.leaves += sc.st.leaves;
        st.maxDepth = std::max(st.maxDepth, sc.st.maxDepth);
    }
    st.nodes = total;
    // Stack bound: at a node with k children the walk keeps one and pushes
    // k-1, then descends into any one child.
    std::vector<uint32_t> maxStackOf(total, 0);
    for (uint32_t i = total; i-- > 0;) {
        uint32_t best = 0;
        const Node4 &nd = nodes[rootIndex + i];
        for (int c = 0; c < 4; ++c) {
            const uint32_t r = nd.child[c];
            if (r == kEmpty || ref_is_leaf(r)) continue;
            best = std::max(best, maxStackOf[r - rootIndex]);
        }
        const uint32_t kc = kidCount[i];
        maxStackOf[i] = (kc > 0 ? kc - 1 : 0) + best;
    }
    st.maxStack = maxStackOf[0];
    if (stats) *stats = st;
    return rootIndex;
}

uint32_t BuildWideBvh(const std::vector<BuildBox> &boxes, const BuildParams &P, uint32_t itemBase,
                      std::vector<Node4> &nodes, std::vector<uint32_t> &order, BuildStats *stats) {
    std::vector<BinNode> b2;
    BuildBinaryBvh(boxes, P, b2, order);
    return CollapseWideBvh(b2, 0, (uint32_t)boxes.size(), P, itemBase, nodes, stats);
}

}  // namespace az
