// Checks CollapseWideBvh on random boxes, below and above the size from which
// it collapses subtrees on host threads:
//   numbering: a depth-first walk from the root that enters the last child
//              first visits the nodes as 0, 1, 2, ... (each subtree is one
//              contiguous run), as the serial collapse numbers them;
//   items    : every item lies in exactly one leaf;
//   boxes    : every child box is the union of the item boxes below it;
//   threads  : from 65536 items, the threaded collapse equals the serial one
//              byte for byte.
// Prints one line per case and "ok" at the end; exits non-zero on a failure.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "bvh_build.h"

using namespace az;

static int fail(const char *what, unsigned n, unsigned w) {
    std::printf("FAIL %s (n=%u width=%u)\n", what, n, w);
    return 1;
}

int main() {
    const unsigned sizes[] = {1, 7, 3000, 65535, 65536, 150000};
    for (unsigned n : sizes)
        for (unsigned width = 2; width <= 4; ++width) {
            std::mt19937 rng(n * 31u + width);
            std::uniform_real_distribution<float> U(-50.f, 50.f), S(0.001f, 3.f);
            std::vector<BuildBox> boxes(n);
            for (BuildBox &b : boxes)
                for (int k = 0; k < 3; ++k) {
                    const float c = U(rng), e = S(rng);
                    b.lo[k] = c - e;
                    b.hi[k] = c + e;
                }
            BuildParams P;
            P.width = width;
            P.maxLeafItems = 4;
            std::vector<Node4> nodes(5);   // the tree starts after other nodes
            std::vector<uint32_t> order;
            BuildStats st;
            const uint32_t itemBase = 100;
            const uint32_t root = BuildWideBvh(boxes, P, itemBase, nodes, order, &st);
            if (root != 5 || nodes.size() != 5 + st.nodes) return fail("root or node count", n, width);
            std::vector<int> seen(n, 0);
            uint32_t expect = root;
            struct Visit {
                uint32_t node;
                int parentSlot;
                uint32_t parent;
            };
            std::vector<Visit> walk{{root, -1, 0}};
            while (!walk.empty()) {
                const Visit v = walk.back();
                walk.pop_back();
                if (v.node != expect++) return fail("numbering", n, width);
                const Node4 &nd = nodes[v.node];
                for (int c = 0; c < 4; ++c) {
                    const uint32_t r = nd.child[c];
                    if (r == kEmpty || !ref_is_leaf(r)) continue;
                    const uint32_t first = leaf_first(r) - itemBase, count = leaf_count(r);
                    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
                    for (uint32_t i = first; i < first + count; ++i) {
                        if (i >= n || seen[order[i]]++) return fail("item coverage", n, width);
                        for (int k = 0; k < 3; ++k) {
                            lo[k] = std::min(lo[k], boxes[order[i]].lo[k]);
                            hi[k] = std::max(hi[k], boxes[order[i]].hi[k]);
                        }
                    }
                    if (lo[0] != nd.lox[c] || lo[1] != nd.loy[c] || lo[2] != nd.loz[c] || hi[0] != nd.hix[c] ||
                        hi[1] != nd.hiy[c] || hi[2] != nd.hiz[c])
                        return fail("leaf box", n, width);
                }
                for (int c = 0; c < 4; ++c) {
                    const uint32_t r = nd.child[c];
                    if (r != kEmpty && !ref_is_leaf(r)) walk.push_back({r, c, v.node});
                }
            }
            if (expect != nodes.size()) return fail("unvisited nodes", n, width);
            for (unsigned i = 0; i < n; ++i)
                if (seen[i] != 1) return fail("item coverage", n, width);
            // Inner child boxes: the union of the child's own child boxes.
            for (size_t v = root; v < nodes.size(); ++v)
                for (int c = 0; c < 4; ++c) {
                    const uint32_t r = nodes[v].child[c];
                    if (r == kEmpty || ref_is_leaf(r)) continue;
                    const Node4 &ch = nodes[r];
                    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
                    for (int d = 0; d < 4; ++d) {
                        if (ch.child[d] == kEmpty) continue;
                        lo[0] = std::min(lo[0], ch.lox[d]);
                        lo[1] = std::min(lo[1], ch.loy[d]);
                        lo[2] = std::min(lo[2], ch.loz[d]);
                        hi[0] = std::max(hi[0], ch.hix[d]);
                        hi[1] = std::max(hi[1], ch.hiy[d]);
                        hi[2] = std::max(hi[2], ch.hiz[d]);
                    }
                    if (lo[0] != nodes[v].lox[c] || lo[1] != nodes[v].loy[c] || lo[2] != nodes[v].loz[c] ||
                        hi[0] != nodes[v].hix[c] || hi[1] != nodes[v].hiy[c] || hi[2] != nodes[v].hiz[c])
                        return fail("inner box", n, width);
                }
            // The threaded collapse gives the serial collapse's nodes byte for byte.
            if (n >= 65536) {
                BuildParams Ps = P;
                Ps.serialCollapse = true;
                std::vector<Node4> serial(5);
                std::vector<uint32_t> serialOrder;
                BuildStats ss;
                const uint32_t sroot = BuildWideBvh(boxes, Ps, itemBase, serial, serialOrder, &ss);
                if (sroot != root || serial.size() != nodes.size() || serialOrder != order)
                    return fail("serial collapse shape", n, width);
                if (std::memcmp(serial.data(), nodes.data(), nodes.size() * sizeof(Node4)) != 0)
                    return fail("serial collapse bytes", n, width);
                if (ss.nodes != st.nodes || ss.leaves != st.leaves || ss.maxDepth != st.maxDepth ||
                    ss.maxStack != st.maxStack)
                    return fail("serial collapse stats", n, width);
            }
            std::printf("n=%u width=%u nodes=%u leaves=%u depth=%u stack=%u\n", n, width, st.nodes, st.leaves,
                        st.maxDepth, st.maxStack);
        }
    std::printf("ok\n");
    return 0;
}
