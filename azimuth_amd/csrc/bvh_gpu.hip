// bvh_gpu.hip -- the binned-SAH binary BVH of bvh_build.cpp, built on the
// device level by level: the same tree, node for node, as BuildBinaryBvh.
//
// Per level, one workgroup per range of items still to split:
//   bounds   : the range's box and centroid box (exact: min and max)
//   binning  : 32 bins per axis in LDS, as orderable integer min/max
//   sweep    : one lane per bin, scans over the lanes in place of the host's
//              right-to-left and left-to-right sweeps, each split priced with
//              the host's float operations; the first minimum over
//              (axis, bin) wins, as in the host loop
//   decision : the host's leaf test
//   partition: stable (ballot ranks within each workgroup-wide tile), through a
//              scratch copy of the range
// then a scan over the split flags numbers the children and lays out the
// next level's ranges.  The host reads the tree back and collapses it into
// 4-wide nodes (CollapseWideBvh).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cfloat>

#include "bvh_build.h"

namespace az {
namespace {

constexpr uint32_t kBuildBlock = 256;
constexpr uint32_t kBuildBins = kDeviceBuildBins;
// Levels with this many ranges or more split each range with one wave.
constexpr uint32_t kOneWaveLevel = 4096;

struct BuildRange {
    uint32_t begin, end, node;
};

// The host's BinNode, field for field, so the tree is copied into it as is.
struct DevBinNode {
    float lo[3], hi[3];
    uint32_t left, right, first, count, leaf;
};
static_assert(sizeof(DevBinNode) == sizeof(BinNode) && offsetof(DevBinNode, left) == offsetof(BinNode, left) &&
                  offsetof(DevBinNode, leaf) == offsetof(BinNode, leaf),
              "device and host binary nodes share one layout");

__device__ inline uint32_t fkey(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ inline float fval(uint32_t k) { return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k); }

__device__ inline float box_area(const float *lo, const float *hi) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0.0f || dy < 0.0f || dz < 0.0f) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

__global__ __launch_bounds__(kBuildBlock) void k_bvh_init(const float *__restrict__ boxes, uint32_t n,
                                                          float *__restrict__ cen, uint32_t *__restrict__ order) {
    const uint32_t i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 3; ++k) cen[3 * (size_t)i + k] = 0.5f * (boxes[6 * (size_t)i + k] + boxes[6 * (size_t)i + 3 + k]);
    order[i] = i;
}

// Block-wide min (or max) of one float per thread, in `red` (one entry per
// wave of the BLOCK-thread workgroup).
template <int BLOCK>
__device__ inline float block_min(float v, float *red) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_down(v, o));
    const uint32_t w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63u) == 0) red[w] = v;
    __syncthreads();
    float r = red[0];
    for (int k = 1; k < BLOCK / 64; ++k) r = fminf(r, red[k]);
    return r;
}

template <int BLOCK>
__device__ inline float block_max(float v, float *red) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_down(v, o));
    const uint32_t w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63u) == 0) red[w] = v;
    __syncthreads();
    float r = red[0];
    for (int k = 1; k < BLOCK / 64; ++k) r = fmaxf(r, red[k]);
    return r;
}

struct LevelParams {
    uint32_t maxLeaf;
    float nodeCost, itemCost;
};

// BLOCK threads per range: 1024 while a level has few ranges (the top
// levels, where one workgroup walks up to all items), 256 below, and one
// wave once the ranges are many and small.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_bvh_level(const float *__restrict__ boxes, const float *__restrict__ cen,
                                                           uint32_t *__restrict__ order, uint32_t *__restrict__ tmp,
                                                           const BuildRange *__restrict__ ranges,
                                                           DevBinNode *__restrict__ nodes, uint32_t *__restrict__ split,
                                                           uint32_t *__restrict__ mids, LevelParams P) {
    constexpr uint32_t W = BLOCK / 64;
    __shared__ float red[W];
    __shared__ float box[6], cb[6];
    __shared__ uint32_t binLo[3][kBuildBins][3], binHi[3][kBuildBins][3], binCnt[3][kBuildBins];
    __shared__ float axCost[3];
    __shared__ uint32_t axSplit[3];
    __shared__ uint32_t tileCount[W];
    __shared__ uint32_t decision[2];   // split (1/0), mid
    const BuildRange rg = ranges[blockIdx.x];
    const uint32_t t = threadIdx.x;
    const uint32_t cnt = rg.end - rg.begin;

    // ---- bounds -------------------------------------------------------------
    float bl[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bh[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    float cl[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, ch[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = rg.begin + t; i < rg.end; i += BLOCK) {
        const uint32_t it = order[i];
        for (int k = 0; k < 3; ++k) {
            bl[k] = fminf(bl[k], boxes[6 * (size_t)it + k]);
            bh[k] = fmaxf(bh[k], boxes[6 * (size_t)it + 3 + k]);
            const float c = cen[3 * (size_t)it + k];
            cl[k] = fminf(cl[k], c);
            ch[k] = fmaxf(ch[k], c);
        }
    }
    for (int k = 0; k < 3; ++k) {
        const float a = block_min<BLOCK>(bl[k], red);
        const float b = block_max<BLOCK>(bh[k], red);
        const float c = block_min<BLOCK>(cl[k], red);
        const float d = block_max<BLOCK>(ch[k], red);
        if (t == 0) {
            box[k] = a;
            box[3 + k] = b;
            cb[k] = c;
            cb[3 + k] = d;
        }
    }
    __syncthreads();

    // ---- binning ------------------------------------------------------------
    if (cnt > 1) {
        for (uint32_t b = t; b < 3 * kBuildBins; b += BLOCK) {
            const uint32_t ax = b / kBuildBins, bb = b % kBuildBins;
            for (int k = 0; k < 3; ++k) {
                binLo[ax][bb][k] = fkey(FLT_MAX);
                binHi[ax][bb][k] = fkey(-FLT_MAX);
            }
            binCnt[ax][bb] = 0u;
        }
        __syncthreads();
        // Each thread bins a contiguous slice of the range and keeps, per
        // axis, the box and count of its current bin, adding them to the
        // shared bins only when the bin changes: neighbouring items tend to
        // share bins, and a wave's atomics on one bin would serialize.
        const uint32_t per = (cnt + BLOCK - 1) / BLOCK;
        const uint32_t s0 = rg.begin + min(cnt, t * per), s1 = rg.begin + min(cnt, (t + 1) * per);
        int curBin[3] = {-1, -1, -1};
        uint32_t runLo[3][3], runHi[3][3], runCnt[3] = {0u, 0u, 0u};
        auto flush = [&](int ax) {
            const int b = curBin[ax];
            if (b < 0) return;
            for (int k = 0; k < 3; ++k) {
                atomicMin(&binLo[ax][b][k], runLo[ax][k]);
                atomicMax(&binHi[ax][b][k], runHi[ax][k]);
            }
            atomicAdd(&binCnt[ax][b], runCnt[ax]);
        };
        for (uint32_t i = s0; i < s1; ++i) {
            const uint32_t it = order[i];
            uint32_t lo[3], hi[3];
            for (int k = 0; k < 3; ++k) {
                lo[k] = fkey(boxes[6 * (size_t)it + k]);
                hi[k] = fkey(boxes[6 * (size_t)it + 3 + k]);
            }
            for (int ax = 0; ax < 3; ++ax) {
                const float ext = cb[3 + ax] - cb[ax];
                if (!(ext > 0.0f)) continue;
                const float scale = (float)kBuildBins / ext;
                int b = (int)((cen[3 * (size_t)it + ax] - cb[ax]) * scale);
                b = min(max(b, 0), (int)kBuildBins - 1);
                if (b != curBin[ax]) {
                    flush(ax);
                    curBin[ax] = b;
                    for (int k = 0; k < 3; ++k) {
                        runLo[ax][k] = lo[k];
                        runHi[ax][k] = hi[k];
                    }
                    runCnt[ax] = 1u;
                } else {
                    for (int k = 0; k < 3; ++k) {
                        runLo[ax][k] = min(runLo[ax][k], lo[k]);
                        runHi[ax][k] = max(runHi[ax][k], hi[k]);
                    }
                    ++runCnt[ax];
                }
            }
        }
        for (int ax = 0; ax < 3; ++ax) flush(ax);
        __syncthreads();
        // The host's two sweeps (BuildBinaryBvh), one lane per bin: the
        // running boxes and counts are scans over the lanes (min and max are
        // exact in any order; lanes 32-63 hold empty bins), each lane prices
        // the split above its bin with the host's float operations, and the
        // cheapest split wins, the lowest bin on ties, as in the host's
        // first-minimum loop.  A cost that is not below FLT_MAX never wins.
        for (uint32_t ax = t >> 6; ax < 3u; ax += W) {
            const uint32_t lane = t & 63u;
            const float ext = cb[3 + ax] - cb[ax];
            float best = FLT_MAX;
            uint32_t bestSplit = 0;
            if (ext > 0.0f) {
                float plo[3], phi[3], slo[3], shi[3];
                uint32_t pc = 0u;
                for (int k = 0; k < 3; ++k) {
                    plo[k] = lane < kBuildBins ? fval(binLo[ax][lane][k]) : FLT_MAX;
                    phi[k] = lane < kBuildBins ? fval(binHi[ax][lane][k]) : -FLT_MAX;
                    slo[k] = plo[k];
                    shi[k] = phi[k];
                }
                if (lane < kBuildBins) pc = binCnt[ax][lane];
                uint32_t sc = pc;
                for (uint32_t o = 1; o < 64u; o <<= 1) {
                    const bool up = lane >= o, dn = lane + o < 64u;
                    for (int k = 0; k < 3; ++k) {
                        const float a0 = __shfl_up(plo[k], o), a1 = __shfl_up(phi[k], o);
                        const float d0 = __shfl_down(slo[k], o), d1 = __shfl_down(shi[k], o);
                        if (up) {
                            plo[k] = fminf(plo[k], a0);
                            phi[k] = fmaxf(phi[k], a1);
                        }
                        if (dn) {
                            slo[k] = fminf(slo[k], d0);
                            shi[k] = fmaxf(shi[k], d1);
                        }
                    }
                    const uint32_t ac = __shfl_up(pc, o), dc = __shfl_down(sc, o);
                    if (up) pc += ac;
                    if (dn) sc += dc;
                }
                // Right side of the split above bin `lane`: the suffix from lane + 1.
                float rlo[3], rhi[3];
                for (int k = 0; k < 3; ++k) {
                    rlo[k] = __shfl_down(slo[k], 1);
                    rhi[k] = __shfl_down(shi[k], 1);
                }
                const uint32_t rc = __shfl_down(sc, 1);
                if (lane + 1u < kBuildBins && pc != 0u && rc != 0u) {
                    const float cost = box_area(plo, phi) * (float)pc + box_area(rlo, rhi) * (float)rc;
                    if (cost < FLT_MAX) {
                        best = cost;
                        bestSplit = lane + 1u;
                    }
                }
                for (uint32_t o = 32; o > 0; o >>= 1) {
                    const float oc = __shfl_xor(best, o);
                    const uint32_t os = __shfl_xor(bestSplit, o);
                    if (oc < best || (oc == best && os < bestSplit)) {
                        best = oc;
                        bestSplit = os;
                    }
                }
            }
            if (lane == 0u) {
                axCost[ax] = best;
                axSplit[ax] = bestSplit;
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        int bestAxis = -1;
        float bestCost = FLT_MAX;
        if (cnt > 1)
            for (int ax = 0; ax < 3; ++ax)
                if (axCost[ax] < bestCost) {
                    bestCost = axCost[ax];
                    bestAxis = ax;
                }
        const float area = box_area(box, box + 3);
        const float splitCost =
            bestAxis >= 0 && area > 0.0f ? 0.5f * P.nodeCost + P.itemCost * bestCost / area : FLT_MAX;
        const float leafCost = P.itemCost * (float)cnt;
        const bool makeLeaf = cnt <= 1 || (cnt <= P.maxLeaf && leafCost <= splitCost);
        DevBinNode &nd = nodes[rg.node];
        for (int k = 0; k < 3; ++k) {
            nd.lo[k] = box[k];
            nd.hi[k] = box[3 + k];
        }
        nd.leaf = makeLeaf ? 1u : 0u;
        nd.first = rg.begin;
        nd.count = cnt;
        nd.left = nd.right = 0u;
        decision[0] = makeLeaf ? 0u : 1u;
        decision[1] = bestAxis >= 0 ? (uint32_t)bestAxis : 3u;
        tileCount[0] = axSplit[bestAxis >= 0 ? bestAxis : 0];
    }
    __syncthreads();
    const bool doSplit = decision[0] != 0u;
    const uint32_t ax = decision[1];
    const uint32_t bestSplit = tileCount[0];
    __syncthreads();
    uint32_t mid = rg.begin + cnt / 2u;
    if (doSplit && ax < 3u) {
        // ---- stable partition ------------------------------------------------
        const float ext = cb[3 + ax] - cb[ax];
        const float scale = (float)kBuildBins / ext;
        auto goesLeft = [&](uint32_t it) {
            int b = (int)((cen[3 * (size_t)it + ax] - cb[ax]) * scale);
            b = min(max(b, 0), (int)kBuildBins - 1);
            return (uint32_t)b < bestSplit;
        };
        uint32_t myLeft = 0;
        for (uint32_t i = rg.begin + t; i < rg.end; i += BLOCK) myLeft += goesLeft(order[i]) ? 1u : 0u;
        for (int o = 32; o > 0; o >>= 1) myLeft += __shfl_down(myLeft, o);
        if ((t & 63u) == 0) tileCount[t >> 6] = myLeft;
        __syncthreads();
        uint32_t nLeft = 0;
        for (uint32_t w = 0; w < W; ++w) nLeft += tileCount[w];
        __syncthreads();
        if (nLeft != 0u && nLeft != cnt) {
            mid = rg.begin + nLeft;
            uint32_t leftBase = 0, rightBase = 0;
            for (uint32_t t0 = rg.begin; t0 < rg.end; t0 += BLOCK) {
                const uint32_t i = t0 + t;
                const bool valid = i < rg.end;
                const uint32_t it = valid ? order[i] : 0u;
                const bool l = valid && goesLeft(it);
                const uint64_t ml = __ballot(l), mv = __ballot(valid);
                const uint32_t lo = (uint32_t)ml, hi = (uint32_t)(ml >> 32);
                const uint32_t belowL = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
                const uint32_t vlo = (uint32_t)mv, vhi = (uint32_t)(mv >> 32);
                const uint32_t belowV = __builtin_amdgcn_mbcnt_hi(vhi, __builtin_amdgcn_mbcnt_lo(vlo, 0u));
                if ((t & 63u) == 0) tileCount[t >> 6] = (uint32_t)__popcll(ml);
                __syncthreads();
                uint32_t waveLeft = 0, tileLeft = 0;
                for (uint32_t w = 0; w < W; ++w) {
                    if (w < (t >> 6)) waveLeft += tileCount[w];
                    tileLeft += tileCount[w];
                }
                if (valid) {
                    const uint32_t leftRank = waveLeft + belowL;
                    const uint32_t rank = (t & ~63u) + belowV;            // position within the tile
                    const uint32_t dst = l ? rg.begin + leftBase + leftRank
                                           : mid + rightBase + (rank - leftRank);
                    tmp[dst] = it;
                }
                leftBase += tileLeft;
                rightBase += min((uint32_t)BLOCK, rg.end - t0) - tileLeft;
                __syncthreads();
            }
            __threadfence_block();
            __syncthreads();
            for (uint32_t i = rg.begin + t; i < rg.end; i += BLOCK) order[i] = tmp[i];
        }
    }
    if (t == 0) {
        split[blockIdx.x] = doSplit ? 1u : 0u;
        mids[blockIdx.x] = mid;
    }
}

// Exclusive scan of u32 in place, for up to kBuildBlock * kBuildBlock * 64 values: per-block
// sums, one block scanning them, then each block adding its offset.
__global__ __launch_bounds__(kBuildBlock) void k_scan_blocks(const uint32_t *__restrict__ in, uint32_t n,
                                                             uint32_t *__restrict__ out, uint32_t *__restrict__ sums) {
    __shared__ uint32_t s[kBuildBlock];
    const uint32_t i = blockIdx.x * kBuildBlock + threadIdx.x;
    const uint32_t v = i < n ? in[i] : 0u;
    s[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t o = 1; o < kBuildBlock; o <<= 1) {
        const uint32_t a = threadIdx.x >= o ? s[threadIdx.x - o] : 0u;
        __syncthreads();
        s[threadIdx.x] += a;
        __syncthreads();
    }
    if (i < n) out[i] = s[threadIdx.x] - v;
    if (threadIdx.x == kBuildBlock - 1) sums[blockIdx.x] = s[threadIdx.x];
}

__global__ __launch_bounds__(kBuildBlock) void k_scan_sums(uint32_t *__restrict__ sums, uint32_t nb,
                                                           uint32_t *__restrict__ total) {
    __shared__ uint32_t s[kBuildBlock];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += kBuildBlock) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < nb ? sums[i] : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t o = 1; o < kBuildBlock; o <<= 1) {
            const uint32_t a = threadIdx.x >= o ? s[threadIdx.x - o] : 0u;
            __syncthreads();
            s[threadIdx.x] += a;
            __syncthreads();
        }
        if (i < nb) sums[i] = carry + s[threadIdx.x] - v;
        carry += s[kBuildBlock - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kBuildBlock) void k_scan_add(uint32_t *__restrict__ out, uint32_t n,
                                                          const uint32_t *__restrict__ sums) {
    const uint32_t i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i < n) out[i] += sums[blockIdx.x];
}

// Children of the split ranges: node ids base + 2 p and base + 2 p + 1 (p =
// the range's rank among the split ones), and the next level's ranges.
__global__ __launch_bounds__(kBuildBlock) void k_bvh_emit(const BuildRange *__restrict__ ranges, uint32_t R,
                                                          const uint32_t *__restrict__ split,
                                                          const uint32_t *__restrict__ rankOf,
                                                          const uint32_t *__restrict__ mids, uint32_t base,
                                                          DevBinNode *__restrict__ nodes, BuildRange *__restrict__ next) {
    const uint32_t r = blockIdx.x * kBuildBlock + threadIdx.x;
    if (r >= R || !split[r]) return;
    const BuildRange rg = ranges[r];
    const uint32_t p = rankOf[r], c0 = base + 2u * p;
    nodes[rg.node].left = c0;
    nodes[rg.node].right = c0 + 1u;
    next[2u * p] = BuildRange{rg.begin, mids[r], c0};
    next[2u * p + 1u] = BuildRange{mids[r], rg.end, c0 + 1u};
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

int BuildBinaryBvhDevice(void *stream, const std::vector<BuildBox> &boxes, const BuildParams &P,
                         std::vector<BinNode> &b2, std::vector<uint32_t> &order) {
    hipStream_t st = (hipStream_t)stream;
    const uint32_t n = (uint32_t)boxes.size();
    b2.clear();
    order.clear();
    if (P.bins != kBuildBins) return (int)hipErrorInvalidValue;   // the host binning would differ
    if (n == 0) {
        b2.emplace_back();
        return hipSuccess;
    }
    const uint32_t nodeCap = 2u * n;
    const uint32_t scanBlocks = (n + kBuildBlock - 1) / kBuildBlock;
    DevBuf dBoxes, dCen, dOrder, dTmp, dNodes, dR0, dR1, dSplit, dRank, dMids, dSums, dTotal;
    hipError_t e = hipSuccess;
#define AZ_BUILD_TRY(call)          \
    do {                            \
        e = (call);                 \
        if (e != hipSuccess) return (int)e; \
    } while (0)
    AZ_BUILD_TRY(hipMalloc(&dBoxes.p, (size_t)n * sizeof(BuildBox)));
    AZ_BUILD_TRY(hipMalloc(&dCen.p, (size_t)n * 3 * sizeof(float)));
    AZ_BUILD_TRY(hipMalloc(&dOrder.p, (size_t)n * sizeof(uint32_t)));
    AZ_BUILD_TRY(hipMalloc(&dTmp.p, (size_t)n * sizeof(uint32_t)));
    AZ_BUILD_TRY(hipMalloc(&dNodes.p, (size_t)nodeCap * sizeof(DevBinNode)));
    AZ_BUILD_TRY(hipMalloc(&dR0.p, (size_t)n * sizeof(BuildRange)));
    AZ_BUILD_TRY(hipMalloc(&dR1.p, (size_t)n * sizeof(BuildRange)));
    AZ_BUILD_TRY(hipMalloc(&dSplit.p, (size_t)n * sizeof(uint32_t)));
    AZ_BUILD_TRY(hipMalloc(&dRank.p, (size_t)n * sizeof(uint32_t)));
    AZ_BUILD_TRY(hipMalloc(&dMids.p, (size_t)n * sizeof(uint32_t)));
    AZ_BUILD_TRY(hipMalloc(&dSums.p, (size_t)scanBlocks * sizeof(uint32_t)));
    AZ_BUILD_TRY(hipMalloc(&dTotal.p, sizeof(uint32_t)));
    AZ_BUILD_TRY(hipMemcpyAsync(dBoxes.p, boxes.data(), (size_t)n * sizeof(BuildBox), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_bvh_init, dim3(scanBlocks), dim3(kBuildBlock), 0, st, (const float *)dBoxes.p, n,
                       (float *)dCen.p, (uint32_t *)dOrder.p);
    const BuildRange root{0u, n, 0u};
    AZ_BUILD_TRY(hipMemcpyAsync(dR0.p, &root, sizeof(root), hipMemcpyHostToDevice, st));
    LevelParams lp{std::max(1u, std::min(P.maxLeafItems, kMaxLeafItems)), P.nodeCost, P.itemCost};
    BuildRange *cur = (BuildRange *)dR0.p, *nxt = (BuildRange *)dR1.p;
    uint32_t R = 1, nodeCount = 1;
    while (R > 0) {
        if (R < 1024u)
            hipLaunchKernelGGL(k_bvh_level<1024>, dim3(R), dim3(1024), 0, st, (const float *)dBoxes.p,
                               (const float *)dCen.p, (uint32_t *)dOrder.p, (uint32_t *)dTmp.p, (const BuildRange *)cur,
                               (DevBinNode *)dNodes.p, (uint32_t *)dSplit.p, (uint32_t *)dMids.p, lp);
        else if (R < kOneWaveLevel)
            hipLaunchKernelGGL(k_bvh_level<256>, dim3(R), dim3(256), 0, st, (const float *)dBoxes.p,
                               (const float *)dCen.p, (uint32_t *)dOrder.p, (uint32_t *)dTmp.p, (const BuildRange *)cur,
                               (DevBinNode *)dNodes.p, (uint32_t *)dSplit.p, (uint32_t *)dMids.p, lp);
        else
            hipLaunchKernelGGL(k_bvh_level<64>, dim3(R), dim3(64), 0, st, (const float *)dBoxes.p,
                               (const float *)dCen.p, (uint32_t *)dOrder.p, (uint32_t *)dTmp.p, (const BuildRange *)cur,
                               (DevBinNode *)dNodes.p, (uint32_t *)dSplit.p, (uint32_t *)dMids.p, lp);
        const uint32_t rb = (R + kBuildBlock - 1) / kBuildBlock;
        hipLaunchKernelGGL(k_scan_blocks, dim3(rb), dim3(kBuildBlock), 0, st, (const uint32_t *)dSplit.p, R,
                           (uint32_t *)dRank.p, (uint32_t *)dSums.p);
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBuildBlock), 0, st, (uint32_t *)dSums.p, rb,
                           (uint32_t *)dTotal.p);
        hipLaunchKernelGGL(k_scan_add, dim3(rb), dim3(kBuildBlock), 0, st, (uint32_t *)dRank.p, R,
                           (const uint32_t *)dSums.p);
        uint32_t S = 0;
        AZ_BUILD_TRY(hipMemcpyAsync(&S, dTotal.p, sizeof(S), hipMemcpyDeviceToHost, st));
        AZ_BUILD_TRY(hipStreamSynchronize(st));
        if (S > 0) {
            if (nodeCount + 2u * S > nodeCap || 2u * S > n) return (int)hipErrorInvalidValue;
            hipLaunchKernelGGL(k_bvh_emit, dim3(rb), dim3(kBuildBlock), 0, st, (const BuildRange *)cur, R,
                               (const uint32_t *)dSplit.p, (const uint32_t *)dRank.p, (const uint32_t *)dMids.p,
                               nodeCount, (DevBinNode *)dNodes.p, nxt);
        }
        AZ_BUILD_TRY(hipGetLastError());
        nodeCount += 2u * S;
        R = 2u * S;
        std::swap(cur, nxt);
    }
    b2.resize(nodeCount);
    order.resize(n);
    AZ_BUILD_TRY(hipMemcpyAsync(b2.data(), dNodes.p, (size_t)nodeCount * sizeof(BinNode), hipMemcpyDeviceToHost, st));
    AZ_BUILD_TRY(hipMemcpyAsync(order.data(), dOrder.p, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    AZ_BUILD_TRY(hipStreamSynchronize(st));
#undef AZ_BUILD_TRY
    return hipSuccess;
}

}  // namespace az
