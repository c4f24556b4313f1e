// Device scene preparation for gfx950: binned-SAH BVH build, the
// leaf-ordered triangle / attribute gather and the 4-wide collapse.
//
// The tree is the host builder's (csrc/host/bvh.cpp) word for word: the
// same 32-bin SAH with the same float expressions, boxes as unions of
// triangle boxes (min / max are exact, so only the triangle set matters),
// stable partitions (leaves list ascending triangle ids) and depth-first
// node numbering. The reference packs on the host in GpuPathTracer::upload
// (src/pathtrace.rs:970-1239) from oxideav-mesh3d's Bvh; the node layout
// is the kernel's (src/shaders/pathtrace.wgsl:638-735).
//
// Build order:
//  1. per-triangle boxes and centroids (-0 turned into +0, NaN skipped);
//  2. large nodes (> kBuildSmall triangles), one level per step: every
//     workgroup bins a 4096-position chunk of one node in LDS and flushes
//     with global atomics; one wave per node evaluates the SAH over the
//     3 x 32 bins; a chunked count / scan / scatter stable-partitions the
//     node's ids. Boxes are kept as order-preserving integer encodings so
//     that min / max are integer atomics;
//  3. small subtrees: one wave per subtree, ids in LDS, depth-first with an
//     LDS stack, the same SAH decision and a ballot stable partition;
//  4. numbering: a node's depth-first rank is the number of inner nodes
//     whose slot range starts earlier, plus its depth below the topmost
//     inner node starting at the same slot (inner nodes starting at one
//     slot form a left spine), so preorder comes from one scan over slots;
//  5. the 4-wide collapse level by level (open the inner child with the
//     largest surface area until four children), numbered the same way.
#include "bvh_build.hpp"

#include <algorithm>
#include <chrono>
#include <vector>

#include "oxpt/oxpt.h"

namespace oxd {
namespace {

constexpr int kBins = 32;
constexpr uint32_t kMaxLeaf = 4;
constexpr uint32_t kHardLeaf = 8;
constexpr uint32_t kMaxDepth = 56;
constexpr float kTravCost = 1.0f;
constexpr float kTriCost = 1.0f;
constexpr uint32_t kWideLeafMax = 30;
constexpr uint32_t kWideStack = 96;
constexpr uint32_t kEmpty = 0xffffffffu;
constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kInner = 0x80000000u;
// Per bin: box lo3, hi3, centroid lo3, hi3, count.
constexpr uint32_t kBinWords = 13;
constexpr uint32_t kNodeBins = 3 * kBins * kBinWords;
constexpr uint32_t kChunk = 4096;  // positions per large-phase workgroup
constexpr uint32_t kLargeBlock = 256;
constexpr uint32_t kPerThread = kChunk / kLargeBlock;
constexpr uint32_t kLoId = 0xff800000u;  // enc(+inf): identity of min
constexpr uint32_t kHiId = 0x007fffffu;  // enc(-inf): identity of max
constexpr uint32_t kStackCap = 64;       // small-phase stack entries (depth <= 56)
constexpr uint32_t kScanBlock = 1024;    // items per scan workgroup

enum Counter : uint32_t {
    C_NODES = 0,
    C_NEXT_OPEN,
    C_NEXT_MAXN,
    C_MEDIAN,
    C_SMALL,
    C_DEPTH,
    C_WIDE,
    C_WIDE_NEXT,
    C_WIDE_BAD,
    C_WIDE_DEPTH,
    C_INNER,
    C_COUNT
};

// Temporary node: encoded box and centroid box, slot range, parent link.
struct TNode {
    uint32_t box[6];  // enc lo3, hi3
    uint32_t cb[6];   // enc centroid lo3, hi3
    uint32_t b, n;    // slot range [b, b + n)
    uint32_t parent;  // parent << 1 | side, kNone for the root
    uint32_t depth;   // | kInner once the node has children
};
static_assert(sizeof(TNode) == 64, "TNode is 64 bytes");

// Order-preserving float <-> u32 (negative values flipped, positive ones
// with the sign bit set): integer min / max are float min / max.
__device__ inline uint32_t enc(float f) {
    uint32_t u = __float_as_uint(f);
    return u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
}
__device__ inline float dec(uint32_t u) { return __uint_as_float(u ^ (((u >> 31) - 1u) | 0x80000000u)); }

__device__ inline bool lo_word(uint32_t w) { return w < 3 || (w >= 6 && w < 9); }
__device__ inline uint32_t identity(uint32_t w) { return w == 12 ? 0u : (lo_word(w) ? kLoId : kHiId); }
__device__ inline uint32_t combine(uint32_t w, uint32_t a, uint32_t b) {
    return w == 12 ? a + b : (lo_word(w) ? min(a, b) : max(a, b));
}

// The host builder's bin index (bvh.cpp bin_of).
__device__ inline int bin_of(float c, float lo, float scale) {
    float f = (c - lo) * scale;
    if (!(f > 0.0f)) return 0;
    if (f >= (float)kBins) return kBins - 1;
    return (int)f;
}

__device__ inline float box_area(const float lo[3], const float hi[3]) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (!(dx >= 0.0f && dy >= 0.0f && dz >= 0.0f)) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Element i of a 3-array by selects (a dynamic index would place the array
// in scratch memory).
template <typename T>
__device__ inline T pick3(const T a[3], uint32_t i) {
    return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]);
}

__device__ inline uint32_t pick6(const uint32_t a[6], uint32_t i) {
    return i < 3 ? pick3(a, i) : pick3(a + 3, i - 3);
}

__device__ inline float coord(const float4 *tlo, const float4 *thi, const float *tcz, uint32_t t, uint32_t ax) {
    return ax == 0 ? tlo[t].w : (ax == 1 ? thi[t].w : tcz[t]);
}

__device__ inline unsigned long long shfl_xor_u64(unsigned long long v, int m) {
    uint32_t lo = __shfl_xor((uint32_t)v, m), hi = __shfl_xor((uint32_t)(v >> 32), m);
    return ((unsigned long long)hi << 32) | lo;
}

// Exclusive scan of one value per thread over a workgroup of `nt` threads
// (a multiple of 64, at most 1024); returns the prefix, *total the sum.
__device__ uint32_t block_exclusive(uint32_t v, uint32_t *lds_waves, uint32_t *total) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t x = v;
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += t;
    }
    if (lane == 63) lds_waves[wave] = x;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (uint32_t w = 0; w < nw; w++) {
        uint32_t s = lds_waves[w];
        if (w < wave) base += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return base + x - v;
}

// ---- 1. triangle boxes ----------------------------------------------

__global__ void k_prep(const float *pos, uint32_t n, float4 *tlo, float4 *thi, float *tcz, uint32_t *ids) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int v = 0; v < 3; v++)
        for (int k = 0; k < 3; k++) {
            float q = pos[9 * (size_t)t + 3 * v + k] + 0.0f;  // -0 -> +0
            lo[k] = q < lo[k] ? q : lo[k];                     // std::min(lo, q): NaN skipped
            hi[k] = hi[k] < q ? q : hi[k];
        }
    float c[3];
    for (int k = 0; k < 3; k++) c[k] = 0.5f * (lo[k] + hi[k]);
    tlo[t] = make_float4(lo[0], lo[1], lo[2], c[0]);
    thi[t] = make_float4(hi[0], hi[1], hi[2], c[1]);
    tcz[t] = c[2];
    ids[t] = t;
}

__global__ void k_init(TNode *nodes, uint32_t n, uint32_t *counters, uint32_t *open, uint32_t *small) {
    TNode r;
    for (int k = 0; k < 3; k++) {
        r.box[k] = kLoId;
        r.box[3 + k] = kHiId;
        r.cb[k] = kLoId;
        r.cb[3 + k] = kHiId;
    }
    r.b = 0;
    r.n = n;
    r.parent = kNone;
    r.depth = 0;
    nodes[0] = r;
    for (uint32_t i = 0; i < C_COUNT; i++) counters[i] = 0;
    counters[C_NODES] = 1;
    if (n <= kBuildSmall) {
        small[0] = 0;
        counters[C_SMALL] = 1;
    } else {
        open[0] = 0;
    }
}

__global__ __launch_bounds__(256) void k_root_bounds(const float4 *tlo, const float4 *thi, const float *tcz, uint32_t n,
                                                     TNode *nodes) {
    uint32_t v[12];
    for (uint32_t w = 0; w < 12; w++) v[w] = lo_word(w) ? kLoId : kHiId;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        float4 l = tlo[t], h = thi[t];
        float c[3] = {l.w, h.w, tcz[t]};
        v[0] = min(v[0], enc(l.x));
        v[1] = min(v[1], enc(l.y));
        v[2] = min(v[2], enc(l.z));
        v[3] = max(v[3], enc(h.x));
        v[4] = max(v[4], enc(h.y));
        v[5] = max(v[5], enc(h.z));
        for (int k = 0; k < 3; k++)
            if (c[k] == c[k]) {
                v[6 + k] = min(v[6 + k], enc(c[k]));
                v[9 + k] = max(v[9 + k], enc(c[k]));
            }
    }
    for (int m = 32; m >= 1; m >>= 1)
        for (uint32_t w = 0; w < 12; w++) {
            uint32_t o = __shfl_xor(v[w], m);
            v[w] = lo_word(w) ? min(v[w], o) : max(v[w], o);
        }
    if ((threadIdx.x & 63) == 0) {
        uint32_t *r = nodes[0].box;  // box then cb: 12 consecutive words
        for (uint32_t w = 0; w < 12; w++) {
            if (lo_word(w)) {
                if (v[w] != kLoId) atomicMin(&r[w], v[w]);
            } else if (v[w] != kHiId) {
                atomicMax(&r[w], v[w]);
            }
        }
    }
}

// ---- the SAH decision (one wave, bins in LDS) ----------------------------

struct Decision {
    uint32_t kind;  // 0 leaf, 1 binned split, 2 object median
    uint32_t axis, split;
};

struct NodeF {
    float blo[3], bhi[3], clo[3], chi[3];
};

__device__ inline NodeF decode(const uint32_t box[6], const uint32_t cb[6]) {
    NodeF f;
    for (int k = 0; k < 3; k++) {
        f.blo[k] = dec(box[k]);
        f.bhi[k] = dec(box[3 + k]);
        f.clo[k] = dec(cb[k]);
        f.chi[k] = dec(cb[3 + k]);
    }
    return f;
}

// All 64 lanes of the wave call this with the same arguments. Follows the
// host builder's Builder::build decision step by step: for each axis with
// a positive centroid extent, the cost of every split between bins from
// left / right unions; the first minimum in (axis, bin) order wins.
__device__ Decision decide(const uint32_t *sb, const NodeF &f, uint32_t n, uint32_t depth) {
    Decision d{0, 0, 0};
    if (n <= 1 || depth >= kMaxDepth) return d;
    const uint32_t lane = threadIdx.x & 63;
    bool ok[3];
    for (int k = 0; k < 3; k++) ok[k] = (f.chi[k] - f.clo[k]) > 0.0f;
    unsigned long long best = ~0ull;
    for (uint32_t pass = 0; pass < 2; pass++) {
        const uint32_t ax = pass * 2 + (lane >> 5), i = lane & 31;
        const bool axv = ax < 3 && pick3(ok, ax);
        uint32_t L[7], R[7];  // lo3 hi3 count
        if (ax < 3) {
            const uint32_t *p = &sb[(ax * kBins + i) * kBinWords];
            for (int w = 0; w < 6; w++) L[w] = p[w];
            L[6] = p[12];
        } else {
            for (int w = 0; w < 3; w++) {
                L[w] = kLoId;
                L[3 + w] = kHiId;
            }
            L[6] = 0;
        }
        for (int w = 0; w < 7; w++) R[w] = L[w];
        for (int s = 1; s < 32; s <<= 1) {
            for (int w = 0; w < 7; w++) {
                uint32_t tl = __shfl_up(L[w], s, 32), tr = __shfl_down(R[w], s, 32);
                if (i >= (uint32_t)s) L[w] = w == 6 ? L[w] + tl : (w < 3 ? min(L[w], tl) : max(L[w], tl));
                if (i + s < 32) R[w] = w == 6 ? R[w] + tr : (w < 3 ? min(R[w], tr) : max(R[w], tr));
            }
        }
        // Split after bin i: left = bins [0, i], right = bins [i + 1, 32).
        uint32_t N[7];
        for (int w = 0; w < 7; w++) N[w] = __shfl_down(R[w], 1, 32);
        const bool valid = axv && i < 31 && L[6] > 0 && N[6] > 0;
        float llo[3], lhi[3], rlo[3], rhi[3];
        for (int k = 0; k < 3; k++) {
            llo[k] = dec(L[k]);
            lhi[k] = dec(L[3 + k]);
            rlo[k] = dec(N[k]);
            rhi[k] = dec(N[3 + k]);
        }
        float cost = box_area(llo, lhi) * (float)L[6] + box_area(rlo, rhi) * (float)N[6];
        if (valid && cost < INFINITY) {
            unsigned long long key = ((unsigned long long)__float_as_uint(cost) << 32) | (ax * 32 + i);
            best = key < best ? key : best;
        }
    }
    for (int m = 32; m >= 1; m >>= 1) {
        unsigned long long o = shfl_xor_u64(best, m);
        best = o < best ? o : best;
    }
    float best_cost = INFINITY;
    int best_axis = -1;
    uint32_t best_split = 0;
    if (best != ~0ull) {
        best_cost = __uint_as_float((uint32_t)(best >> 32));
        best_axis = (int)((uint32_t)best / 32u);
        best_split = (uint32_t)best % 32u + 1u;
    }
    float parent_area = box_area(f.blo, f.bhi);
    float leaf_cost = kTriCost * (float)n;
    float split_cost = parent_area > 0.0f ? kTravCost + kTriCost * best_cost / parent_area : INFINITY;
    if (n <= kMaxLeaf && !(split_cost < leaf_cost)) return d;
    if (best_axis >= 0) {
        d.kind = 1;
        d.axis = (uint32_t)best_axis;
        d.split = best_split;
    } else if (n > kHardLeaf) {
        d.kind = 2;
    }
    return d;
}

// Lanes 0..12 (left side) and 16..28 (right side) each fold one bin word of
// the split axis over their side's bins; returns that word (others: 0).
__device__ inline uint32_t side_word(const uint32_t *sb, uint32_t axis, uint32_t split, uint32_t lane) {
    const uint32_t w = lane & 15, side = lane >> 4;
    if (w >= kBinWords || side > 1) return 0;
    uint32_t v = identity(w);
    const uint32_t from = side ? split : 0, to = side ? kBins : split;
    for (uint32_t j = from; j < to; j++) v = combine(w, v, sb[(axis * kBins + j) * kBinWords + w]);
    return v;
}

// Writes both children of a split node (lanes of one wave): boxes and
// centroid boxes from the bins (binned split) or identity boxes and the
// parent's centroid box (median; boxes are reduced afterwards).
// mbox: the halves' boxes of a median split (left 6 words, right 6), or
// null to leave identity boxes for a later reduction.
__device__ void write_children(TNode *nodes, uint32_t parent, const TNode &pn, uint32_t c, const Decision &d,
                               const uint32_t *sb, const uint32_t *mbox, uint32_t m, uint32_t lane) {
    uint32_t *L = reinterpret_cast<uint32_t *>(&nodes[c]);
    uint32_t *R = reinterpret_cast<uint32_t *>(&nodes[c + 1]);
    if (d.kind == 1) {
        uint32_t v = side_word(sb, d.axis, d.split, lane);
        const uint32_t w = lane & 15, side = lane >> 4;
        if (w < 12 && side <= 1) (side ? R : L)[w] = v;
    } else if (lane < 12) {
        L[lane] = lane < 6 ? (mbox ? mbox[lane] : identity(lane)) : pick6(pn.cb, lane - 6);
        R[lane] = lane < 6 ? (mbox ? mbox[6 + lane] : identity(lane)) : pick6(pn.cb, lane - 6);
    }
    const uint32_t depth = (pn.depth & ~kInner) + 1;
    if (lane == 32) {
        L[12] = pn.b;
        L[13] = m;
        L[14] = parent << 1;
        L[15] = depth;
    } else if (lane == 33) {
        R[12] = pn.b + m;
        R[13] = pn.n - m;
        R[14] = parent << 1 | 1u;
        R[15] = depth;
    } else if (lane == 34) {
        nodes[parent].depth = (pn.depth & ~kInner) | kInner;
    }
}

// ---- 2. large nodes, one level per step ----------------------------------

__global__ void k_clear_bins(uint32_t *bins, uint32_t words) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words) bins[i] = identity(i % kBinWords);
}

__global__ __launch_bounds__(kLargeBlock) void k_bin_large(const uint32_t *open, const TNode *nodes, const uint32_t *ids,
                                                           const float4 *tlo, const float4 *thi, const float *tcz,
                                                           uint32_t *bins) {
    __shared__ uint32_t sb[kNodeBins];
    const uint32_t k = blockIdx.y;
    const TNode &nd = nodes[open[k]];
    const uint32_t b = nd.b, n = nd.n, start = blockIdx.x * kChunk;
    if (start >= n) return;
    const uint32_t end = min(n, start + kChunk);
    for (uint32_t w = threadIdx.x; w < kNodeBins; w += blockDim.x) sb[w] = identity(w % kBinWords);
    float clo[3], scale[3];
    bool ok[3];
    for (int a = 0; a < 3; a++) {
        clo[a] = dec(nd.cb[a]);
        const float ext = dec(nd.cb[3 + a]) - clo[a];
        ok[a] = ext > 0.0f;
        scale[a] = (float)kBins / ext;
    }
    __syncthreads();
    for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x) {
        const uint32_t t = ids[b + i];
        const float4 l = tlo[t], h = thi[t];
        const float c[3] = {l.w, h.w, tcz[t]};
        const uint32_t e[6] = {enc(l.x), enc(l.y), enc(l.z), enc(h.x), enc(h.y), enc(h.z)};
        for (int a = 0; a < 3; a++) {
            if (!ok[a]) continue;
            uint32_t *p = &sb[(a * kBins + bin_of(c[a], clo[a], scale[a])) * kBinWords];
            atomicMin(&p[0], e[0]);
            atomicMin(&p[1], e[1]);
            atomicMin(&p[2], e[2]);
            atomicMax(&p[3], e[3]);
            atomicMax(&p[4], e[4]);
            atomicMax(&p[5], e[5]);
            for (int q = 0; q < 3; q++)
                if (c[q] == c[q]) {
                    atomicMin(&p[6 + q], enc(c[q]));
                    atomicMax(&p[9 + q], enc(c[q]));
                }
            atomicAdd(&p[12], 1u);
        }
    }
    __syncthreads();
    uint32_t *g = bins + (size_t)k * kNodeBins;
    for (uint32_t w = threadIdx.x; w < kNodeBins; w += blockDim.x) {
        const uint32_t v = sb[w], ww = w % kBinWords;
        if (v == identity(ww)) continue;
        if (ww == 12) atomicAdd(&g[w], v);
        else if (lo_word(ww)) atomicMin(&g[w], v);
        else atomicMax(&g[w], v);
    }
}

// One wave per open node: decide, create and route the children.
// dec[k] = (kind | axis << 8, split, left count, first child).
__global__ __launch_bounds__(64) void k_split_large(const uint32_t *open, TNode *nodes, const uint32_t *bins,
                                                    uint4 *decs, uint32_t *next_open, uint32_t *small,
                                                    uint32_t *counters) {
    __shared__ uint32_t sb[kNodeBins];
    __shared__ uint32_t sh[2];
    const uint32_t k = blockIdx.x, lane = threadIdx.x, v = open[k];
    const TNode pn = nodes[v];
    for (uint32_t w = lane; w < kNodeBins; w += 64) sb[w] = bins[(size_t)k * kNodeBins + w];
    __syncthreads();
    const Decision d = decide(sb, decode(pn.box, pn.cb), pn.n, pn.depth & ~kInner);
    if (d.kind == 0) {
        if (lane == 0) decs[k] = make_uint4(0, 0, 0, 0);
        return;
    }
    uint32_t m = pn.n / 2;
    if (d.kind == 1) {
        // Left count: the count word folded over bins [0, split).
        const uint32_t v12 = side_word(sb, d.axis, d.split, lane);
        if (lane == 12) sh[1] = v12;
    }
    if (lane == 0) sh[0] = atomicAdd(&counters[C_NODES], 2u);
    __syncthreads();
    const uint32_t c = sh[0];
    if (d.kind == 1) m = sh[1];
    write_children(nodes, v, pn, c, d, sb, nullptr, m, lane);
    if (lane == 0) {
        decs[k] = make_uint4(d.kind | d.axis << 8, d.split, m, c);
        atomicAdd(&counters[C_INNER], 1u);
        if (d.kind == 2) counters[C_MEDIAN] = 1;
        const uint32_t depth = (pn.depth & ~kInner) + 1;
        atomicMax(&counters[C_DEPTH], depth);
        const uint32_t cn[2] = {m, pn.n - m};
        for (uint32_t s = 0; s < 2; s++) {
            if (cn[s] > kBuildSmall) {
                next_open[atomicAdd(&counters[C_NEXT_OPEN], 1u)] = c + s;
                atomicMax(&counters[C_NEXT_MAXN], cn[s]);
            } else {
                small[atomicAdd(&counters[C_SMALL], 1u)] = c + s;
            }
        }
    }
}

__global__ __launch_bounds__(kLargeBlock) void k_part_count(const uint32_t *open, const TNode *nodes,
                                                            const uint4 *decs, const uint32_t *ids, const float4 *tlo,
                                                            const float4 *thi, const float *tcz, uint32_t *chunk_left,
                                                            uint32_t max_chunks) {
    __shared__ uint32_t lw[16];
    const uint32_t k = blockIdx.y;
    const uint4 d = decs[k];
    if ((d.x & 0xff) != 1) return;
    const TNode &nd = nodes[open[k]];
    const uint32_t b = nd.b, n = nd.n, start = blockIdx.x * kChunk;
    if (start >= n) return;
    const uint32_t end = min(n, start + kChunk), ax = d.x >> 8;
    const float lo = dec(nd.cb[ax]), scale = (float)kBins / (dec(nd.cb[3 + ax]) - lo);
    uint32_t cnt = 0;
    for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x)
        cnt += bin_of(coord(tlo, thi, tcz, ids[b + i], ax), lo, scale) < (int)d.y ? 1u : 0u;
    uint32_t total;
    block_exclusive(cnt, lw, &total);
    if (threadIdx.x == 0) chunk_left[(size_t)k * max_chunks + blockIdx.x] = total;
}

__global__ __launch_bounds__(64) void k_part_scan(const uint32_t *open, const TNode *nodes, const uint4 *decs,
                                                  uint32_t *chunk_left, uint32_t max_chunks) {
    const uint32_t k = blockIdx.x, lane = threadIdx.x;
    if ((decs[k].x & 0xff) != 1) return;
    const uint32_t nc = (nodes[open[k]].n + kChunk - 1) / kChunk;
    uint32_t *a = chunk_left + (size_t)k * max_chunks;
    uint32_t run = 0;
    for (uint32_t base = 0; base < nc; base += 64) {
        const uint32_t i = base + lane, v = i < nc ? a[i] : 0u;
        uint32_t x = v;
        for (int s = 1; s < 64; s <<= 1) {
            uint32_t t = __shfl_up(x, s, 64);
            if (lane >= (uint32_t)s) x += t;
        }
        if (i < nc) a[i] = run + x - v;
        run += __shfl(x, 63, 64);
    }
}

__global__ __launch_bounds__(kLargeBlock) void k_part_scatter(const uint32_t *open, const TNode *nodes,
                                                              const uint4 *decs, const uint32_t *ids, uint32_t *ids2,
                                                              const float4 *tlo, const float4 *thi, const float *tcz,
                                                              const uint32_t *chunk_left, uint32_t max_chunks) {
    __shared__ uint32_t lw[16];
    const uint32_t k = blockIdx.y;
    const uint4 d = decs[k];
    if ((d.x & 0xff) != 1) return;
    const TNode &nd = nodes[open[k]];
    const uint32_t b = nd.b, n = nd.n, start = blockIdx.x * kChunk;
    if (start >= n) return;
    const uint32_t ax = d.x >> 8, m = d.z;
    const float lo = dec(nd.cb[ax]), scale = (float)kBins / (dec(nd.cb[3 + ax]) - lo);
    const uint32_t first = start + threadIdx.x * kPerThread;
    uint32_t t[kPerThread];
    uint32_t flags = 0, cnt = 0;
    for (uint32_t j = 0; j < kPerThread; j++) {
        const uint32_t i = first + j;
        t[j] = 0;
        if (i < n) {
            t[j] = ids[b + i];
            if (bin_of(coord(tlo, thi, tcz, t[j], ax), lo, scale) < (int)d.y) {
                flags |= 1u << j;
                cnt++;
            }
        }
    }
    uint32_t total;
    uint32_t lb = chunk_left[(size_t)k * max_chunks + blockIdx.x] + block_exclusive(cnt, lw, &total);
    for (uint32_t j = 0; j < kPerThread; j++) {
        const uint32_t i = first + j;
        if (i >= n) break;
        // The flags repeat the binning's bin_of, so lb < m for every left
        // id and m + (i - lb) < n for every right one; the guard only keeps
        // writes inside the node's range.
        const uint32_t dst = (flags >> j & 1u) ? lb++ : m + (i - lb);
        if (dst < n) ids2[b + dst] = t[j];
    }
}

__global__ __launch_bounds__(kLargeBlock) void k_copy_back(const uint32_t *open, const TNode *nodes, const uint4 *decs,
                                                           uint32_t *ids, const uint32_t *ids2) {
    const uint32_t k = blockIdx.y;
    if ((decs[k].x & 0xff) != 1) return;
    const TNode &nd = nodes[open[k]];
    const uint32_t b = nd.b, n = nd.n, start = blockIdx.x * kChunk;
    const uint32_t end = min(n, start + kChunk);
    for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x) ids[b + i] = ids2[b + i];
}

// Boxes of the two halves of every object-median split of the level.
__global__ __launch_bounds__(kLargeBlock) void k_median_bounds(const uint32_t *open, const TNode *nodes_c,
                                                               const uint4 *decs, const uint32_t *ids,
                                                               const float4 *tlo, const float4 *thi, TNode *nodes) {
    const uint32_t k = blockIdx.y;
    const uint4 d = decs[k];
    if ((d.x & 0xff) != 2) return;
    const TNode &nd = nodes_c[open[k]];
    const uint32_t b = nd.b, n = nd.n, m = d.z, start = blockIdx.x * kChunk;
    if (start >= n) return;
    const uint32_t end = min(n, start + kChunk);
    uint32_t v[2][6];
    for (int s = 0; s < 2; s++)
        for (uint32_t w = 0; w < 6; w++) v[s][w] = identity(w);
    for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x) {
        const uint32_t t = ids[b + i], s = i >= m ? 1u : 0u;
        const float4 l = tlo[t], h = thi[t];
        const uint32_t e[6] = {enc(l.x), enc(l.y), enc(l.z), enc(h.x), enc(h.y), enc(h.z)};
        for (uint32_t w = 0; w < 6; w++) {
            const uint32_t a = s ? v[1][w] : v[0][w];
            const uint32_t r = w < 3 ? min(a, e[w]) : max(a, e[w]);
            if (s) v[1][w] = r;
            else v[0][w] = r;
        }
    }
    for (int s = 0; s < 2; s++)
        for (uint32_t w = 0; w < 6; w++) {
            uint32_t x = v[s][w];
            for (int q = 32; q >= 1; q >>= 1) {
                uint32_t o = __shfl_xor(x, q);
                x = w < 3 ? min(x, o) : max(x, o);
            }
            if ((threadIdx.x & 63) == 0 && x != identity(w)) {
                uint32_t *dst = &nodes[d.w + s].box[w];
                if (w < 3) atomicMin(dst, x);
                else atomicMax(dst, x);
            }
        }
}

// ---- 3. small subtrees, one wave each -------------------------------------

// A pending node of the one-wave build: its record in TNode word order
// (box, centroid box, b, n, parent, depth).
constexpr uint32_t kRec = 16;

__global__ __launch_bounds__(64) void k_small(const uint32_t *small, TNode *nodes, uint32_t *ids, const float4 *tlo,
                                              const float4 *thi, const float *tcz, uint32_t *counters) {
    __shared__ uint32_t s_ids[kBuildSmall];
    __shared__ uint32_t sb[kNodeBins];            // bins, then the partition's scratch
    __shared__ uint32_t stk[kStackCap * kRec];    // pending node records
    __shared__ uint32_t sh[4];
    const uint32_t lane = threadIdx.x;
    const uint32_t root = small[blockIdx.x];
    const uint32_t b0 = nodes[root].b, n0 = nodes[root].n;
    for (uint32_t i = lane; i < n0; i += 64) s_ids[i] = ids[b0 + i];
    // The subtree's nodes: at most 2 n0 - 2 below the root, reserved at once.
    if (lane == 0) sh[0] = n0 >= 2 ? atomicAdd(&counters[C_NODES], 2 * n0 - 2) : 0u;
    if (lane < kRec) stk[lane] = reinterpret_cast<const uint32_t *>(&nodes[root])[lane];
    __syncthreads();
    const uint32_t base = sh[0], reserved = n0 >= 2 ? 2 * n0 - 2 : 0u;
    uint32_t used = 0, splits = 0, max_depth = 0, sp = 1;
    uint32_t node = root;  // temp index of the record on top of the stack
    __shared__ uint32_t s_node[kStackCap];
    if (lane == 0) s_node[0] = root;
    __syncthreads();
    while (sp > 0) {
        sp--;
        const uint32_t *rec = &stk[sp * kRec];
        node = s_node[sp];
        uint32_t box[6], cb[6];
        for (int w = 0; w < 6; w++) {
            box[w] = rec[w];
            cb[w] = rec[6 + w];
        }
        const uint32_t nb = rec[12], n = rec[13], depth = rec[15] & ~kInner, rb = nb - b0;
        max_depth = max(max_depth, depth);
        __syncthreads();  // the record is read before anything overwrites its slot
        if (n <= 1 || depth >= kMaxDepth) continue;
        for (uint32_t w = lane; w < kNodeBins; w += 64) sb[w] = identity(w % kBinWords);
        const NodeF f = decode(box, cb);
        float scale[3];
        bool ok[3];
        for (int a = 0; a < 3; a++) {
            const float ext = f.chi[a] - f.clo[a];
            ok[a] = ext > 0.0f;
            scale[a] = (float)kBins / ext;
        }
        __syncthreads();
        for (uint32_t i = lane; i < n; i += 64) {
            const uint32_t t = s_ids[rb + i];
            const float4 l = tlo[t], h = thi[t];
            const float c[3] = {l.w, h.w, tcz[t]};
            const uint32_t e[6] = {enc(l.x), enc(l.y), enc(l.z), enc(h.x), enc(h.y), enc(h.z)};
            for (int a = 0; a < 3; a++) {
                if (!ok[a]) continue;
                uint32_t *p = &sb[(a * kBins + bin_of(c[a], f.clo[a], scale[a])) * kBinWords];
                atomicMin(&p[0], e[0]);
                atomicMin(&p[1], e[1]);
                atomicMin(&p[2], e[2]);
                atomicMax(&p[3], e[3]);
                atomicMax(&p[4], e[4]);
                atomicMax(&p[5], e[5]);
                for (int q = 0; q < 3; q++)
                    if (c[q] == c[q]) {
                        atomicMin(&p[6 + q], enc(c[q]));
                        atomicMax(&p[9 + q], enc(c[q]));
                    }
                atomicAdd(&p[12], 1u);
            }
        }
        __syncthreads();
        const Decision d = decide(sb, f, n, depth);
        if (d.kind == 0) continue;
        const uint32_t c = base + used;  // left child; right child c + 1
        used += 2;
        splits++;
        // Child records: lanes 0..12 / 16..28 fold the left / right bins (binned
        // split), or the halves' boxes are reduced over their triangles (median).
        uint32_t word = 0;
        const uint32_t w = lane & 15, side = lane >> 4;
        uint32_t m = n / 2;
        if (d.kind == 1) {
            word = side_word(sb, d.axis, d.split, lane);
            if (lane == 12) sh[1] = word;
        } else {
            uint32_t bx[2][6];
            for (int s2 = 0; s2 < 2; s2++)
                for (uint32_t q = 0; q < 6; q++) bx[s2][q] = identity(q);
            for (uint32_t i = lane; i < n; i += 64) {
                const uint32_t t = s_ids[rb + i];
                const float4 l = tlo[t], h = thi[t];
                const uint32_t e[6] = {enc(l.x), enc(l.y), enc(l.z), enc(h.x), enc(h.y), enc(h.z)};
                const bool right = i >= m;
                for (uint32_t q = 0; q < 6; q++) {
                    const uint32_t a = right ? bx[1][q] : bx[0][q];
                    const uint32_t r = q < 3 ? min(a, e[q]) : max(a, e[q]);
                    if (right) bx[1][q] = r;
                    else bx[0][q] = r;
                }
            }
            for (int s2 = 0; s2 < 2; s2++)
                for (uint32_t q = 0; q < 6; q++) {
                    uint32_t x = bx[s2][q];
                    for (int o = 32; o >= 1; o >>= 1) {
                        uint32_t y = __shfl_xor(x, o);
                        x = q < 3 ? min(x, y) : max(x, y);
                    }
                    if (side == (uint32_t)s2 && w == q) word = x;
                }
            if (w >= 6 && w < 12 && side <= 1) word = pick6(cb, w - 6);
        }
        __syncthreads();
        if (d.kind == 1) m = sh[1];
        // Push right (below) and left (on top): records in LDS, and the
        // temp nodes in global memory for the numbering.
        uint32_t *R = &stk[sp * kRec], *L = &stk[(sp + 1) * kRec];
        uint32_t *gL = reinterpret_cast<uint32_t *>(&nodes[c]), *gR = reinterpret_cast<uint32_t *>(&nodes[c + 1]);
        if (w < 12 && side <= 1) {
            (side ? R : L)[w] = word;
            (side ? gR : gL)[w] = word;
        }
        if (lane >= 32 && lane < 40) {
            const uint32_t q = lane - 32, s2 = q >> 2, k = 12 + (q & 3);
            const uint32_t v = k == 12 ? (s2 ? nb + m : nb)
                             : k == 13 ? (s2 ? n - m : m)
                             : k == 14 ? (node << 1 | s2)
                                       : depth + 1;
            (s2 ? R : L)[k] = v;
            (s2 ? gR : gL)[k] = v;
        }
        if (lane == 40) nodes[node].depth = depth | kInner;
        if (lane == 41) {
            s_node[sp] = c + 1;
            s_node[sp + 1] = c;
        }
        if (d.kind == 1) {
            // Ballot stable partition of [rb, rb + n) through the bin area.
            const float lo = pick3(f.clo, d.axis), sc = pick3(scale, d.axis);
            uint32_t lcount = 0;
            for (uint32_t b2 = 0; b2 < n; b2 += 64) {
                const uint32_t i = b2 + lane;
                const bool in = i < n;
                const uint32_t t = in ? s_ids[rb + i] : 0u;
                const bool fl = in && bin_of(coord(tlo, thi, tcz, t, d.axis), lo, sc) < (int)d.split;
                const unsigned long long mask = __ballot(fl);
                const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
                if (in) {
                    const uint32_t lb = lcount + rank, dst = fl ? lb : m + (i - lb);
                    if (dst < n) sb[dst] = t;
                }
                lcount += (uint32_t)__popcll(mask);
            }
            __syncthreads();
            for (uint32_t i = lane; i < n; i += 64) s_ids[rb + i] = sb[i];
        }
        sp += 2;
        __syncthreads();
    }
    __syncthreads();
    for (uint32_t i = lane; i < n0; i += 64) ids[b0 + i] = s_ids[i];
    // Reserved slots left unused: empty records the numbering skips.
    for (uint32_t i = used + lane; i < reserved; i += 64) {
        nodes[base + i].n = 0;
        nodes[base + i].depth = 0;
        nodes[base + i].parent = kNone;
    }
    for (int q = 32; q >= 1; q >>= 1) max_depth = max(max_depth, __shfl_xor(max_depth, q));
    if (lane == 0) {
        atomicMax(&counters[C_DEPTH], max_depth);
        if (splits) atomicAdd(&counters[C_INNER], splits);
    }
}

// ---- scans ----------------------------------------------------------------

__global__ __launch_bounds__(256) void k_scan_local(const uint32_t *in, uint32_t *out, uint32_t *sums, uint32_t n) {
    __shared__ uint32_t lw[16];
    const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * 4;
    uint32_t v[4], s = 0;
    for (int j = 0; j < 4; j++) {
        v[j] = base + j < n ? in[base + j] : 0u;
        s += v[j];
    }
    uint32_t total;
    uint32_t run = block_exclusive(s, lw, &total);
    for (int j = 0; j < 4; j++) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t *sums, uint32_t nb) {
    __shared__ uint32_t lw[16];
    constexpr uint32_t kPer = 16;
    const uint32_t base = threadIdx.x * kPer;
    uint32_t v[kPer], s = 0;
    for (uint32_t j = 0; j < kPer; j++) {
        v[j] = base + j < nb ? sums[base + j] : 0u;
        s += v[j];
    }
    uint32_t total;
    uint32_t run = block_exclusive(s, lw, &total);
    for (uint32_t j = 0; j < kPer; j++) {
        if (base + j < nb) sums[base + j] = run;
        run += v[j];
    }
}

__global__ void k_scan_add(uint32_t *out, const uint32_t *sums, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += sums[i / kScanBlock];
}

// ---- 4. depth-first numbering and the node words --------------------------

__global__ void k_fill(uint32_t *a, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_rank_count(const TNode *nodes, uint32_t T, uint32_t *cnt, uint32_t *d0) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= T) return;
    const TNode &nd = nodes[v];
    if (nd.n == 0 || !(nd.depth & kInner)) return;
    atomicAdd(&cnt[nd.b], 1u);
    atomicMin(&d0[nd.b], nd.depth & ~kInner);
}

__global__ void k_rank(const TNode *nodes, uint32_t T, const uint32_t *base, const uint32_t *d0, uint32_t *rank) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= T) return;
    const TNode &nd = nodes[v];
    if (nd.n == 0 || !(nd.depth & kInner)) return;
    rank[v] = base[nd.b] + (nd.depth & ~kInner) - d0[nd.b];
}

// Node words: lo = (bmin, first), hi = (bmax, count); node 1 stays unused.
__global__ void k_emit_nodes(const TNode *nodes, uint32_t T, uint32_t node_count, const uint32_t *rank, uint32_t *words,
                             uint32_t *bstart) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v == 0) {
        const uint32_t dummy[8] = {__float_as_uint(INFINITY), __float_as_uint(INFINITY), __float_as_uint(INFINITY), 0u,
                                   __float_as_uint(-INFINITY), __float_as_uint(-INFINITY), __float_as_uint(-INFINITY),
                                   0u};
        for (int w = 0; w < 8; w++) words[8 + w] = dummy[w];
        bstart[1] = 0;
    }
    if (v >= T) return;
    const TNode &nd = nodes[v];
    if (nd.n == 0) return;  // a reserved slot a subtree did not use
    const uint32_t idx = nd.parent == kNone ? 0u : 2u + 2u * rank[nd.parent >> 1] + (nd.parent & 1u);
    if (idx >= node_count) return;  // ranks are < inner count by construction
    const bool inner = (nd.depth & kInner) != 0;
    uint32_t *o = words + 8 * (size_t)idx;
    for (int k = 0; k < 3; k++) {
        o[k] = __float_as_uint(dec(nd.box[k]));
        o[4 + k] = __float_as_uint(dec(nd.box[3 + k]));
    }
    o[3] = inner ? 2u + 2u * rank[v] : nd.b;
    o[7] = inner ? 0u : nd.n;
    bstart[idx] = nd.b;
}

// ---- gather: triangles and attributes in leaf-slot order -------------------

__global__ void k_gather(SoupDev s, const uint32_t *ids, float4 *tris, float4 *attrs) {
    const uint32_t sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= s.n) return;
    const size_t g = ids[sl];
    const float *p = s.pos + 9 * g;
    const uint32_t ws[3] = {(uint32_t)g, s.light_of ? s.light_of[g] : OXPT_NONE, s.material[g]};
    float4 *gw = tris + 3 * (size_t)sl;
    for (int v = 0; v < 3; v++) gw[v] = make_float4(p[3 * v], p[3 * v + 1], p[3 * v + 2], __uint_as_float(ws[v]));
    const uint32_t fl = s.flags[g];
    float4 *a = attrs + 12 * (size_t)sl;
    for (int v = 0; v < 3; v++) {
        const float w = v == 0 ? __uint_as_float(fl) : 0.0f;
        if (fl & OXPT_F_NORMALS) {
            const float *q = s.normals + 9 * g + 3 * v;
            a[v] = make_float4(q[0], q[1], q[2], w);
        } else {
            a[v] = make_float4(0.0f, 0.0f, 0.0f, w);
        }
    }
    for (int v = 0; v < 3; v++) {
        if (fl & OXPT_F_TANGENTS) {
            const float *q = s.tangents + 12 * g + 4 * v;
            a[3 + v] = make_float4(q[0], q[1], q[2], q[3]);
        } else {
            a[3 + v] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
    for (int v = 0; v < 3; v++) {
        float4 u = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (fl & OXPT_F_UV0) {
            u.x = s.uv0[6 * g + 2 * v];
            u.y = s.uv0[6 * g + 2 * v + 1];
        }
        if (fl & OXPT_F_UV1) {
            u.z = s.uv1[6 * g + 2 * v];
            u.w = s.uv1[6 * g + 2 * v + 1];
        }
        a[6 + v] = u;
    }
    for (int v = 0; v < 3; v++) {
        if (fl & OXPT_F_COLORS) {
            const float *q = s.colors + 12 * g + 4 * v;
            a[9 + v] = make_float4(q[0], q[1], q[2], q[3]);
        } else {
            a[9 + v] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        }
    }
}

__global__ void k_emissive_slots(const uint32_t *ids, const uint32_t *light_of, uint32_t n, uint32_t n_emissive,
                                 uint32_t *out) {
    const uint32_t sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= n) return;
    const uint32_t li = light_of[ids[sl]];
    if (li < n_emissive) out[li] = sl;
}

// ---- emissive lights (EmissiveLights::build) --------------------------------

// Weight of triangle t: area (f64 cross product) x its material's emission
// luminance, the host packer's arithmetic; flag = weight > 0.
__global__ void k_emit_weight(const float *pos, const uint32_t *material, const float *mat_lum, uint32_t n_mat,
                              uint32_t n, double *wgt, uint32_t *flag) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t m = material[t];
    const float lum = m < n_mat ? mat_lum[m] : 0.0f;
    double w = 0.0;
    if (lum != 0.0f) {
        const float *p = pos + 9 * (size_t)t;
        double e1[3], e2[3];
        for (int k = 0; k < 3; k++) {
            e1[k] = (double)p[3 + k] - p[k];
            e2[k] = (double)p[6 + k] - p[k];
        }
        const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                     cz = e1[0] * e2[1] - e1[1] * e2[0];
        const double area = 0.5 * __builtin_sqrt(cx * cx + cy * cy + cz * cz);
        w = area * (double)lum;
    }
    wgt[t] = w;
    flag[t] = w > 0.0 ? 1u : 0u;
}

// rank: exclusive prefix of the flags. Light index = rank, in input order.
__global__ void k_light_of(const uint32_t *flag, const uint32_t *rank, const double *wgt, uint32_t n,
                           uint32_t *light_of, double *wgt_c, uint32_t *tri_of, uint32_t *count) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (flag[t]) {
        light_of[t] = rank[t];
        wgt_c[rank[t]] = wgt[t];
        tri_of[rank[t]] = t;
    } else {
        light_of[t] = OXPT_NONE;
    }
    if (t == n - 1) *count = rank[t] + flag[t];
}

// Records {cdf, pmf, triangle, slot}: the CDF over the weights in light
// order, sums in that order (one thread), the last entry exactly 1.
__global__ void k_emissive_records(const double *wgt_c, const uint32_t *tri_of, const uint32_t *slots, uint32_t n,
                                   uint32_t *rec) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double total = 0.0;
    for (uint32_t i = 0; i < n; i++) total += wgt_c[i];
    double run = 0.0;
    float prev = 0.0f;
    for (uint32_t i = 0; i < n; i++) {
        run += wgt_c[i];
        const float cdf = i + 1 == n ? 1.0f : (float)(run / total);
        rec[4 * i] = __float_as_uint(cdf);
        rec[4 * i + 1] = __float_as_uint(cdf - prev);
        rec[4 * i + 2] = tri_of[i];
        rec[4 * i + 3] = slots[i];
        prev = cdf;
    }
}

// ---- 5. 4-wide collapse ------------------------------------------------------

__device__ inline float wide_area(const uint32_t *words, uint32_t node) {
    float d[3];
    for (int k = 0; k < 3; k++) {
        const float x = __uint_as_float(words[8 * (size_t)node + 4 + k]) - __uint_as_float(words[8 * (size_t)node + k]);
        d[k] = x < 0.0f ? 0.0f : x;  // std::max(x, 0.0f)
    }
    return d[0] * d[1] + d[1] * d[2] + d[2] * d[0];
}

// One thread per binary node that roots a wide node of this level.
// max_w: capacity of the wide arrays and level lists (the binary inner
// node count); a tree that would exceed it is flagged, never written past.
__global__ void k_wide_level(const uint32_t *words, uint32_t node_count, const uint32_t *bstart, const uint32_t *cur,
                             uint32_t n_cur, uint32_t depth, uint32_t root_leaf, uint32_t max_w, uint32_t *next,
                             uint32_t *wkids, uint32_t *wdepth, uint32_t *wb, uint32_t *wide_of_bin,
                             uint32_t *counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_cur) return;
    const uint32_t s = cur[i];
    if (s >= node_count) {
        counters[C_WIDE_BAD] = 2;
        return;
    }
    uint32_t kids[4] = {kNone, kNone, kNone, kNone}, nk;
    if (root_leaf) {
        kids[0] = s;
        nk = 1;
    } else {
        const uint32_t f = words[8 * (size_t)s + 3];
        if (f < 2 || f + 1 >= node_count) {  // an inner node's children are a pair after the root
            counters[C_WIDE_BAD] = 2;
            return;
        }
        kids[0] = f;
        kids[1] = f + 1;
        nk = 2;
    }
    while (nk < 4) {
        int best = -1;
        float ba = -1.0f;
        for (uint32_t j = 0; j < nk; j++) {
            if (words[8 * (size_t)kids[j] + 7] != 0) continue;
            const float a = wide_area(words, kids[j]);
            if (a > ba) {
                ba = a;
                best = (int)j;
            }
        }
        if (best < 0) break;
        const uint32_t f = words[8 * (size_t)kids[best] + 3];
        if (f < 2 || f + 1 >= node_count) {
            counters[C_WIDE_BAD] = 2;
            return;
        }
        for (int j = (int)nk; j > best + 1; j--) kids[j] = kids[j - 1];
        kids[best] = f;
        kids[best + 1] = f + 1;
        nk++;
    }
    const uint32_t w = atomicAdd(&counters[C_WIDE], 1u);
    if (w >= max_w) {
        counters[C_WIDE_BAD] = 2;
        return;
    }
    wdepth[w] = depth;
    wb[w] = bstart[s];
    wide_of_bin[s] = w;
    for (uint32_t j = 0; j < 4; j++) wkids[4 * (size_t)w + j] = kids[j];
    for (uint32_t j = 0; j < nk; j++) {
        const uint32_t c = words[8 * (size_t)kids[j] + 7];
        if (c == 0) {
            const uint32_t at = atomicAdd(&counters[C_WIDE_NEXT], 1u);
            if (at < max_w) next[at] = kids[j];
            else counters[C_WIDE_BAD] = 2;
        } else if (c > kWideLeafMax) {
            counters[C_WIDE_BAD] = 1;
        }
    }
    atomicMax(&counters[C_WIDE_DEPTH], depth);
}

__global__ void k_wide_rank_count(const uint32_t *wb, const uint32_t *wdepth, uint32_t W, uint32_t *cnt, uint32_t *d0) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    atomicAdd(&cnt[wb[w]], 1u);
    atomicMin(&d0[wb[w]], wdepth[w]);
}

__global__ void k_wide_rank(const uint32_t *wb, const uint32_t *wdepth, uint32_t W, const uint32_t *base,
                            const uint32_t *d0, uint32_t *rank) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    rank[w] = base[wb[w]] + wdepth[w] - d0[wb[w]];
}

// Wide node = 8 float4: lo.x[4] lo.y[4] lo.z[4] hi.x[4] hi.y[4] hi.z[4]
// ref[4] count[4] (the host collapse's layout, csrc/host/bvh.cpp).
__global__ void k_wide_emit(const uint32_t *words, const uint32_t *wkids, const uint32_t *rank,
                            const uint32_t *wide_of_bin, uint32_t W, float *out) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W || rank[w] >= W) return;
    float *o = out + 32 * (size_t)rank[w];
    for (uint32_t j = 0; j < 4; j++) {
        const uint32_t kid = wkids[4 * (size_t)w + j];
        uint32_t ref = 0, cnt = kEmpty;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        if (kid != kNone) {
            const uint32_t *nw = words + 8 * (size_t)kid;
            for (int k = 0; k < 3; k++) {
                lo[k] = __uint_as_float(nw[k]);
                hi[k] = __uint_as_float(nw[4 + k]);
            }
            cnt = nw[7];
            ref = cnt == 0 ? rank[wide_of_bin[kid]] : nw[3];
        }
        for (int k = 0; k < 3; k++) {
            o[4 * k + j] = lo[k];
            o[4 * (3 + k) + j] = hi[k];
        }
        o[24 + j] = __uint_as_float(ref);
        o[28 + j] = __uint_as_float(cnt);
    }
}

// ---- host driver ---------------------------------------------------------------

inline uint32_t blocks(size_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

struct Scratch {
    const DevAlloc &a;
    std::vector<std::pair<void *, size_t>> held;
    explicit Scratch(const DevAlloc &al) : a(al) {}
    ~Scratch() {
        for (auto &h : held) a.release(h.first, h.second);
    }
    template <typename T>
    int get(size_t count, T **p, const char *what) {
        void *q = nullptr;
        const size_t bytes = std::max<size_t>(count * sizeof(T), 16);
        if (int rc = a.alloc(bytes, &q, what)) return rc;
        held.push_back({q, bytes});
        *p = static_cast<T *>(q);
        return 0;
    }
};

int launch_check(std::string &err, const char *what) {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) return 0;
    err = std::string("hip: device BVH build (") + what + "): " + hipGetErrorString(e);
    return OXPT_E_BACKEND;
}

int read_counters(uint32_t *dev, uint32_t *host, hipStream_t s, std::string &err) {
    hipError_t e = hipMemcpyAsync(host, dev, C_COUNT * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        err = std::string("hip: device BVH build: ") + hipGetErrorString(e);
        return OXPT_E_BACKEND;
    }
    return 0;
}

// Exclusive scan of n (<= kBuildMaxTris) words in place.
int scan(uint32_t *a, uint32_t n, uint32_t *sums, hipStream_t s, std::string &err) {
    const uint32_t nb = blocks(n, kScanBlock);
    k_scan_local<<<nb, 256, 0, s>>>(a, a, sums, n);
    k_scan_sums<<<1, 1024, 0, s>>>(sums, nb);
    k_scan_add<<<blocks(n, 256), 256, 0, s>>>(a, sums, n);
    return launch_check(err, "scan");
}

}  // namespace

int build_geometry(const SoupDev &soup, uint32_t n_emissive, uint32_t *emissive_slots, hipStream_t s,
                   const DevAlloc &al, BuiltGeometry *out, std::string &err) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t N = soup.n;
    if (N == 0 || N > kBuildMaxTris) {
        err = "device BVH build takes 1 to " + std::to_string(kBuildMaxTris) + " triangles";
        return OXPT_E_INVALID;
    }
    Scratch sc(al);
    float4 *tlo, *thi;
    float *tcz;
    uint32_t *ids, *ids2, *counters, *open, *next_open, *small, *bins, *sums;
    TNode *nodes;
    uint4 *decs;
    const size_t max_open = N / (kBuildSmall + 1) + 2;
    const size_t max_nodes = 2 * (size_t)N;
    int rc;
    if ((rc = sc.get(N, &tlo, "BVH build: triangle boxes")) || (rc = sc.get(N, &thi, "BVH build: triangle boxes")) ||
        (rc = sc.get(N, &tcz, "BVH build: triangle boxes")) || (rc = sc.get(N, &ids, "BVH build: ids")) ||
        (rc = sc.get(N, &ids2, "BVH build: ids")) || (rc = sc.get(C_COUNT, &counters, "BVH build: counters")) ||
        (rc = sc.get(max_open, &open, "BVH build: node lists")) ||
        (rc = sc.get(max_open, &next_open, "BVH build: node lists")) ||
        (rc = sc.get(N, &small, "BVH build: node lists")) ||
        (rc = sc.get(max_open * kNodeBins, &bins, "BVH build: bins")) ||
        (rc = sc.get(max_open, &decs, "BVH build: decisions")) ||
        (rc = sc.get(max_nodes, &nodes, "BVH build: nodes")) ||
        (rc = sc.get(blocks(N, kScanBlock) + 1, &sums, "BVH build: scan")))
        return rc;
    uint32_t hc[C_COUNT];

    // Emissive lights found here (before the gather, which writes light_of).
    SoupDev sd = soup;
    uint32_t found = 0, *found_slots = nullptr, *tri_of = nullptr;
    double *wgt_c = nullptr;
    if (soup.mat_lum) {
        double *wgt;
        uint32_t *flag, *rank, *lof;
        if ((rc = sc.get(N, &wgt, "BVH build: light weights")) || (rc = sc.get(N, &wgt_c, "BVH build: light weights")) ||
            (rc = sc.get(N, &flag, "BVH build: lights")) || (rc = sc.get(N, &rank, "BVH build: lights")) ||
            (rc = sc.get(N, &lof, "BVH build: lights")) || (rc = sc.get(N, &tri_of, "BVH build: lights")))
            return rc;
        k_emit_weight<<<blocks(N, 256), 256, 0, s>>>(soup.pos, soup.material, soup.mat_lum, soup.n_mat, N, wgt, flag);
        hipError_t e = hipMemcpyAsync(rank, flag, 4ull * N, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) {
            err = std::string("hip: device light search: ") + hipGetErrorString(e);
            return OXPT_E_BACKEND;
        }
        if ((rc = scan(rank, N, sums, s, err))) return rc;
        k_light_of<<<blocks(N, 256), 256, 0, s>>>(flag, rank, wgt, N, lof, wgt_c, tri_of, counters + C_WIDE_BAD);
        if ((rc = launch_check(err, "light search"))) return rc;
        if ((rc = read_counters(counters, hc, s, err))) return rc;
        found = hc[C_WIDE_BAD];
        sd.light_of = lof;
        if (found > 0 && (rc = sc.get(found, &found_slots, "BVH build: light slots"))) return rc;
    }
    k_prep<<<blocks(N, 256), 256, 0, s>>>(soup.pos, N, tlo, thi, tcz, ids);
    k_init<<<1, 1, 0, s>>>(nodes, N, counters, open, small);
    k_root_bounds<<<std::min(1024u, blocks(N, 256)), 256, 0, s>>>(tlo, thi, tcz, N, nodes);
    if ((rc = launch_check(err, "prepare"))) return rc;

    // Large nodes, level by level.
    uint32_t n_open = N > kBuildSmall ? 1 : 0, max_n = N;
    uint32_t *chunk_left = nullptr;
    size_t chunk_cap = 0;
    while (n_open > 0) {
        const uint32_t max_chunks = blocks(max_n, kChunk);
        const size_t need = (size_t)n_open * max_chunks;
        if (need > chunk_cap) {
            chunk_cap = std::max(need, 2 * chunk_cap);
            if ((rc = sc.get(chunk_cap, &chunk_left, "BVH build: chunk counts"))) return rc;
        }
        const dim3 grid(max_chunks, n_open);
        k_fill<<<1, 64, 0, s>>>(counters + C_NEXT_OPEN, 3, 0u);  // next open, its max size, median flag
        k_clear_bins<<<blocks((size_t)n_open * kNodeBins, 256), 256, 0, s>>>(bins, n_open * kNodeBins);
        k_bin_large<<<grid, kLargeBlock, 0, s>>>(open, nodes, ids, tlo, thi, tcz, bins);
        k_split_large<<<n_open, 64, 0, s>>>(open, nodes, bins, decs, next_open, small, counters);
        k_part_count<<<grid, kLargeBlock, 0, s>>>(open, nodes, decs, ids, tlo, thi, tcz, chunk_left, max_chunks);
        k_part_scan<<<n_open, 64, 0, s>>>(open, nodes, decs, chunk_left, max_chunks);
        k_part_scatter<<<grid, kLargeBlock, 0, s>>>(open, nodes, decs, ids, ids2, tlo, thi, tcz, chunk_left,
                                                    max_chunks);
        k_copy_back<<<grid, kLargeBlock, 0, s>>>(open, nodes, decs, ids, ids2);
        k_median_bounds<<<grid, kLargeBlock, 0, s>>>(open, nodes, decs, ids, tlo, thi, nodes);
        if ((rc = launch_check(err, "large nodes"))) return rc;
        if ((rc = read_counters(counters, hc, s, err))) return rc;
        std::swap(open, next_open);
        n_open = hc[C_NEXT_OPEN];
        max_n = hc[C_NEXT_MAXN];
    }
    if ((rc = read_counters(counters, hc, s, err))) return rc;
    if (hc[C_SMALL] > 0) k_small<<<hc[C_SMALL], 64, 0, s>>>(small, nodes, ids, tlo, thi, tcz, counters);
    if ((rc = launch_check(err, "small subtrees"))) return rc;
    if ((rc = read_counters(counters, hc, s, err))) return rc;
    const uint32_t T = hc[C_NODES];  // temp slots, unused reserved ones included
    const uint32_t inner = hc[C_INNER];
    const uint32_t node_count = 2 + 2 * inner;
    out->bvh_depth = hc[C_DEPTH];

    // Depth-first numbering.
    uint32_t *cnt, *d0, *rank, *words, *bstart;
    if ((rc = sc.get(N, &cnt, "BVH build: ranks")) || (rc = sc.get(N, &d0, "BVH build: ranks")) ||
        (rc = sc.get(T, &rank, "BVH build: ranks")) || (rc = sc.get((size_t)node_count * 8, &words, "BVH build: node words")) ||
        (rc = sc.get(node_count, &bstart, "BVH build: node ranges")))
        return rc;
    k_fill<<<blocks(N, 256), 256, 0, s>>>(cnt, N, 0u);
    k_fill<<<blocks(N, 256), 256, 0, s>>>(d0, N, kNone);
    k_rank_count<<<blocks(T, 256), 256, 0, s>>>(nodes, T, cnt, d0);
    if ((rc = scan(cnt, N, sums, s, err))) return rc;
    k_rank<<<blocks(T, 256), 256, 0, s>>>(nodes, T, cnt, d0, rank);
    k_emit_nodes<<<blocks(T, 256), 256, 0, s>>>(nodes, T, node_count, rank, words, bstart);
    if ((rc = launch_check(err, "numbering"))) return rc;

    // 4-wide collapse, one wide level per step.
    uint32_t *wcur, *wnext, *wkids, *wdepth, *wb, *wide_of_bin;
    const uint32_t max_wide = std::max(inner, 1u);
    if ((rc = sc.get(max_wide, &wcur, "BVH build: wide lists")) ||
        (rc = sc.get(max_wide, &wnext, "BVH build: wide lists")) ||
        (rc = sc.get((size_t)max_wide * 4, &wkids, "BVH build: wide nodes")) ||
        (rc = sc.get(max_wide, &wdepth, "BVH build: wide nodes")) ||
        (rc = sc.get(max_wide, &wb, "BVH build: wide nodes")) ||
        (rc = sc.get(node_count, &wide_of_bin, "BVH build: wide nodes")))
        return rc;
    k_fill<<<1, 64, 0, s>>>(wcur, 1, 0u);  // the binary root roots the wide tree
    uint32_t n_cur = 1, depth = 1;
    const uint32_t root_leaf = inner == 0 ? 1u : 0u;  // T also counts reserved slots left unused
    while (n_cur > 0) {
        if (depth > kMaxDepth + 1) {
            err = "device BVH build: 4-wide collapse deeper than the binary tree";
            return OXPT_E_BACKEND;
        }
        k_fill<<<1, 64, 0, s>>>(counters + C_WIDE_NEXT, 1, 0u);
        k_wide_level<<<blocks(n_cur, 256), 256, 0, s>>>(words, node_count, bstart, wcur, n_cur, depth,
                                                        depth == 1 ? root_leaf : 0u,
                                                        max_wide, wnext, wkids, wdepth, wb, wide_of_bin, counters);
        if ((rc = launch_check(err, "wide levels"))) return rc;
        if ((rc = read_counters(counters, hc, s, err))) return rc;
        if (hc[C_WIDE_BAD] == 2 || hc[C_WIDE] > max_wide || hc[C_WIDE_NEXT] > max_wide) {
            err = "device BVH build: 4-wide collapse exceeds the binary tree's inner nodes";
            return OXPT_E_BACKEND;
        }
        std::swap(wcur, wnext);
        n_cur = hc[C_WIDE_NEXT];
        depth++;
    }
    const uint32_t W = hc[C_WIDE];
    const bool wide_ok = hc[C_WIDE_BAD] == 0 && 3u * hc[C_WIDE_DEPTH] <= kWideStack;

    // Final buffers: binary nodes | triangles | pad | wide nodes.
    out->node_count = node_count;
    out->tri_base = node_count * 2;
    const size_t tri_end = (size_t)out->tri_base + 3 * (size_t)N;
    out->wide_base = wide_ok ? (uint32_t)((tri_end + 7) / 8 * 8) : 0u;
    out->wide_depth = wide_ok ? hc[C_WIDE_DEPTH] : 0u;
    out->geom_vec4 = wide_ok ? out->wide_base + 8 * (size_t)W : tri_end;
    out->attrs_vec4 = 12 * (size_t)N;
    void *g = nullptr, *at = nullptr;
    if ((rc = al.alloc(out->geom_vec4 * 16, &g, "BVH + triangle buffer"))) return rc;
    if ((rc = al.alloc(out->attrs_vec4 * 16, &at, "vertex attribute buffer"))) {
        al.release(g, out->geom_vec4 * 16);
        return rc;
    }
    out->geom = static_cast<float4 *>(g);
    out->attrs = static_cast<float4 *>(at);
    auto fail_out = [&](const char *what, hipError_t e) {
        err = std::string("hip: device BVH build (") + what + "): " + hipGetErrorString(e);
        al.release(out->geom, out->geom_vec4 * 16);
        al.release(out->attrs, out->attrs_vec4 * 16);
        out->geom = out->attrs = nullptr;
        return OXPT_E_BACKEND;
    };
    hipError_t e = hipMemcpyAsync(out->geom, words, (size_t)node_count * 32, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return fail_out("node copy", e);
    if (wide_ok && out->wide_base > tri_end) {
        e = hipMemsetAsync(out->geom + tri_end, 0, (out->wide_base - tri_end) * 16, s);
        if (e != hipSuccess) return fail_out("pad", e);
    }
    k_gather<<<blocks(N, 256), 256, 0, s>>>(sd, ids, out->geom + out->tri_base, out->attrs);
    if (wide_ok) {
        k_fill<<<blocks(N, 256), 256, 0, s>>>(cnt, N, 0u);
        k_fill<<<blocks(N, 256), 256, 0, s>>>(d0, N, kNone);
        k_wide_rank_count<<<blocks(W, 256), 256, 0, s>>>(wb, wdepth, W, cnt, d0);
        std::string serr;
        if (scan(cnt, N, sums, s, serr)) {
            err = serr;
            al.release(out->geom, out->geom_vec4 * 16);
            al.release(out->attrs, out->attrs_vec4 * 16);
            out->geom = out->attrs = nullptr;
            return OXPT_E_BACKEND;
        }
        k_wide_rank<<<blocks(W, 256), 256, 0, s>>>(wb, wdepth, W, cnt, d0, rank);
        k_wide_emit<<<blocks(W, 256), 256, 0, s>>>(words, wkids, rank, wide_of_bin, W,
                                                   reinterpret_cast<float *>(out->geom + out->wide_base));
    }
    uint32_t *records = nullptr;
    if (found > 0) {
        if (sc.get(4ull * found, &records, "BVH build: light records")) {
            err = "hip: device BVH build: light records";
            al.release(out->geom, out->geom_vec4 * 16);
            al.release(out->attrs, out->attrs_vec4 * 16);
            out->geom = out->attrs = nullptr;
            return OXPT_E_BACKEND;
        }
        k_emissive_slots<<<blocks(N, 256), 256, 0, s>>>(ids, sd.light_of, N, found, found_slots);
        k_emissive_records<<<1, 64, 0, s>>>(wgt_c, tri_of, found_slots, found, records);
    } else if (emissive_slots && n_emissive > 0 && soup.light_of) {
        k_emissive_slots<<<blocks(N, 256), 256, 0, s>>>(ids, soup.light_of, N, n_emissive, emissive_slots);
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    out->n_emissive = found;
    out->emissive_records.assign(4ull * found, 0u);
    if (e == hipSuccess && found > 0)
        e = hipMemcpy(out->emissive_records.data(), records, 16ull * found, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_out("emit", e);
    out->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

}  // namespace oxd
