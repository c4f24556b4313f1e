// Binned-SAH BVH over the triangle soup, emitted in the reference's node
// layout (src/shaders/pathtrace.wgsl:638-735): 32-byte nodes
// lo = (bmin.xyz, first), hi = (bmax.xyz, count); count > 0 is a leaf over
// leaf slots [first, first + count), otherwise the children are the
// adjacent pair (first, first + 1); the root is node 0. Siblings are
// adjacent so one traversal step reads one contiguous 64-byte pair.
//
// The reference takes its tree from oxideav-mesh3d's `Bvh` (not vendored);
// only the layout is fixed by the kernel. Images do not depend on the
// topology except through exact-tie order, which the brute-force
// closest-hit test covers.
//
// The tree is a function of the triangle set alone, so the device builder
// (csrc/device/bvh_build.hip) reproduces it word for word:
//  - triangle boxes hold no -0 and no NaN (coordinates are turned from -0
//    into +0, a NaN coordinate is skipped), so box unions do not depend on
//    the order triangles are visited in;
//  - split decisions use only per-bin minima, maxima and counts;
//  - partitions are stable, so every node's triangles stay in ascending id
//    order: leaves list ascending ids, and the object-median fallback (all
//    centroids equal) splits a node at its middle position;
//  - nodes are numbered depth first, left subtree first (each inner node's
//    child pair is appended when the node is visited).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "scene.hpp"

namespace oxpt {
namespace {

constexpr int kBins = 32;
constexpr uint32_t kMaxLeaf = 4;      // SAH may stop at or below this size
constexpr uint32_t kHardLeaf = 8;     // never leave more than this in a leaf ...
constexpr uint32_t kMaxDepth = 56;    // ... unless the depth cap (stack is 64) forces it
constexpr float kTravCost = 1.0f;
constexpr float kTriCost = 1.0f;

// Boxes grow by std::min / std::max, which skip a NaN operand (it compares
// false) and are order-independent on the values they see: triangle boxes
// are built from coordinates with -0 turned into +0 (see build_bvh), and
// centroids of such boxes are never -0.
struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const float *p) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    void grow(const Box &b) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    float area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0 && dy >= 0 && dz >= 0)) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

// Bin of centroid coordinate c in [lo, lo + kBins / scale]; NaN goes to
// bin 0 and overflow to the last bin, the same on every compiler.
inline int bin_of(float c, float lo, float scale) {
    float f = (c - lo) * scale;
    if (!(f > 0.0f)) return 0;
    if (f >= (float)kBins) return kBins - 1;
    return (int)f;
}

struct Node {
    Box box;
    uint32_t first = 0, count = 0;
};

struct Builder {
    const std::vector<Box> &tb;
    const std::vector<float> &cen;  // 3 per tri
    std::vector<uint32_t> &ids;
    std::vector<Node> nodes;
    std::vector<uint32_t> scratch;  // right-hand ids of the partition in progress
    uint32_t depth = 0;

    Builder(const std::vector<Box> &t, const std::vector<float> &c, std::vector<uint32_t> &i)
        : tb(t), cen(c), ids(i), scratch(i.size()) {}

    // Builds the subtree for ids[b, e) into `nodes[ni]`, children appended.
    void build(uint32_t ni, uint32_t b, uint32_t e, uint32_t d) {
        depth = std::max(depth, d);
        Box box, cb;
        for (uint32_t i = b; i < e; i++) {
            box.grow(tb[ids[i]]);
            cb.grow(&cen[3 * (size_t)ids[i]]);
        }
        nodes[ni].box = box;
        uint32_t n = e - b;
        if (n <= 1 || d >= kMaxDepth) {
            nodes[ni].first = b;
            nodes[ni].count = n;
            return;
        }
        // Binned SAH on every axis with a non-degenerate centroid extent.
        float best_cost = INFINITY;
        int best_axis = -1, best_split = 0;
        for (int ax = 0; ax < 3; ax++) {
            float ext = cb.hi[ax] - cb.lo[ax];
            if (!(ext > 0.0f)) continue;
            float scale = (float)kBins / ext;
            Box bins[kBins];
            uint32_t cnt[kBins] = {0};
            for (uint32_t i = b; i < e; i++) {
                uint32_t t = ids[i];
                int bi = bin_of(cen[3 * (size_t)t + ax], cb.lo[ax], scale);
                bins[bi].grow(tb[t]);
                cnt[bi]++;
            }
            float right_area[kBins];
            uint32_t right_cnt[kBins];
            Box acc;
            uint32_t c = 0;
            for (int i = kBins - 1; i > 0; i--) {
                acc.grow(bins[i]);
                c += cnt[i];
                right_area[i] = acc.area();
                right_cnt[i] = c;
            }
            Box lacc;
            uint32_t lc = 0;
            for (int i = 0; i < kBins - 1; i++) {
                lacc.grow(bins[i]);
                lc += cnt[i];
                if (lc == 0 || right_cnt[i + 1] == 0) continue;
                float cost = lacc.area() * (float)lc + right_area[i + 1] * (float)right_cnt[i + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = ax;
                    best_split = i + 1;
                }
            }
        }
        float parent_area = box.area();
        float leaf_cost = kTriCost * (float)n;
        float split_cost = parent_area > 0.0f ? kTravCost + kTriCost * best_cost / parent_area : INFINITY;
        if (n <= kMaxLeaf && !(split_cost < leaf_cost)) {
            nodes[ni].first = b;
            nodes[ni].count = n;
            return;
        }
        uint32_t m;
        if (best_axis >= 0) {
            // Both sides are non-empty: the winning split has triangles on each.
            int ax = best_axis;
            float scale = (float)kBins / (cb.hi[ax] - cb.lo[ax]);
            float lo = cb.lo[ax];
            // Stable partition through the scratch buffer: left ids in
            // place, right ids after them, both in their original order.
            uint32_t l = b, r = 0;
            for (uint32_t i = b; i < e; i++) {
                uint32_t t = ids[i];
                if (bin_of(cen[3 * (size_t)t + ax], lo, scale) < best_split) ids[l++] = t;
                else scratch[r++] = t;
            }
            std::copy(scratch.begin(), scratch.begin() + r, ids.begin() + l);
            m = l;
        } else {
            // All centroids coincide: leaf, or object median. The ids are in
            // ascending order, so the median of (centroid, id) is the middle.
            if (n <= kHardLeaf) {
                nodes[ni].first = b;
                nodes[ni].count = n;
                return;
            }
            m = b + n / 2;
        }
        uint32_t c = (uint32_t)nodes.size();
        nodes.emplace_back();
        nodes.emplace_back();
        nodes[ni].first = c;
        nodes[ni].count = 0;
        build(c, b, m, d + 1);
        build(c + 1, m, e, d + 1);
    }
};

inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

}  // namespace

uint32_t build_bvh(const Soup &soup, std::vector<uint32_t> &node_words, std::vector<uint32_t> &order) {
    size_t n = soup.count();
    node_words.clear();
    order.clear();
    if (n == 0) return 0;
    std::vector<Box> tb(n);
    std::vector<float> cen(3 * n);
    auto prep = [&](size_t b, size_t e) {
        for (size_t t = b; t < e; t++) {
            Box bx;
            for (int v = 0; v < 3; v++) {
                float q[3];
                for (int k = 0; k < 3; k++) q[k] = soup.pos[9 * t + 3 * v + k] + 0.0f;  // -0 -> +0
                bx.grow(q);
            }
            tb[t] = bx;
            for (int k = 0; k < 3; k++) cen[3 * t + k] = 0.5f * (bx.lo[k] + bx.hi[k]);
        }
    };
    unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n > 65536 && hw > 1) {
        std::vector<std::thread> pool;
        size_t chunk = (n + hw - 1) / hw;
        for (unsigned i = 0; i < hw; i++) {
            size_t b = i * chunk, e = std::min(n, b + chunk);
            if (b < e) pool.emplace_back(prep, b, e);
        }
        for (auto &t : pool) t.join();
    } else {
        prep(0, n);
    }
    order.resize(n);
    for (size_t i = 0; i < n; i++) order[i] = (uint32_t)i;
    Builder bld(tb, cen, order);
    bld.nodes.reserve(2 * n / kMaxLeaf + 8);
    // Root, then one unused node, so every child pair starts at an even
    // index: a pair is 64 bytes, and 64-byte aligned it never straddles a
    // 128-byte cache line.
    bld.nodes.emplace_back();
    bld.nodes.emplace_back();
    bld.build(0, 0, (uint32_t)n, 0);
    node_words.resize(bld.nodes.size() * 8);
    for (size_t i = 0; i < bld.nodes.size(); i++) {
        const Node &nd = bld.nodes[i];
        uint32_t *w = &node_words[8 * i];
        for (int k = 0; k < 3; k++) {
            w[k] = fbits(nd.box.lo[k]);
            w[4 + k] = fbits(nd.box.hi[k]);
        }
        w[3] = nd.first;
        w[7] = nd.count;
    }
    return bld.depth;
}



// ---- 4-wide BVH collapsed from the binary tree -------------------------
//
// Node = 8 float4 (128 bytes, one cache line): lo.x[4], lo.y[4], lo.z[4],
// hi.x[4], hi.y[4], hi.z[4], ref[4], count[4]. Child i is an inner wide
// node (count 0, ref = wide node index), a leaf (count 1..30, ref = first
// leaf slot) or empty (count 0xffffffff). Child boxes and leaves are the
// binary tree's, bit for bit, so the slab and triangle tests see the same
// numbers. Each wide node takes the binary node's two children and
// repeatedly opens the inner child with the largest surface area until it
// has 4 children.
namespace {

constexpr uint32_t kWideLeafMax = 30;
constexpr uint32_t kEmpty = 0xffffffffu;

struct WideChild {
    float lo[3], hi[3];
    uint32_t first, count;  // the binary node's record
};

struct WideBuilder {
    const std::vector<uint32_t> &nw;  // binary node words
    std::vector<float> out;           // 32 words per wide node
    uint32_t depth = 0;
    bool ok = true;

    explicit WideBuilder(const std::vector<uint32_t> &n) : nw(n) {}

    WideChild of_bin(uint32_t n) const {
        WideChild c;
        for (int k = 0; k < 3; k++) {
            std::memcpy(&c.lo[k], &nw[8 * (size_t)n + k], 4);
            std::memcpy(&c.hi[k], &nw[8 * (size_t)n + 4 + k], 4);
        }
        c.first = nw[8 * (size_t)n + 3];
        c.count = nw[8 * (size_t)n + 7];
        return c;
    }
    static float area(const WideChild &c) {
        float d[3];
        for (int k = 0; k < 3; k++) d[k] = std::max(c.hi[k] - c.lo[k], 0.0f);
        return d[0] * d[1] + d[1] * d[2] + d[2] * d[0];
    }
    std::vector<WideChild> open(const WideChild &c) const { return {of_bin(c.first), of_bin(c.first + 1)}; }

    // Emits the wide node over the given children; returns its index.
    uint32_t emit(std::vector<WideChild> kids, uint32_t d) {
        depth = std::max(depth, d);
        while (kids.size() < 4) {
            int best = -1;
            float ba = -1.0f;
            for (size_t i = 0; i < kids.size(); i++) {
                if (kids[i].count != 0) continue;
                float a = area(kids[i]);
                if (a > ba) {
                    ba = a;
                    best = (int)i;
                }
            }
            if (best < 0) break;
            std::vector<WideChild> sub = open(kids[best]);
            kids.erase(kids.begin() + best);
            kids.insert(kids.begin() + best, sub.begin(), sub.end());
        }
        const uint32_t me = (uint32_t)(out.size() / 32);
        out.resize(out.size() + 32, 0.0f);
        uint32_t refs[4], cnts[4];
        float box[6][4];
        for (int i = 0; i < 4; i++) {
            const bool real = i < (int)kids.size();
            for (int k = 0; k < 3; k++) {
                box[k][i] = real ? kids[i].lo[k] : INFINITY;
                box[3 + k][i] = real ? kids[i].hi[k] : -INFINITY;
            }
            refs[i] = 0;
            cnts[i] = kEmpty;
        }
        for (int i = 0; i < (int)kids.size(); i++) {
            if (kids[i].count == 0) {
                refs[i] = emit(open(kids[i]), d + 1);
                cnts[i] = 0;
            } else {
                refs[i] = kids[i].first;
                cnts[i] = kids[i].count;
                if (cnts[i] > kWideLeafMax) ok = false;
            }
            if (refs[i] >= (1u << 27)) ok = false;
        }
        float *o = &out[32 * (size_t)me];
        for (int r = 0; r < 6; r++)
            for (int i = 0; i < 4; i++) o[4 * r + i] = box[r][i];
        std::memcpy(o + 24, refs, 16);
        std::memcpy(o + 28, cnts, 16);
        return me;
    }
};

}  // namespace

uint32_t build_wide(const std::vector<uint32_t> &node_words, std::vector<float> &wide) {
    wide.clear();
    if (node_words.size() < 8) return 0;
    WideBuilder wb(node_words);
    const WideChild root = wb.of_bin(0);
    std::vector<WideChild> kids;
    if (root.count == 0) kids = wb.open(root);
    else kids.push_back(root);
    wb.emit(kids, 1);
    // At most three entries per level go on the stack.
    if (!wb.ok || 3u * wb.depth > kWideStack) return 0;
    wide.swap(wb.out);
    return wb.depth;
}

}  // namespace oxpt
