// bvh_build.cpp - host builder of the compressed 8-wide BVH that replaces the
// driver-built BLAS of the reference (examples/cornell_box/src/main.rs:464-563;
// the reference's ray queries traverse it at general.wesl:1052-1078).
//
// 0. Pre-splitting: the triangle references whose boxes waste the most area
//    are split by clipping, up to half as many extra references as
//    triangles.
// 1. Binned-SAH binary BVH (16 bins per axis) down to single references,
//    subtrees above 64k primitives built on separate threads.
// 2. Collapse to 8-wide by dynamic programming over the binary tree
//    (the cost-optimal collapse of Ylitie, Karras and Laine 2017, "Efficient
//    Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs"): every
//    binary subtree is covered by at most i child slots, i = 1..8, at least
//    surface-area cost, where a visited 8-wide node costs kNodeCost and a
//    triangle of a hit leaf slot kTriCost.
// 3. Slot assignment by direction octant (greedy maximum alignment of the
//    child centroid offset with the slot's octant direction).
// 4. 8-bit quantization of the child boxes against a per-node origin that is
//    a multiple of the power-of-two scale, so that the dequantized planes
//    fmaf(q, scale, p) are exact and enclose the true boxes.
#include "bvh_build.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <string>
#include <vector>

namespace lmb {

namespace {

struct PrimRef {
    float lo[3];
    float hi[3];
    float c[3];
    uint32_t id;
};

struct BNode {
    float lo[3];
    float hi[3];
    uint32_t left;   // index of the left child; right = left + 1
    uint32_t first;  // leaf: first PrimRef
    uint32_t count;  // leaf: number of primitives (0 for internal nodes)
};

constexpr int kBins = 16;
// Triangles of one leaf slot (the node format's limit).
constexpr uint32_t kMaxLeaf = 3;
// The binary tree is built down to leaves of this many triangles; the
// collapse chooses the leaf slots (subtrees of at most kMaxLeaf triangles).
#ifndef LMB_BVH_BINARY_LEAF
#define LMB_BVH_BINARY_LEAF 1
#endif
constexpr uint32_t kBinaryLeaf = LMB_BVH_BINARY_LEAF;
static_assert(kBinaryLeaf >= 1 && kBinaryLeaf <= kMaxLeaf, "binary leaves must fit a leaf slot");

// Relative costs of the collapse: one visit of an 8-wide node (its eight
// child-box tests and the choice of the next node) and one triangle test of a
// hit leaf slot.  The traversal tests triangles in full rounds shared by the
// 64 lanes of a wave, so one more triangle costs far less than a node visit:
// on the 1M-triangle soup at 1080p the traversal takes 0.418, 0.415, 0.413,
// 0.413, 0.415, 0.418 and 0.423 ms per launch at 0.2, 0.25, 0.3, 0.35, 0.4,
// 0.5 and 0.75.
#ifndef LMB_BVH_TRI_COST
#define LMB_BVH_TRI_COST 0.3
#endif
constexpr float kNodeCost = 1.0f;
constexpr float kTriCost = (float)(LMB_BVH_TRI_COST);
constexpr uint32_t kParallelThreshold = 1u << 16;

// Pre-splitting budget: up to this many extra references per 100 triangles
// (0: none).  On the 1M-triangle soup at 1080p, k_extend takes 2.05, 1.99,
// 1.976 and 1.972 ms per frame at 0, 25, 50 and 100 (108k, 139k, 169k and
// 226k nodes; 101, 115, 130 and 158 MB of BVH).
#ifndef LMB_BVH_PRESPLIT
#define LMB_BVH_PRESPLIT 50
#endif
constexpr uint32_t kPresplitPercent = LMB_BVH_PRESPLIT;
#ifndef LMB_BVH_PRESPLIT_RATIO
#define LMB_BVH_PRESPLIT_RATIO 3.0
#endif
constexpr double kPresplitRatio = LMB_BVH_PRESPLIT_RATIO;
// Scenes below this many triangles are not pre-split: there the duplicates
// cost more triangle tests than the tighter boxes save (the 36-triangle
// Cornell box at 256 x 256, 16 spp: 1.14 -> 1.20 ms per frame when split;
// the 64-scene batch of 1k-500k triangles: 59.3 -> 60.6 ms when split, and
// 60.4 ms when only its scenes of 2^17 triangles or more are).
#ifndef LMB_BVH_PRESPLIT_MIN
#define LMB_BVH_PRESPLIT_MIN 524288
#endif
constexpr size_t kPresplitMinTriangles = LMB_BVH_PRESPLIT_MIN;

inline float half_area(const float lo[3], const float hi[3]) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0.f || dy < 0.f || dz < 0.f) return 0.f;
    return dx * dy + dy * dz + dz * dx;
}

struct Box {
    float lo[3], hi[3];
    void reset() {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::numeric_limits<float>::infinity();
            hi[a] = -std::numeric_limits<float>::infinity();
        }
    }
    void grow(const float l[3], const float h[3]) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], l[a]);
            hi[a] = std::max(hi[a], h[a]);
        }
    }
    void grow_point(const float p[3]) { grow(p, p); }
};

class BinaryBuilder {
  public:
    BinaryBuilder(std::vector<PrimRef>& refs) : refs_(refs) {
        nodes_.resize(std::max<size_t>(1, 2 * refs.size()));
        next_.store(1);
    }

    void build() { build_node(0, 0, (uint32_t)refs_.size()); }
    uint32_t node_count() const { return next_.load(); }
    std::vector<BNode>& nodes() { return nodes_; }

  private:
    void build_node(uint32_t ni, uint32_t first, uint32_t count) {
        Box b, cb;
        b.reset();
        cb.reset();
        for (uint32_t i = first; i < first + count; ++i) {
            b.grow(refs_[i].lo, refs_[i].hi);
            cb.grow_point(refs_[i].c);
        }
        BNode& n = nodes_[ni];
        std::memcpy(n.lo, b.lo, sizeof n.lo);
        std::memcpy(n.hi, b.hi, sizeof n.hi);
        if (count <= kBinaryLeaf) {
            n.first = first;
            n.count = count;
            n.left = 0;
            return;
        }
        // Binned SAH over the three axes.
        float best_cost = std::numeric_limits<float>::infinity();
        int best_axis = -1, best_split = -1;
        for (int a = 0; a < 3; ++a) {
            float ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0.f)) continue;
            float scale = kBins / ext;
            uint32_t cnt[kBins] = {};
            Box bb[kBins];
            for (auto& x : bb) x.reset();
            for (uint32_t i = first; i < first + count; ++i) {
                int k = (int)((refs_[i].c[a] - cb.lo[a]) * scale);
                k = std::min(std::max(k, 0), kBins - 1);
                cnt[k]++;
                bb[k].grow(refs_[i].lo, refs_[i].hi);
            }
            float right_area[kBins];
            uint32_t right_cnt[kBins];
            Box acc;
            acc.reset();
            uint32_t c = 0;
            for (int k = kBins - 1; k > 0; --k) {
                acc.grow(bb[k].lo, bb[k].hi);
                c += cnt[k];
                right_area[k] = half_area(acc.lo, acc.hi);
                right_cnt[k] = c;
            }
            acc.reset();
            c = 0;
            for (int k = 0; k < kBins - 1; ++k) {
                acc.grow(bb[k].lo, bb[k].hi);
                c += cnt[k];
                if (c == 0 || right_cnt[k + 1] == 0) continue;
                float cost = half_area(acc.lo, acc.hi) * (float)c +
                             right_area[k + 1] * (float)right_cnt[k + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = a;
                    best_split = k + 1;
                }
            }
        }
        uint32_t mid = first + count / 2;
        if (best_axis >= 0) {
            float ext = cb.hi[best_axis] - cb.lo[best_axis];
            float scale = kBins / ext;
            float lo = cb.lo[best_axis];
            auto it = std::partition(
                refs_.begin() + first, refs_.begin() + first + count,
                [&](const PrimRef& r) {
                    int k = (int)((r.c[best_axis] - lo) * scale);
                    k = std::min(std::max(k, 0), kBins - 1);
                    return k < best_split;
                });
            mid = (uint32_t)(it - refs_.begin());
            if (mid == first || mid == first + count) mid = first + count / 2;
        } else {
            // All centroids coincide: split by position in the list.
            mid = first + count / 2;
        }
        uint32_t left = next_.fetch_add(2);
        n.left = left;
        n.first = 0;
        n.count = 0;
        uint32_t lc = mid - first, rc = first + count - mid;
        if (count >= kParallelThreshold) {
            auto fut = std::async(std::launch::async,
                                  [=] { build_node(left, first, lc); });
            build_node(left + 1, mid, rc);
            fut.get();
        } else {
            build_node(left, first, lc);
            build_node(left + 1, mid, rc);
        }
    }

    std::vector<PrimRef>& refs_;
    std::vector<BNode> nodes_;
    std::atomic<uint32_t> next_;
};

// Cost-optimal covering of every binary subtree by at most i slots of an
// 8-wide node.  A slot is either a leaf (a subtree of at most kMaxLeaf
// triangles, tested when the slot's box is hit) or an 8-wide node of its
// own.  Children of a binary node have larger indices than the node, and a
// subtree's primitives are contiguous, left child first.
class CollapsePlan {
  public:
    CollapsePlan(const std::vector<BNode>& bn, uint32_t count)
        : bn_(bn), cost_(count), split_(count), wide_left_(count), leaf_(count), prims_(count),
          first_(count) {
        for (int64_t n = (int64_t)count - 1; n >= 0; --n) {
            const BNode& b = bn[n];
            const float area = half_area(b.lo, b.hi);
            if (b.count > 0) {
                prims_[n] = b.count;
                first_[n] = b.first;
                leaf_[n] = 1;
                cost_[n].fill(area * (float)b.count * kTriCost);
                split_[n].fill(0);
                wide_left_[n] = 0;
                continue;
            }
            const uint32_t l = b.left, r = b.left + 1;
            prims_[n] = prims_[l] + prims_[r];
            first_[n] = first_[l];
            float best = std::numeric_limits<float>::infinity();
            uint8_t best_k = 1;
            for (int k = 1; k < 8; ++k) {
                const float c = cost_[l][k - 1] + cost_[r][8 - k - 1];
                if (c < best) {
                    best = c;
                    best_k = (uint8_t)k;
                }
            }
            wide_left_[n] = best_k;
            float single = area * kNodeCost + best;
            leaf_[n] = 0;
            if (prims_[n] <= kMaxLeaf) {
                const float as_leaf = area * (float)prims_[n] * kTriCost;
                if (as_leaf <= single) {
                    single = as_leaf;
                    leaf_[n] = 1;
                }
            }
            cost_[n][0] = single;
            split_[n][0] = 0;
            for (int i = 2; i <= 8; ++i) {
                float c = single;
                uint8_t sk = 0;
                for (int k = 1; k < i; ++k) {
                    const float d = cost_[l][k - 1] + cost_[r][i - k - 1];
                    if (d < c) {
                        c = d;
                        sk = (uint8_t)k;
                    }
                }
                cost_[n][i - 1] = c;
                split_[n][i - 1] = sk;
            }
        }
    }

    // The child slots (binary node indices) of the 8-wide node made from the
    // internal binary node n.
    int children(uint32_t n, uint32_t out[8]) const {
        int m = 0;
        const uint32_t l = bn_[n].left;
        expand(l, wide_left_[n], out, &m);
        expand(l + 1, 8 - wide_left_[n], out, &m);
        return m;
    }
    // Slot n is a leaf of prims(n) primitives starting at first(n).
    bool leaf(uint32_t n) const { return leaf_[n] != 0; }
    uint32_t prims(uint32_t n) const { return prims_[n]; }
    uint32_t first(uint32_t n) const { return first_[n]; }

  private:
    void expand(uint32_t n, int i, uint32_t out[8], int* m) const {
        const uint8_t k = split_[n][i - 1];
        if (k == 0) {
            out[(*m)++] = n;
            return;
        }
        expand(bn_[n].left, k, out, m);
        expand(bn_[n].left + 1, i - k, out, m);
    }

    const std::vector<BNode>& bn_;
    std::vector<std::array<float, 8>> cost_;    // [n][i - 1]: least cost with at most i slots
    std::vector<std::array<uint8_t, 8>> split_;  // [n][i - 1]: 0 = one slot, else slots of the left child
    std::vector<uint8_t> wide_left_;             // slots of the left child when n is an 8-wide node
    std::vector<uint8_t> leaf_;                  // n as one slot is a leaf
    std::vector<uint32_t> prims_;                // primitives of the subtree
    std::vector<uint32_t> first_;                // its first primitive
};

inline bool finite3(const float* v) {
    return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
}

// Quantization scale exponent for one axis of a node covering [lo, hi]:
// 2^e >= extent / 253 so that q_hi <= 255 with a plane to spare, and
// 2^e >= max|coord| / 2^22 so that every dequantized plane (k + q) * 2^e is
// an exact float.
// Largest float <= x and smallest float >= x.
inline float float_down(double x) {
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -std::numeric_limits<float>::infinity());
    return f;
}
inline float float_up(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, std::numeric_limits<float>::infinity());
    return f;
}

// Bounds of the part of triangle (a, b, c) inside the box [lo, hi] and on
// one side of the plane x_axis = plane (side 0: <=, 1: >=), by clipping the
// triangle polygon in double precision (Sutherland-Hodgman); rounded outward
// to floats.  False when the part is empty.
bool clipped_bounds(const double tri[3][3], const float lo[3], const float hi[3], int axis,
                    double plane, int side, float out_lo[3], float out_hi[3]) {
    double poly[16][3], tmp[16][3];
    int n = 3;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) poly[i][k] = tri[i][k];
    // planes: the box's six, then the split plane
    for (int pl = 0; pl < 7 && n > 0; ++pl) {
        int ax;
        double v;
        bool keep_below;
        if (pl < 6) {
            ax = pl >> 1;
            keep_below = (pl & 1) != 0;  // even: keep >= lo, odd: keep <= hi
            v = keep_below ? (double)hi[ax] : (double)lo[ax];
        } else {
            ax = axis;
            keep_below = side == 0;
            v = plane;
        }
        int m = 0;
        for (int i = 0; i < n; ++i) {
            const double* p = poly[i];
            const double* q = poly[(i + 1) % n];
            const double dp = keep_below ? v - p[ax] : p[ax] - v;
            const double dq = keep_below ? v - q[ax] : q[ax] - v;
            if (dp >= 0.0 && m < 16) {
                for (int k = 0; k < 3; ++k) tmp[m][k] = p[k];
                ++m;
            }
            if ((dp >= 0.0) != (dq >= 0.0) && m < 16) {
                const double s = dp / (dp - dq);
                for (int k = 0; k < 3; ++k) tmp[m][k] = p[k] + s * (q[k] - p[k]);
                tmp[m][ax] = v;
                ++m;
            }
        }
        n = m;
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) poly[i][k] = tmp[i][k];
    }
    if (n == 0) return false;
    for (int k = 0; k < 3; ++k) {
        double l = poly[0][k], h = poly[0][k];
        for (int i = 1; i < n; ++i) {
            l = std::min(l, poly[i][k]);
            h = std::max(h, poly[i][k]);
        }
        // within the reference's box, whatever the clipping's rounding
        out_lo[k] = std::max(lo[k], float_down(l));
        out_hi[k] = std::min(hi[k], float_up(h));
        if (!(out_lo[k] <= out_hi[k])) return false;
    }
    return true;
}

// Pre-splitting of references (early split clipping: Ernst and Greiner 2007,
// Karras and Aila 2013): the references whose boxes waste the most area
// beyond their triangle (box half-area minus the triangle's area) are split
// in half along their longest axis, each half bounding the triangle's part
// inside it, until `extra` references have been added.  Every point of a
// triangle stays inside one of its references' boxes, so a leaf of any of
// them finds every hit of the triangle; the leaves store the triangle once
// per reference.
void presplit(std::vector<PrimRef>& refs, const float* vertices, const uint32_t* indices,
              size_t extra) {
    if (extra == 0) return;
    auto tri_of = [&](uint32_t t, double tri[3][3]) {
        for (int v = 0; v < 3; ++v) {
            const uint64_t i = indices ? indices[3 * (uint64_t)t + v] : 3 * (uint64_t)t + v;
            for (int k = 0; k < 3; ++k) tri[v][k] = vertices[4 * i + k];
        }
    };
    auto tri_area = [&](const double tri[3][3]) {
        double e1[3], e2[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = tri[1][k] - tri[0][k];
            e2[k] = tri[2][k] - tri[0][k];
        }
        const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                     cz = e1[0] * e2[1] - e1[1] * e2[0];
        return 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
    };
    // (waste, reference index); a split part's triangle area is taken as the
    // whole triangle's times its share of the parent box's half-area
    struct Item {
        double waste;
        uint32_t ref;
        double area;
        bool operator<(const Item& o) const { return waste < o.waste; }
    };
    std::vector<Item> heap;
    heap.reserve(refs.size() + extra);
    for (uint32_t i = 0; i < refs.size(); ++i) {
        double tri[3][3];
        tri_of(refs[i].id, tri);
        const double a = tri_area(tri);
        heap.push_back({(double)half_area(refs[i].lo, refs[i].hi) - a, i, a});
    }
    std::make_heap(heap.begin(), heap.end());
    size_t added = 0;
    while (added < extra && !heap.empty()) {
        std::pop_heap(heap.begin(), heap.end());
        const Item it = heap.back();
        heap.pop_back();
        if (!(it.waste > 0.0)) break;
        // Only boxes at least kPresplitRatio times their part's area: an
        // axis-aligned right triangle (a wall's half) has a box of twice its
        // area, and splitting it only adds triangle tests.
        if (!(it.waste + it.area > kPresplitRatio * it.area)) continue;
        PrimRef r = refs[it.ref];
        int axis = 0;
        for (int k = 1; k < 3; ++k)
            if (r.hi[k] - r.lo[k] > r.hi[axis] - r.lo[axis]) axis = k;
        const double plane = 0.5 * ((double)r.lo[axis] + (double)r.hi[axis]);
        double tri[3][3];
        tri_of(r.id, tri);
        PrimRef a = r, b = r;
        const bool ha = clipped_bounds(tri, r.lo, r.hi, axis, plane, 0, a.lo, a.hi);
        const bool hb = clipped_bounds(tri, r.lo, r.hi, axis, plane, 1, b.lo, b.hi);
        if (!ha || !hb) continue;  // one side is empty: nothing to gain, drop the item
        const double area_r = std::max(1e-30, (double)half_area(r.lo, r.hi));
        for (PrimRef* p : {&a, &b})
            for (int k = 0; k < 3; ++k) p->c[k] = 0.5f * p->lo[k] + 0.5f * p->hi[k];
        refs[it.ref] = a;
        refs.push_back(b);
        ++added;
        const double ta = it.area * half_area(a.lo, a.hi) / area_r;
        const double tb = it.area * half_area(b.lo, b.hi) / area_r;
        heap.push_back({(double)half_area(a.lo, a.hi) - ta, it.ref, ta});
        std::push_heap(heap.begin(), heap.end());
        heap.push_back({(double)half_area(b.lo, b.hi) - tb, (uint32_t)(refs.size() - 1), tb});
        std::push_heap(heap.begin(), heap.end());
    }
}

int pick_exponent(float lo, float hi) {
    double extent = (double)hi - (double)lo;
    double need = std::max(extent / 253.0,
                           std::max(std::fabs((double)lo), std::fabs((double)hi)) /
                               4194304.0);
    need = std::max(need, 1e-30);
    int e = (int)std::ceil(std::log2(need));
    while (std::ldexp(1.0, e) < need) ++e;
    e = std::max(e, -126);
    e = std::min(e, 127);
    return e;
}

}  // namespace

void assign_octant_slots(const float lo[3], const float hi[3], const float (*child_lo)[3],
                         const float (*child_hi)[3], int n, int slot_of[8]) {
    float center[3];
    for (int a = 0; a < 3; ++a) center[a] = 0.5f * (lo[a] + hi[a]);
    struct Cand {
        float score;
        int child, slot;
    };
    Cand cands[64];
    int nc = 0;
    for (int i = 0; i < n; ++i) {
        float d[3];
        for (int a = 0; a < 3; ++a) d[a] = 0.5f * (child_lo[i][a] + child_hi[i][a]) - center[a];
        for (int s = 0; s < 8; ++s) {
            float sc = ((s & 1) ? d[0] : -d[0]) + ((s & 2) ? d[1] : -d[1]) +
                       ((s & 4) ? d[2] : -d[2]);
            cands[nc++] = {sc, i, s};
        }
    }
    std::stable_sort(cands, cands + nc,
                     [](const Cand& x, const Cand& y) { return x.score > y.score; });
    int child_in_slot[8];
    for (int i = 0; i < 8; ++i) {
        slot_of[i] = -1;
        child_in_slot[i] = -1;
    }
    for (int k = 0; k < nc; ++k) {
        const Cand& c = cands[k];
        if (slot_of[c.child] >= 0 || child_in_slot[c.slot] >= 0) continue;
        slot_of[c.child] = c.slot;
        child_in_slot[c.slot] = c.child;
    }
}

int build_bvh8(const float* vertices, uint64_t num_vertices, const uint32_t* indices,
               uint64_t num_triangles, BvhBuildResult* out, std::string* err) {
    auto t0 = std::chrono::steady_clock::now();
    out->nodes.clear();
    out->tris.clear();
    out->max_depth = 0;
    out->sah_cost = 0.0;

    // Primitive references (non-finite triangles can never be hit and are
    // left out).
    std::vector<PrimRef> refs;
    refs.reserve(num_triangles);
    for (uint64_t t = 0; t < num_triangles; ++t) {
        uint64_t i0 = indices ? indices[3 * t + 0] : 3 * t + 0;
        uint64_t i1 = indices ? indices[3 * t + 1] : 3 * t + 1;
        uint64_t i2 = indices ? indices[3 * t + 2] : 3 * t + 2;
        if (i0 >= num_vertices || i1 >= num_vertices || i2 >= num_vertices) {
            *err = "triangle " + std::to_string(t) + " indexes a vertex out of range";
            return -1;
        }
        const float* a = vertices + 4 * i0;
        const float* b = vertices + 4 * i1;
        const float* c = vertices + 4 * i2;
        if (!finite3(a) || !finite3(b) || !finite3(c)) continue;
        PrimRef r;
        for (int k = 0; k < 3; ++k) {
            r.lo[k] = std::min(a[k], std::min(b[k], c[k]));
            r.hi[k] = std::max(a[k], std::max(b[k], c[k]));
            r.c[k] = 0.5f * r.lo[k] + 0.5f * r.hi[k];
        }
        r.id = (uint32_t)t;
        refs.push_back(r);
    }
    if (refs.empty()) {
        out->build_ms = std::chrono::duration<double, std::milli>(
                            std::chrono::steady_clock::now() - t0)
                            .count();
        return 0;
    }
    if (refs.size() >= kPresplitMinTriangles)
        presplit(refs, vertices, indices, refs.size() * kPresplitPercent / 100);

    BinaryBuilder bb(refs);
    bb.build();
    std::vector<BNode>& bn = bb.nodes();
    const CollapsePlan plan(bn, bb.node_count());

    // Collapse into 8-wide nodes, breadth first so that the internal children
    // of every node are contiguous.
    struct Pending {
        uint32_t bnode;
        uint32_t node8;
        uint32_t depth;
    };
    std::vector<Pending> queue;
    queue.push_back({0, 0, 1});
    out->nodes.resize(1);
    double root_area = std::max(1e-30, (double)half_area(bn[0].lo, bn[0].hi));
    double sah = 0.0;
    size_t head = 0;
    while (head < queue.size()) {
        Pending pd = queue[head++];
        out->max_depth = std::max(out->max_depth, pd.depth);
        const BNode& b = bn[pd.bnode];
        sah += 1.0 * half_area(b.lo, b.hi) / root_area;

        uint32_t ch[8];
        int nch = 0;
        if (b.count > 0) {
            ch[nch++] = pd.bnode;  // only a leaf root gets here
        } else {
            nch = plan.children(pd.bnode, ch);
        }

        // Octant slot assignment.
        float clo[8][3], chi[8][3];
        for (int i = 0; i < nch; ++i)
            for (int a = 0; a < 3; ++a) {
                clo[i][a] = bn[ch[i]].lo[a];
                chi[i][a] = bn[ch[i]].hi[a];
            }
        int slot_of[8];
        int child_in_slot[8];
        assign_octant_slots(b.lo, b.hi, clo, chi, nch, slot_of);
        for (int i = 0; i < 8; ++i) child_in_slot[i] = -1;
        for (int i = 0; i < nch; ++i) child_in_slot[slot_of[i]] = i;

        Bvh8Node node;
        std::memset(&node, 0, sizeof node);
        // Quantization frame.
        float scale[3];
        for (int a = 0; a < 3; ++a) {
            int e = pick_exponent(b.lo[a], b.hi[a]);
            double s = std::ldexp(1.0, e);
            double p = std::floor((double)b.lo[a] / s) * s;
            node.p[a] = (float)p;
            node.e[a] = (uint8_t)(e + 127);
            scale[a] = (float)s;
        }
        // Children in slot order: internal nodes get consecutive indices,
        // leaf triangles consecutive positions.
        uint32_t child_base = (uint32_t)out->nodes.size();
        uint32_t tri_base = (uint32_t)out->tris.size();
        node.child_base = child_base;
        node.tri_base = tri_base;
        uint32_t n_internal = 0, tri_off = 0;
        for (int s = 0; s < 8; ++s) {
            int ci = child_in_slot[s];
            if (ci < 0) {
                node.meta[s] = kMetaEmpty;
                for (int a = 0; a < 3; ++a) {
                    node.qlo[a][s] = 0;
                    node.qhi[a][s] = 0;
                }
                continue;
            }
            const BNode& c = bn[ch[ci]];
            for (int a = 0; a < 3; ++a) {
                double s_a = scale[a];
                double ql = std::floor(((double)c.lo[a] - (double)node.p[a]) / s_a);
                double qh = std::ceil(((double)c.hi[a] - (double)node.p[a]) / s_a);
                ql = std::min(std::max(ql, 0.0), 255.0);
                qh = std::min(std::max(qh, 0.0), 255.0);
                // (c - p) is not exact in double when |c| << |p|: step the
                // planes outward until the kernel's expression encloses c.
                while (ql > 0.0 && !(std::fmaf((float)ql, scale[a], node.p[a]) <= c.lo[a]))
                    ql -= 1.0;
                while (qh < 255.0 && !(std::fmaf((float)qh, scale[a], node.p[a]) >= c.hi[a]))
                    qh += 1.0;
                float dl = std::fmaf((float)ql, scale[a], node.p[a]);
                float dh = std::fmaf((float)qh, scale[a], node.p[a]);
                if (!(dl <= c.lo[a]) || !(dh >= c.hi[a])) {
                    *err = "BVH quantization is not conservative (internal error): axis " +
                           std::to_string(a) + " child [" + std::to_string(c.lo[a]) + ", " +
                           std::to_string(c.hi[a]) + "] dequantized [" + std::to_string(dl) +
                           ", " + std::to_string(dh) + "] p " + std::to_string(node.p[a]) +
                           " scale " + std::to_string(scale[a]) + " q " + std::to_string(ql) +
                           " " + std::to_string(qh);
                    return -1;
                }
                node.qlo[a][s] = (uint8_t)ql;
                node.qhi[a][s] = (uint8_t)qh;
            }
            if (!plan.leaf(ch[ci])) {
                node.meta[s] = kMetaInternal;
                node.imask |= (uint8_t)(1u << s);
                n_internal++;
            } else {
                const uint32_t count = plan.prims(ch[ci]), first = plan.first(ch[ci]);
                uint32_t unary = (1u << count) - 1u;
                node.meta[s] = (uint8_t)((unary << 5) | tri_off);
                for (uint32_t k = 0; k < count; ++k) {
                    const PrimRef& r = refs[first + k];
                    uint64_t t = r.id;
                    uint64_t i0 = indices ? indices[3 * t + 0] : 3 * t + 0;
                    uint64_t i1 = indices ? indices[3 * t + 1] : 3 * t + 1;
                    uint64_t i2 = indices ? indices[3 * t + 2] : 3 * t + 2;
                    LeafTri lt;
                    for (int q = 0; q < 3; ++q) {
                        lt.v0[q] = vertices[4 * i0 + q];
                        lt.v1[q] = vertices[4 * i1 + q];
                        lt.v2[q] = vertices[4 * i2 + q];
                    }
                    uint32_t id = r.id;
                    std::memcpy(&lt.v0[3], &id, 4);
                    lt.v1[3] = 0.f;
                    lt.v2[3] = 0.f;
                    out->tris.push_back(lt);
                }
                tri_off += count;
            }
        }
        if (tri_off > 24) {
            *err = "BVH node references more than 24 triangles (internal error)";
            return -1;
        }
        out->nodes.resize(out->nodes.size() + n_internal);
        uint32_t k = 0;
        for (int s = 0; s < 8; ++s) {
            int ci = child_in_slot[s];
            if (ci < 0 || node.meta[s] != kMetaInternal) continue;
            queue.push_back({ch[ci], child_base + k, pd.depth + 1});
            ++k;
        }
        out->nodes[pd.node8] = node;
    }
    out->sah_cost = sah;
    out->build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
            .count();
    return 0;
}

}  // namespace lmb
