// api_instances.cpp - C ABI: instanced scenes (lmb_scene_create_instanced)
// and moving their instances (lmb_scene_update_instances).
//
// The reference builds one BLAS per geometry and a TLAS over instances
// (examples/cornell_box/src/main.rs:464-563), and a moved instance means a
// new TlasInstance transform and a TLAS rebuild.  Here the scene is one
// 8-wide BVH in the traversal's format: a top tree over the instances, built
// on the host, above one instance-pure subtree per instance whose triangles
// are stored in world space.  Moving instances rebuilds the top tree on the
// host and, on the device, transforms each triangle from its object-space
// copy and refits every node from the leaves up (lmb_instances.h); the
// traversal kernels are those of flat scenes, unchanged.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lambent.h"
#include "bvh_build.h"
#include "lmb_instances.h"
#include "lmb_internal.h"

namespace lmb {

namespace {

struct WBox {
    float lo[3], hi[3];
};

double half_area(const WBox& b) {
    const double dx = (double)b.hi[0] - b.lo[0], dy = (double)b.hi[1] - b.lo[1],
                 dz = (double)b.hi[2] - b.lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0;
    return dx * dy + dy * dz + dz * dx;
}

WBox merge(const WBox& a, const WBox& b) {
    WBox r;
    for (int k = 0; k < 3; ++k) {
        r.lo[k] = std::min(a.lo[k], b.lo[k]);
        r.hi[k] = std::max(a.hi[k], b.hi[k]);
    }
    return r;
}

double det3(const float* m) {
    return (double)m[0] * ((double)m[5] * m[10] - (double)m[6] * m[9]) -
           (double)m[1] * ((double)m[4] * m[10] - (double)m[6] * m[8]) +
           (double)m[2] * ((double)m[4] * m[9] - (double)m[5] * m[8]);
}

// World box of an object-space box under a 3x4 transform (its 8 corners in
// double, widened by a relative 1e-6 so it also holds the float-rounded
// vertices), or false when it is not finite or too large to refit.
bool world_box(const InstancePart& p, const float* m, WBox* out) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int c = 0; c < 8; ++c) {
        const double x = (c & 1) ? p.obj_hi[0] : p.obj_lo[0];
        const double y = (c & 2) ? p.obj_hi[1] : p.obj_lo[1];
        const double z = (c & 4) ? p.obj_hi[2] : p.obj_lo[2];
        for (int r = 0; r < 3; ++r) {
            const double w = m[4 * r] * x + m[4 * r + 1] * y + m[4 * r + 2] * z + m[4 * r + 3];
            lo[r] = std::min(lo[r], w);
            hi[r] = std::max(hi[r], w);
        }
    }
    for (int r = 0; r < 3; ++r) {
        const double pad = 1e-6 * (std::fabs(lo[r]) + std::fabs(hi[r]) + 1e-30);
        lo[r] -= pad;
        hi[r] += pad;
        if (!(std::fabs(lo[r]) < 1e30 && std::fabs(hi[r]) < 1e30)) return false;
        out->lo[r] = (float)lo[r];
        out->hi[r] = (float)hi[r];
    }
    return true;
}

// The top tree over the visible parts: binary SAH over their world boxes
// (one part per leaf), collapsed to 8-wide nodes and laid out breadth first
// from node position 0, each node's children at consecutive positions.
struct TopTree {
    std::vector<RefitEntry> entries;  // sorted by height
    std::vector<uint32_t> level_off;  // entries of height h: [off[h], off[h + 1])
    std::vector<uint32_t> root_pos;   // per part
    uint32_t max_depth = 0;
    uint32_t num_nodes = 0;
    double sah = 0.0;  // node areas relative to the root box
};

struct BinNode {
    WBox box;
    int left = -1, right = -1;  // binary children
    int part = -1;              // leaf: the part
};

int build_binary(std::vector<int>& idx, int b, int e, const std::vector<WBox>& boxes,
                 std::vector<BinNode>& out) {
    const int ni = (int)out.size();
    out.emplace_back();
    WBox box = boxes[idx[b]];
    for (int i = b + 1; i < e; ++i) box = merge(box, boxes[idx[i]]);
    out[ni].box = box;
    if (e - b == 1) {
        out[ni].part = idx[b];
        return ni;
    }
    auto centre = [&](int k, int a) {
        return 0.5 * ((double)boxes[k].lo[a] + (double)boxes[k].hi[a]);
    };
    double clo[3], chi[3];
    for (int a = 0; a < 3; ++a) {
        clo[a] = INFINITY;
        chi[a] = -INFINITY;
        for (int i = b; i < e; ++i) {
            clo[a] = std::min(clo[a], centre(idx[i], a));
            chi[a] = std::max(chi[a], centre(idx[i], a));
        }
    }
    constexpr int kBins = 16;
    double best = INFINITY;
    int best_axis = -1, best_split = -1;
    for (int a = 0; a < 3; ++a) {
        const double ext = chi[a] - clo[a];
        if (!(ext > 0.0)) continue;
        int cnt[kBins] = {};
        WBox bb[kBins];
        bool used[kBins] = {};
        for (int i = b; i < e; ++i) {
            int k = (int)((centre(idx[i], a) - clo[a]) / ext * kBins);
            k = std::min(std::max(k, 0), kBins - 1);
            bb[k] = used[k] ? merge(bb[k], boxes[idx[i]]) : boxes[idx[i]];
            used[k] = true;
            cnt[k]++;
        }
        for (int s = 1; s < kBins; ++s) {
            WBox l{}, r{};
            bool lu = false, ru = false;
            int lc = 0, rc = 0;
            for (int k = 0; k < s; ++k)
                if (used[k]) {
                    l = lu ? merge(l, bb[k]) : bb[k];
                    lu = true;
                    lc += cnt[k];
                }
            for (int k = s; k < kBins; ++k)
                if (used[k]) {
                    r = ru ? merge(r, bb[k]) : bb[k];
                    ru = true;
                    rc += cnt[k];
                }
            if (lc == 0 || rc == 0) continue;
            const double cost = half_area(l) * lc + half_area(r) * rc;
            if (cost < best) {
                best = cost;
                best_axis = a;
                best_split = s;
            }
        }
    }
    int mid = b + (e - b) / 2;
    if (best_axis >= 0) {
        const double ext = chi[best_axis] - clo[best_axis];
        auto it = std::stable_partition(idx.begin() + b, idx.begin() + e, [&](int k) {
            int bin = (int)((centre(k, best_axis) - clo[best_axis]) / ext * kBins);
            return std::min(std::max(bin, 0), kBins - 1) < best_split;
        });
        mid = (int)(it - idx.begin());
        if (mid == b || mid == e) mid = b + (e - b) / 2;
    }
    const int l = build_binary(idx, b, mid, boxes, out);
    const int r = build_binary(idx, mid, e, boxes, out);
    out[ni].left = l;
    out[ni].right = r;
    return ni;
}

// parts[i] with its world box; visible[i]: whether it takes part.
TopTree build_top(const InstanceState& st, const std::vector<WBox>& boxes,
                  const std::vector<uint8_t>& visible) {
    TopTree t;
    const uint32_t n = (uint32_t)st.parts.size();
    t.root_pos.assign(n, 0);
    std::vector<int> vis;
    for (uint32_t i = 0; i < n; ++i)
        if (visible[i]) vis.push_back((int)i);
    uint32_t next = 0;
    struct Node8 {
        RefitEntry e;
        uint32_t depth;
        std::vector<int> kids;  // child node8 index, or -(part + 1)
    };
    std::vector<Node8> top;
    if (vis.empty()) {
        // Nothing visible: a root with no children (the traversal finds no hit).
        Node8 z{};
        z.e.dst = 0;
        z.depth = 1;
        top.push_back(z);
        next = 1;
    } else if (vis.size() == 1) {
        t.root_pos[vis[0]] = 0;
        next = 1;
        t.max_depth = st.parts[vis[0]].levels;
    } else {
        std::vector<BinNode> bn;
        bn.reserve(2 * vis.size());
        build_binary(vis, 0, (int)vis.size(), boxes, bn);
        const double root_area = std::max(1e-30, half_area(bn[0].box));
        struct Pending {
            int bnode;
            uint32_t pos, depth;
        };
        std::vector<Pending> queue{{0, 0, 1}};
        next = 1;
        for (size_t head = 0; head < queue.size(); ++head) {
            const Pending pd = queue[head];
            const BinNode& b = bn[pd.bnode];
            t.sah += half_area(b.box) / root_area;
            int ch[8], nch = 0;
            ch[nch++] = b.left;
            ch[nch++] = b.right;
            while (nch < 8) {
                int pick = -1;
                double best = -1.0;
                for (int i = 0; i < nch; ++i) {
                    if (bn[ch[i]].part >= 0) continue;
                    const double a = half_area(bn[ch[i]].box);
                    if (a > best) {
                        best = a;
                        pick = i;
                    }
                }
                if (pick < 0) break;
                const int l = bn[ch[pick]].left;
                ch[nch++] = bn[ch[pick]].right;
                ch[pick] = l;
            }
            float clo[8][3], chi[8][3];
            for (int i = 0; i < nch; ++i)
                for (int a = 0; a < 3; ++a) {
                    clo[i][a] = bn[ch[i]].box.lo[a];
                    chi[i][a] = bn[ch[i]].box.hi[a];
                }
            int slot_of[8];
            assign_octant_slots(b.box.lo, b.box.hi, clo, chi, nch, slot_of);
            int in_slot[8];
            for (int s = 0; s < 8; ++s) in_slot[s] = -1;
            for (int i = 0; i < nch; ++i) in_slot[slot_of[i]] = i;
            Node8 node{};
            node.e.dst = pd.pos;
            node.e.child_base = next;
            node.depth = pd.depth;
            for (int s = 0; s < 8; ++s) {
                if (in_slot[s] < 0) continue;
                node.e.meta[s] = kMetaInternal;
                node.e.imask |= (uint8_t)(1u << s);
                const BinNode& c = bn[ch[in_slot[s]]];
                const uint32_t pos = next++;
                if (c.part >= 0) {
                    t.root_pos[c.part] = pos;
                    node.kids.push_back(-(c.part + 1));
                    t.max_depth = std::max(t.max_depth, pd.depth + st.parts[c.part].levels);
                } else {
                    node.kids.push_back((int)queue.size());  // its index in `top`
                    queue.push_back({ch[in_slot[s]], pos, pd.depth + 1});
                }
            }
            top.push_back(node);
        }
    }
    // Hidden parts' roots take the remaining positions (refit, unreachable).
    for (uint32_t i = 0; i < n; ++i)
        if (!visible[i]) t.root_pos[i] = next++;
    if (top.empty()) return t;
    if (vis.empty()) t.max_depth = 1;
    // Heights: top nodes were appended breadth first, so children follow
    // their parents.
    std::vector<uint32_t> height(top.size(), 0);
    for (size_t k = top.size(); k-- > 0;) {
        uint32_t h = 0;
        for (int kid : top[k].kids)
            h = std::max(h, 1 + (kid < 0 ? st.parts[-kid - 1].height : height[kid]));
        height[k] = h;
    }
    const uint32_t hmax = *std::max_element(height.begin(), height.end());
    t.level_off.assign(hmax + 2, 0);
    for (uint32_t h : height) t.level_off[h + 1]++;
    for (uint32_t h = 0; h <= hmax; ++h) t.level_off[h + 1] += t.level_off[h];
    t.entries.resize(top.size());
    std::vector<uint32_t> fill(t.level_off.begin(), t.level_off.end() - 1);
    for (size_t k = 0; k < top.size(); ++k) t.entries[fill[height[k]]++] = top[k].e;
    t.num_nodes = (uint32_t)top.size();
    return t;
}

// Places the parts with the given transforms and visibility: the top tree on
// the host, then triangle placement and the refit on the device.  Nothing on
// the device changes when it fails before the first launch.
lmb_status place_parts(lmb_scene s, const std::vector<PartTransform>& xf,
                       const std::vector<uint8_t>& visible, double host_ms_before) {
    auto t0 = std::chrono::steady_clock::now();
    InstanceState& st = *s->inst;
    const uint32_t n = (uint32_t)st.parts.size();
    std::vector<WBox> boxes(n);
    for (uint32_t i = 0; i < n; ++i) {
        if (!world_box(st.parts[i], xf[i].m, &boxes[i]))
            return fail(LMB_ERR_INVALID_ARGUMENT,
                        "instance " + std::to_string(st.parts[i].instance) +
                            ": the transformed geometry is not finite or exceeds 1e30");
    }
    TopTree top = build_top(st, boxes, visible);
    if (top.max_depth > (uint32_t)kStackMax)
        return fail(LMB_ERR_BVH_TOO_DEEP, "BVH depth " + std::to_string(top.max_depth) +
                                              " exceeds the traversal stack (" +
                                              std::to_string(kStackMax) + ")");
    hipError_t e = hipSetDevice(s->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    // Work already queued may read the BVH being replaced.
    e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail_hip(e, "hipDeviceSynchronize");
    e = hipMemcpy(st.xf.ptr, xf.data(), n * sizeof(PartTransform), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy transforms");
    e = hipMemcpy(st.root_pos.ptr, top.root_pos.data(), n * sizeof(uint32_t),
                  hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy root positions");
    if (!top.entries.empty()) {
        e = hipMemcpy(st.top.ptr, top.entries.data(), top.entries.size() * sizeof(RefitEntry),
                      hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail_hip(e, "hipMemcpy top tree");
    }
    e = hipMemset(st.error.ptr, 0, sizeof(uint32_t));
    if (e != hipSuccess) return fail_hip(e, "hipMemset");
    e = launch_instance_place(s->obj_tris.as<float4>(), s->tris.as<LeafTri>(),
                              (uint32_t)(s->tris.bytes / sizeof(LeafTri)), st.first.as<uint32_t>(),
                              n, st.xf.as<PartTransform>(), nullptr);
    if (e != hipSuccess) return fail_hip(e, "instance placement launch");
    const uint32_t hs = (uint32_t)st.level_off.size() - 1;
    const uint32_t ht = top.level_off.empty() ? 0 : (uint32_t)top.level_off.size() - 1;
    const RefitEntry* stat = st.entries.as<RefitEntry>();
    const RefitEntry* dyn = st.top.as<RefitEntry>();
    for (uint32_t h = 0; h < std::max(hs, ht); ++h) {
        const uint32_t a0 = h < hs ? st.level_off[h] : 0, a1 = h < hs ? st.level_off[h + 1] : 0;
        const uint32_t b0 = h < ht ? top.level_off[h] : 0, b1 = h < ht ? top.level_off[h + 1] : 0;
        e = launch_refit(stat + a0, a1 - a0, dyn + b0, b1 - b0, st.root_pos.as<uint32_t>(),
                         s->tris.as<LeafTri>(), s->nodes.as<Bvh8Node>(), st.box.as<float4>(),
                         st.error.as<uint32_t>(), nullptr);
        if (e != hipSuccess) return fail_hip(e, "refit launch");
    }
    uint32_t err = 0;
    e = hipMemcpy(&err, st.error.ptr, sizeof err, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy refit status");
    if (err) return fail(LMB_ERR_DEVICE, "BVH refit produced a non-conservative box (internal error)");
    // The scene's box (the cache's input normalisation) is the root's exact box.
    float4 root_box[2];
    e = hipMemcpy(root_box, st.box.ptr, sizeof root_box, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy root box");
    if (root_box[0].x <= root_box[1].x) {
        const float lo[3] = {root_box[0].x, root_box[0].y, root_box[0].z};
        const float hi[3] = {root_box[1].x, root_box[1].y, root_box[1].z};
        for (int k = 0; k < 3; ++k) {
            s->bounds_lo[k] = lo[k];
            s->bounds_hi[k] = hi[k];
        }
    }
    // Statistics of the tree now in place.
    uint64_t nodes = top.num_nodes;
    double sah = top.sah;
    WBox all{};
    bool any = false;
    for (uint32_t i = 0; i < n; ++i)
        if (visible[i]) {
            all = any ? merge(all, boxes[i]) : boxes[i];
            any = true;
        }
    const double root_area = any ? std::max(1e-30, half_area(all)) : 1.0;
    for (uint32_t i = 0; i < n; ++i)
        if (visible[i]) {
            nodes += st.parts[i].num_nodes;
            sah += st.parts[i].sah * half_area(boxes[i]) / root_area;
        }
    s->stats.num_nodes = nodes;
    s->stats.max_depth = top.max_depth;
    s->stats.sah_cost = sah;
    s->stats.build_ms =
        host_ms_before +
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return LMB_OK;
}

PartTransform part_transform(const float* m) {
    PartTransform x{};
    std::memcpy(x.m, m, sizeof x.m);
    x.mirrored = det3(m) < 0.0 ? 1u : 0u;
    return x;
}

bool finite_transform(const float* m) {
    for (int k = 0; k < 12; ++k)
        if (!std::isfinite(m[k])) return false;
    return true;
}

}  // namespace

}  // namespace lmb

using namespace lmb;

// TLAS instances over BLAS geometries (main.rs:464-563, general.wesl:1052-1059).
lmb_status lmb_scene_create_instanced(const lmb_instanced_desc* desc, lmb_scene* out) {
    auto t0 = std::chrono::steady_clock::now();
    if (!desc || !out) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (desc->num_geometries > 0 && !desc->geometries)
        return fail(LMB_ERR_INVALID_ARGUMENT, "geometries is NULL");
    if (desc->num_instances > 0 && !desc->instances)
        return fail(LMB_ERR_INVALID_ARGUMENT, "instances is NULL");
    if (desc->num_materials == 0 || !desc->materials)
        return fail(LMB_ERR_INVALID_ARGUMENT, "at least one material is required");
    if (desc->flags & ~LMB_SCENE_ALPHA_TEST)
        return fail(LMB_ERR_INVALID_ARGUMENT, "unknown scene flags");
    for (uint32_t gi = 0; gi < desc->num_geometries; ++gi) {
        const lmb_geometry_desc& g = desc->geometries[gi];
        if (g.num_aabbs == 0) continue;
        if (!g.aabbs) return fail(LMB_ERR_INVALID_ARGUMENT, "geometry aabbs are NULL");
        if (g.num_vertices > 0 || g.num_indices > 0)
            return fail(LMB_ERR_INVALID_ARGUMENT, "geometry " + std::to_string(gi) +
                                                      " holds both triangles and AABBs");
        for (uint64_t b = 0; b < g.num_aabbs; ++b) {
            const float* a = g.aabbs + 6 * b;
            for (int c = 0; c < 3; ++c)
                if (!(a[c] <= a[c + 3]) || !std::isfinite(a[c]) || !std::isfinite(a[c + 3]))
                    return fail(LMB_ERR_INVALID_ARGUMENT,
                                "AABB " + std::to_string(b) + " of geometry " +
                                    std::to_string(gi) + " is not finite with min <= max");
        }
    }
    auto state = std::make_unique<InstanceState>();
    state->num_instances = desc->num_instances;
    state->geometry.resize(desc->num_instances);
    state->part_of.assign(desc->num_instances, -1);
    state->hidden_at_creation.assign(desc->num_instances, 0);
    std::vector<float> verts, obj;
    std::vector<uint32_t> mat_idx;
    struct Pending {
        uint32_t instance;
        uint64_t prim0, ntri;
    };
    std::vector<Pending> pending;
    for (uint32_t k = 0; k < desc->num_instances; ++k) {
        const lmb_instance& in = desc->instances[k];
        if (in.geometry >= desc->num_geometries)
            return fail(LMB_ERR_INVALID_ARGUMENT,
                        "instance " + std::to_string(k) + " names a missing geometry");
        state->geometry[k] = in.geometry;
        if (!finite_transform(in.transform))
            return fail(LMB_ERR_INVALID_ARGUMENT,
                        "instance " + std::to_string(k) + " has a non-finite transform");
        const lmb_geometry_desc& g = desc->geometries[in.geometry];
        if (g.num_aabbs > 0) {
            // Procedural (AABB) primitives: the reference hands each candidate
            // to its intersection handler (general.wesl:1068-1076), and the
            // default handler (default_intersection_handler.wesl) reports no
            // hit, so they never take part in a closest hit.
            if ((in.mask & 0xFFu) && !in.material_idx)
                return fail(LMB_ERR_INVALID_ARGUMENT,
                            "instance " + std::to_string(k) + " has no material_idx");
            continue;
        }
        if (g.num_vertices > 0 && !g.vertices)
            return fail(LMB_ERR_INVALID_ARGUMENT, "geometry vertices are NULL");
        if (g.indices && g.num_indices % 3 != 0)
            return fail(LMB_ERR_INVALID_ARGUMENT, "index count is not a multiple of 3");
        const uint64_t ntri = g.indices ? g.num_indices / 3 : g.num_vertices / 3;
        if ((in.mask & 0xFFu) == 0) {
            state->hidden_at_creation[k] = ntri > 0 ? 1 : 0;
            continue;
        }
        if (ntri > 0 && !in.material_idx)
            return fail(LMB_ERR_INVALID_ARGUMENT,
                        "instance " + std::to_string(k) + " has no material_idx");
        const float* m = in.transform;
        const double det = det3(m);
        for (uint64_t t = 0; t < 3 * ntri; ++t) {
            const uint64_t vi = g.indices ? g.indices[t] : t;
            if (vi >= g.num_vertices)
                return fail(LMB_ERR_INVALID_ARGUMENT, "index out of range");
            const float x = g.vertices[4 * vi], y = g.vertices[4 * vi + 1], z = g.vertices[4 * vi + 2];
            // 3x4 row-major object-to-world, summed left to right (the
            // placement kernel computes the same expression)
            verts.push_back(((m[0] * x + m[1] * y) + m[2] * z) + m[3]);
            verts.push_back(((m[4] * x + m[5] * y) + m[6] * z) + m[7]);
            verts.push_back(((m[8] * x + m[9] * y) + m[10] * z) + m[11]);
            verts.push_back(0.0f);
            obj.insert(obj.end(), {x, y, z, (t % 3 == 0 && det < 0.0) ? 1.0f : 0.0f});
        }
        if (ntri > 0) pending.push_back({k, mat_idx.size(), ntri});
        mat_idx.insert(mat_idx.end(), in.material_idx, in.material_idx + ntri);
    }
    const uint64_t ntri_all = mat_idx.size();
    if (ntri_all >= 0xFFFFFFFFull) return fail(LMB_ERR_INVALID_ARGUMENT, "too many triangles");
    for (uint64_t t = 0; t < ntri_all; ++t)
        if (mat_idx[t] >= desc->num_materials)
            return fail(LMB_ERR_INVALID_ARGUMENT,
                        "material_idx of primitive " + std::to_string(t) + " is out of range");

    // One subtree per instance, built over its world-space triangles, on up
    // to hardware_concurrency threads.
    std::vector<BvhBuildResult> builds(pending.size());
    std::vector<std::string> errs(pending.size());
    std::vector<int> rcs(pending.size(), 0);
    {
        const size_t workers =
            std::max<size_t>(1, std::min<size_t>(pending.size(), std::thread::hardware_concurrency()));
        std::atomic<size_t> next_build{0};
        auto work = [&] {
            for (size_t i; (i = next_build.fetch_add(1)) < pending.size();) {
                const Pending& pd = pending[i];
                rcs[i] = build_bvh8(verts.data() + 4 * 3 * pd.prim0, 3 * pd.ntri, nullptr,
                                    pd.ntri, &builds[i], &errs[i]);
            }
        };
        std::vector<std::thread> pool;
        for (size_t w = 1; w < workers; ++w) pool.emplace_back(work);
        work();
        for (std::thread& th : pool) th.join();
    }
    for (size_t i = 0; i < pending.size(); ++i)
        if (rcs[i] != 0) return fail(LMB_ERR_INVALID_ARGUMENT, errs[i]);
    // Layout: node positions [0, region) for the top tree and the part roots,
    // then each part's other nodes in a block of its own.
    std::vector<uint32_t> first{0};
    uint32_t num_parts = 0;
    for (size_t i = 0; i < pending.size(); ++i) {
        if (builds[i].nodes.empty()) continue;  // no finite triangle
        num_parts++;
    }
    const uint32_t region = num_parts ? 2 * num_parts : 0;
    uint64_t capacity = region;
    std::vector<LeafTri> tris;
    tris.reserve(ntri_all);
    std::vector<RefitEntry> entries;
    std::vector<uint32_t> heights;
    for (size_t i = 0; i < pending.size(); ++i) {
        const BvhBuildResult& b = builds[i];
        if (b.nodes.empty()) continue;
        const Pending& pd = pending[i];
        InstancePart part;
        part.instance = pd.instance;
        part.first_tri = (uint32_t)tris.size();
        part.num_tris = (uint32_t)b.tris.size();
        part.levels = b.max_depth;
        part.num_nodes = (uint32_t)b.nodes.size();
        part.sah = b.sah_cost;
        for (int a = 0; a < 3; ++a) {
            part.obj_lo[a] = INFINITY;
            part.obj_hi[a] = -INFINITY;
        }
        const uint32_t pidx = (uint32_t)state->parts.size();
        const uint64_t block = capacity;  // position of local node 1
        capacity += b.nodes.size() - 1;
        if (capacity >= 0x7FFFFFFFull) return fail(LMB_ERR_INVALID_ARGUMENT, "too many BVH nodes");
        for (const LeafTri& lt : b.tris) {
            LeafTri g = lt;
            uint32_t id;
            std::memcpy(&id, &g.v0[3], 4);
            id += (uint32_t)pd.prim0;
            std::memcpy(&g.v0[3], &id, 4);
            std::memcpy(&g.v1[3], &mat_idx[id], 4);
            for (int v = 0; v < 3; ++v)
                for (int a = 0; a < 3; ++a) {
                    const float c = obj[12 * (uint64_t)id + 4 * v + a];
                    part.obj_lo[a] = std::min(part.obj_lo[a], c);
                    part.obj_hi[a] = std::max(part.obj_hi[a], c);
                }
            tris.push_back(g);
        }
        first.push_back((uint32_t)tris.size());
        // Node heights: children follow their parents (breadth-first layout).
        std::vector<uint32_t> h(b.nodes.size(), 0);
        for (size_t j = b.nodes.size(); j-- > 0;) {
            const Bvh8Node& nd = b.nodes[j];
            uint32_t hh = 0;
            for (uint32_t k = 0; k < (uint32_t)__builtin_popcount(nd.imask); ++k)
                hh = std::max(hh, 1 + h[nd.child_base + k]);
            h[j] = hh;
        }
        part.height = h[0];
        for (size_t j = 0; j < b.nodes.size(); ++j) {
            const Bvh8Node& nd = b.nodes[j];
            RefitEntry e{};
            e.dst = j == 0 ? (kRootOfPart | pidx) : (uint32_t)(block + j - 1);
            e.child_base = nd.imask ? (uint32_t)(block + nd.child_base - 1) : 0u;
            e.tri_base = nd.tri_base + part.first_tri;
            e.imask = nd.imask;
            std::memcpy(e.meta, nd.meta, 8);
            entries.push_back(e);
            heights.push_back(h[j]);
        }
        state->part_of[pd.instance] = (int32_t)pidx;
        state->parts.push_back(part);
    }
    state->region = region;
    // Static entries by height.
    uint32_t hmax = 0;
    for (uint32_t h : heights) hmax = std::max(hmax, h);
    state->level_off.assign(entries.empty() ? 1 : hmax + 2, 0);
    for (uint32_t h : heights) state->level_off[h + 1]++;
    for (size_t h = 0; h + 1 < state->level_off.size(); ++h)
        state->level_off[h + 1] += state->level_off[h];
    std::vector<RefitEntry> sorted(entries.size());
    {
        std::vector<uint32_t> fill(state->level_off.begin(), state->level_off.end() - 1);
        for (size_t k = 0; k < entries.size(); ++k) sorted[fill[heights[k]]++] = entries[k];
    }

    lmb_scene_desc flat{};
    flat.vertices = verts.data();
    flat.num_vertices = verts.size() / 4;
    flat.material_idx = mat_idx.data();
    flat.materials = desc->materials;
    flat.num_materials = desc->num_materials;
    flat.textures = desc->textures;
    flat.flags = desc->flags;
    BvhBuildResult bvh;
    bvh.nodes.assign(capacity, Bvh8Node{});  // filled by the refit
    bvh.tris = std::move(tris);
    lmb_scene s = new lmb_scene_s();
    auto cleanup = [&](lmb_status code) {
        scene_free(s);
        delete s;
        return code;
    };
    lmb_status st = scene_upload(s, &flat, ntri_all, bvh);
    if (st) return cleanup(st);
    s->dev.alpha_test = (desc->flags & LMB_SCENE_ALPHA_TEST) ? 1u : 0u;
    if (!obj.empty()) {
        // Shading reads the object-space triangle (the reference's vertex
        // fetch returns object space, bindings.wesl:102-110).
        if ((st = s->obj_tris.upload(obj.data(), obj.size() * sizeof(float)))) return cleanup(st);
        s->dev.obj_tris = s->obj_tris.as<float4>();
        s->stats.device_bytes += obj.size() * sizeof(float);
    }
    InstanceState& is = *state;
    s->inst = std::move(state);
    const uint32_t np = (uint32_t)is.parts.size();
    if (np == 0) {
        s->stats.num_nodes = 0;
        s->stats.max_depth = 0;
        s->dev.num_nodes = 0;
        *out = s;
        return LMB_OK;
    }
    if ((st = is.entries.upload(sorted.data(), sorted.size() * sizeof(RefitEntry))) ||
        (st = is.first.upload(first.data(), first.size() * sizeof(uint32_t))) ||
        (st = is.top.alloc((size_t)(np + 1) * sizeof(RefitEntry))) ||
        (st = is.root_pos.alloc((size_t)np * sizeof(uint32_t))) ||
        (st = is.xf.alloc((size_t)np * sizeof(PartTransform))) ||
        (st = is.box.alloc((size_t)capacity * 2 * sizeof(float4))) ||
        (st = is.error.alloc(sizeof(uint32_t))))
        return cleanup(st);
    s->stats.device_bytes += is.entries.bytes + is.first.bytes + is.top.bytes + is.root_pos.bytes +
                             is.xf.bytes + is.box.bytes + is.error.bytes;
    std::vector<PartTransform> xf(np);
    std::vector<uint8_t> visible(np, 1);
    for (uint32_t i = 0; i < np; ++i) xf[i] = part_transform(desc->instances[is.parts[i].instance].transform);
    const double host_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if ((st = place_parts(s, xf, visible, host_ms))) return cleanup(st);
    *out = s;
    return LMB_OK;
}

// New transforms and masks for the instances of an instanced scene: the
// counterpart of writing TlasInstance entries and rebuilding the TLAS
// (main.rs:560-563).
lmb_status lmb_scene_update_instances(lmb_scene scene, const lmb_instance* instances,
                                      uint32_t num_instances) {
    if (!scene) return fail(LMB_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (!scene->inst)
        return fail(LMB_ERR_INVALID_ARGUMENT,
                    "the scene was not made by lmb_scene_create_instanced");
    InstanceState& is = *scene->inst;
    if (num_instances != is.num_instances)
        return fail(LMB_ERR_INVALID_ARGUMENT,
                    "the scene has " + std::to_string(is.num_instances) + " instances, not " +
                        std::to_string(num_instances));
    if (num_instances > 0 && !instances) return fail(LMB_ERR_INVALID_ARGUMENT, "instances is NULL");
    for (uint32_t k = 0; k < num_instances; ++k) {
        const lmb_instance& in = instances[k];
        if (in.geometry != is.geometry[k])
            return fail(LMB_ERR_INVALID_ARGUMENT,
                        "instance " + std::to_string(k) + " names another geometry than at creation");
        if (!finite_transform(in.transform))
            return fail(LMB_ERR_INVALID_ARGUMENT,
                        "instance " + std::to_string(k) + " has a non-finite transform");
        if (is.hidden_at_creation[k] && (in.mask & 0xFFu))
            return fail(LMB_ERR_INVALID_ARGUMENT,
                        "instance " + std::to_string(k) +
                            " was hidden (mask 0) when the scene was made, so its triangles are "
                            "not in it");
    }
    const uint32_t np = (uint32_t)is.parts.size();
    if (np == 0) return LMB_OK;
    std::vector<PartTransform> xf(np);
    std::vector<uint8_t> visible(np);
    for (uint32_t i = 0; i < np; ++i) {
        const lmb_instance& in = instances[is.parts[i].instance];
        xf[i] = part_transform(in.transform);
        visible[i] = (in.mask & 0xFFu) ? 1 : 0;
    }
    return place_parts(scene, xf, visible, 0.0);
}
