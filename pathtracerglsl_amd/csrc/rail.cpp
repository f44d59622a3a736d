// rail.cpp — the binned-SAH binary BVH core, the data ledger and the mesh-channel packing
// (host, C++17). Every function here restates a reference function named beside it, in its
// operation order and precision: the outputs are the reference's bytes.
#include "rail.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <utility>

namespace ptmi {

pt_status fail(pt_status s, const std::string& msg);

namespace {

constexpr int BINS = 12;   // BVH_BINS (accel/bvh/bvh.ts:37)

// buildBVHCore (accel/bvh/bvh.ts:132-320): count-weighted for the ray trees (Sigma 1 is exact in
// f64, so the weighted cost is the area x count cost), emitting the flat node format of
// buildBVHNodesFlat (:344-383); power-weighted with leaf size 1, always split, the 48-level
// depth cap, the sorted-median fallback and the bit trails for the light tree
// (accel/light_tree/light_tree.ts:46-77), emitting its (min, phi) (max, link) nodes.
struct Core {
    const double* boxes;
    const double* weights = nullptr;  // per-item split-cost weight (null: the count)
    int leaf_size;
    bool always_split, sorted_median;
    int depth_cap = -1;               // < 0: none
    float* trails = nullptr;          // light tree: per item (lo 24 bits, hi 24 bits, 0, 0)
    std::vector<float> c[3];          // f32 centroid stores
    std::vector<uint32_t> order;
    std::vector<float>* nodes;
    int node_count = 0, max_depth = 0;
    static int ceil_log2(int c) { return c <= 1 ? 0 : 32 - __builtin_clz((unsigned)(c - 1)); }

    static void reset(double* b) { b[0] = b[1] = b[2] = INFINITY; b[3] = b[4] = b[5] = -INFINITY; }
    static void grow(double* b, const double* s) {
        for (int a = 0; a < 3; a++) {
            if (s[a] < b[a]) b[a] = s[a];
            if (s[3 + a] > b[3 + a]) b[3 + a] = s[3 + a];
        }
    }
    static double area(const double* b) {
        const double dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0;
        return 2 * (dx * dy + dy * dz + dz * dx);
    }
    void set(int ni, const double* b, double A, double B) {
        float* nd = &(*nodes)[8 * (size_t)ni];
        nd[0] = (float)b[0]; nd[1] = (float)b[1]; nd[2] = (float)b[2]; nd[3] = (float)A;
        nd[4] = (float)b[3]; nd[5] = (float)b[4]; nd[6] = (float)b[5]; nd[7] = (float)B;
    }

    // The leaf over order[start, end): the ray trees' (count, start), or the light tree's
    // (phi, -(light + 1)) with the light's trail.
    int leaf(int ni, const double* bounds, int start, int count, double w, uint32_t lo, uint32_t hi) {
        if (!trails) { set(ni, bounds, count, start); return ni; }
        const uint32_t light = order[start];
        set(ni, bounds, w, -(double)light - 1.0);
        trails[4 * (size_t)light] = (float)lo;
        trails[4 * (size_t)light + 1] = (float)hi;
        return ni;
    }

    // lo / hi: the trail down to this node, one bit per level, LSB first, 24 per half
    int emit(int start, int end, int depth, uint32_t lo = 0, uint32_t hi = 0) {
        if (depth > max_depth) max_depth = depth;
        const int ni = node_count++;
        double bounds[6], cb[6];
        reset(bounds); reset(cb);
        double w = 0;
        for (int i = start; i < end; i++) {
            const uint32_t t = order[i];
            const double* bx = boxes + 6 * (size_t)t;
            for (int a = 0; a < 3; a++) {
                if (bx[a] < bounds[a]) bounds[a] = bx[a];
                if (bx[3 + a] > bounds[3 + a]) bounds[3 + a] = bx[3 + a];
                const double cc = c[a][t];
                if (cc < cb[a]) cb[a] = cc;
                if (cc > cb[3 + a]) cb[3 + a] = cc;
            }
            if (weights) w += weights[t];
        }
        const int count = end - start;
        if (count <= leaf_size) return leaf(ni, bounds, start, count, w, lo, hi);
        // the depth cap: once depth + ceil(log2 count) reaches it, only balanced splits
        const bool balanced_only = depth_cap >= 0 && depth + ceil_log2(count) >= depth_cap;

        // binned SAH over every axis with a centroid extent >= 1e-12: per-axis prefix and
        // suffix sweeps, f32 area stores, axes then bins ascending, strict < (the tie order);
        // cost = A_L w_L + A_R w_R with w the item count or the summed weights
        int best_axis = -1, best_split = -1;
        double best_cost = INFINITY;
        bool valid[3];
        for (int a = 0; a < 3; a++) valid[a] = cb[3 + a] - cb[a] >= 1e-12;
        if (!balanced_only && (valid[0] || valid[1] || valid[2])) {
            double bin_box[3][BINS][6];
            int bin_cnt[3][BINS];
            double bin_w[3][BINS];
            double scale[3];
            for (int a = 0; a < 3; a++) {
                scale[a] = valid[a] ? BINS / (cb[3 + a] - cb[a]) : 0;
                for (int b = 0; b < BINS; b++) { reset(bin_box[a][b]); bin_cnt[a][b] = 0; bin_w[a][b] = 0; }
            }
            for (int i = start; i < end; i++) {
                const uint32_t t = order[i];
                const double* bx = boxes + 6 * (size_t)t;
                for (int a = 0; a < 3; a++) {
                    if (!valid[a]) continue;
                    int b = (int)std::floor(((double)c[a][t] - cb[a]) * scale[a]);
                    if (b < 0) b = 0;
                    if (b >= BINS) b = BINS - 1;
                    bin_cnt[a][b] += 1;
                    bin_w[a][b] += weights ? weights[t] : 1;
                    grow(bin_box[a][b], bx);
                }
            }
            for (int a = 0; a < 3; a++) {
                if (!valid[a]) continue;
                float left_area[BINS - 1], right_area[BINS - 1];
                int left_cnt[BINS - 1], right_cnt[BINS - 1];
                double left_w[BINS - 1], right_w[BINS - 1];
                double acc[6];
                int n = 0;
                double wa = 0;
                reset(acc);
                for (int b = 0; b < BINS - 1; b++) {
                    grow(acc, bin_box[a][b]); n += bin_cnt[a][b]; wa += bin_w[a][b];
                    left_area[b] = (float)area(acc); left_cnt[b] = n; left_w[b] = wa;
                }
                reset(acc); n = 0; wa = 0;
                for (int b = BINS - 1; b > 0; b--) {
                    grow(acc, bin_box[a][b]); n += bin_cnt[a][b]; wa += bin_w[a][b];
                    right_area[b - 1] = (float)area(acc); right_cnt[b - 1] = n; right_w[b - 1] = wa;
                }
                for (int b = 0; b < BINS - 1; b++) {
                    if (left_cnt[b] == 0 || right_cnt[b] == 0) continue;
                    const double cost = (double)left_area[b] * left_w[b] + (double)right_area[b] * right_w[b];
                    if (cost < best_cost) { best_cost = cost; best_axis = a; best_split = b; }
                }
            }
        }
        if (!always_split) {
            const double leaf_cost = area(bounds) * (weights ? w : count);
            if (best_axis == -1 || best_cost >= leaf_cost) return leaf(ni, bounds, start, count, w, lo, hi);
        }
        int mid = -1;
        if (best_axis != -1) {
            // the in-place swap partition by the winning axis and plane
            const double plo = cb[best_axis], sc = BINS / (cb[3 + best_axis] - cb[best_axis]);
            const std::vector<float>& cen = c[best_axis];
            int m = start;
            for (int i = start; i < end; i++) {
                const uint32_t t = order[i];
                int b = (int)std::floor(((double)cen[t] - plo) * sc);
                if (b < 0) b = 0;
                if (b >= BINS) b = BINS - 1;
                if (b <= best_split) { std::swap(order[i], order[m]); m++; }
            }
            if (m > start && m < end) mid = m;
        }
        if (mid == -1) {
            if (sorted_median) {
                // the slice sorted by centroid along the widest centroid axis (a stable sort, as
                // Array.prototype.sort is), then halved
                const double e0 = cb[3] - cb[0], e1 = cb[4] - cb[1], e2 = cb[5] - cb[2];
                const int axis = e0 >= e1 && e0 >= e2 ? 0 : e1 >= e2 ? 1 : 2;
                const std::vector<float>& cen = c[axis];
                std::stable_sort(order.begin() + start, order.begin() + end,
                                 [&](uint32_t x, uint32_t y) { return (double)cen[x] - (double)cen[y] < 0; });
            }
            mid = (start + end) >> 1;
        }
        double keep[6];
        std::memcpy(keep, bounds, sizeof keep);
        const uint32_t rlo = depth < 24 ? lo | (1u << depth) : lo;
        const uint32_t rhi = depth < 24 ? hi : hi | (1u << (depth - 24));
        emit(start, mid, depth + 1, lo, hi);                 // left child at ni + 1: trail bit 0
        const int right = emit(mid, end, depth + 1, rlo, rhi);   // right: bit 1 at this level
        if (trails) set(ni, keep, w, right);                 // the light tree's (phi, right child)
        else set(ni, keep, -1 - std::max(best_axis, 0), right);   // a fallback split stores axis 0
        return ni;
    }
};

}  // namespace

void build_bvh_flat(const double* boxes, int n, int leaf_size, bool always_split, bool sorted_median, FlatBVH& out) {
    out = FlatBVH();
    out.nodes.assign(n > 0 ? 8 * (size_t)(2 * n - 1) : 0, 0.0f);
    out.order.resize((size_t)std::max(n, 0));
    if (n <= 0) return;
    Core C;
    C.boxes = boxes; C.leaf_size = leaf_size; C.always_split = always_split; C.sorted_median = sorted_median;
    for (int a = 0; a < 3; a++) C.c[a].resize((size_t)n);
    for (int t = 0; t < n; t++) {
        const double* b = boxes + 6 * (size_t)t;
        C.c[0][t] = (float)((b[0] + b[3]) * 0.5);
        C.c[1][t] = (float)((b[1] + b[4]) * 0.5);
        C.c[2][t] = (float)((b[2] + b[5]) * 0.5);
    }
    C.order.resize((size_t)n);
    for (int i = 0; i < n; i++) C.order[i] = (uint32_t)i;
    C.nodes = &out.nodes;
    C.emit(0, n, 0);
    out.nodes.resize(8 * (size_t)C.node_count);
    out.order = std::move(C.order);
    out.node_count = C.node_count;
    out.max_depth = C.max_depth;
}

void build_light_tree(const double* boxes, const double* powers, int n, std::vector<float>& nodes,
                      std::vector<float>& trails, int& node_count, int& max_depth) {
    nodes.assign(n > 0 ? 8 * (size_t)(2 * n - 1) : 0, 0.0f);
    trails.assign(4 * (size_t)std::max(n, 0), 0.0f);
    node_count = max_depth = 0;
    if (n <= 0) return;
    Core C;
    C.boxes = boxes; C.weights = powers; C.leaf_size = 1; C.always_split = true; C.sorted_median = true;
    C.depth_cap = LIGHT_TREE_MAX_DEPTH; C.trails = trails.data();
    for (int a = 0; a < 3; a++) C.c[a].resize((size_t)n);
    for (int t = 0; t < n; t++) {
        const double* b = boxes + 6 * (size_t)t;
        C.c[0][t] = (float)((b[0] + b[3]) * 0.5);
        C.c[1][t] = (float)((b[1] + b[4]) * 0.5);
        C.c[2][t] = (float)((b[2] + b[5]) * 0.5);
    }
    C.order.resize((size_t)n);
    for (int i = 0; i < n; i++) C.order[i] = (uint32_t)i;
    C.nodes = &nodes;
    C.emit(0, n, 0);
    node_count = C.node_count;
    max_depth = C.max_depth;
}

void build_mesh_bvh(const float* positions, const uint32_t* indices, int n_tris, FlatBVH& out, std::vector<uint32_t>& reindexed) {
    std::vector<double> box(6 * (size_t)std::max(n_tris, 0));
    for (int t = 0; t < n_tris; t++) {
        double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 3; k++) {
            const float* v = positions + 3 * (size_t)indices[3 * (size_t)t + k];
            for (int a = 0; a < 3; a++) {
                if (v[a] < mn[a]) mn[a] = v[a];
                if (v[a] > mx[a]) mx[a] = v[a];
            }
        }
        for (int a = 0; a < 3; a++) { box[6 * (size_t)t + a] = mn[a]; box[6 * (size_t)t + 3 + a] = mx[a]; }
    }
    build_bvh_flat(box.data(), n_tris, 2, false, false, out);   // BVH_LEAF_SIZE
    reindexed.resize(3 * (size_t)std::max(n_tris, 0));
    for (int i = 0; i < n_tris; i++)
        for (int k = 0; k < 3; k++) reindexed[3 * (size_t)i + k] = indices[3 * (size_t)out.order[i] + k];
}

}  // namespace ptmi

using namespace ptmi;

namespace {
// nodeTexelBound (components/data/ledger.ts:117-119): <= 2T - 1 nodes, 2 texels each.
int64_t node_texel_bound(int64_t leaves) { return leaves > 0 ? 2 * (2 * leaves - 1) : 0; }
bool fits32(int64_t v) { return v >= 0 && v <= INT32_MAX; }
}  // namespace

extern "C" {

pt_status pt_bvh_build(const double* boxes, int32_t n, int32_t leaf_size, uint32_t flags, float* nodes_out,
                       uint32_t* order_out, int32_t* node_count, int32_t* max_depth) {
    if ((n > 0 && !boxes) || !nodes_out || !order_out || !node_count) return fail(PT_E_INVALID, "pt_bvh_build: null argument");
    if (n < 0 || leaf_size < 1) return fail(PT_E_INVALID, "pt_bvh_build: n must be >= 0 and leaf_size >= 1");
    if (flags & ~(uint32_t)(PT_BVH_ALWAYS_SPLIT | PT_BVH_SORTED_MEDIAN)) return fail(PT_E_INVALID, "pt_bvh_build: unknown flags");
    FlatBVH b;
    build_bvh_flat(boxes, n, leaf_size, (flags & PT_BVH_ALWAYS_SPLIT) != 0, (flags & PT_BVH_SORTED_MEDIAN) != 0, b);
    std::memcpy(nodes_out, b.nodes.data(), b.nodes.size() * sizeof(float));
    if (n > 0) std::memcpy(order_out, b.order.data(), (size_t)n * sizeof(uint32_t));
    *node_count = b.node_count;
    if (max_depth) *max_depth = b.max_depth;
    return PT_OK;
}

pt_status pt_light_tree_build(const double* boxes, const double* powers, int32_t n, float* nodes_out, float* trails_out,
                              int32_t* node_count, int32_t* max_depth) {
    if ((n > 0 && (!boxes || !powers || !nodes_out || !trails_out)) || !node_count)
        return fail(PT_E_INVALID, "pt_light_tree_build: null argument");
    if (n < 0 || n > (1 << 22)) return fail(PT_E_INVALID, "pt_light_tree_build: n must be in [0, 2^22]");
    for (int32_t i = 0; i < n; i++) {
        if (!(powers[i] > 0.0) || !std::isfinite(powers[i]))
            return fail(PT_E_INVALID, "pt_light_tree_build: light " + std::to_string(i) + " has a power that is not finite and > 0");
        for (int k = 0; k < 3; k++)
            if (!std::isfinite(boxes[6 * (size_t)i + k]) || !std::isfinite(boxes[6 * (size_t)i + 3 + k]) ||
                boxes[6 * (size_t)i + k] > boxes[6 * (size_t)i + 3 + k])
                return fail(PT_E_INVALID, "pt_light_tree_build: light " + std::to_string(i) + " has a bad box");
    }
    std::vector<float> nodes, trails;
    int nc = 0, md = 0;
    build_light_tree(boxes, powers, n, nodes, trails, nc, md);
    if (n > 0) {
        std::memcpy(nodes_out, nodes.data(), nodes.size() * sizeof(float));
        std::memcpy(trails_out, trails.data(), trails.size() * sizeof(float));
    }
    *node_count = nc;
    if (max_depth) *max_depth = md;
    return PT_OK;
}

pt_status pt_mesh_bvh_build(const float* positions, int32_t n_vertices, const uint32_t* indices, int32_t n_tris,
                            float* nodes_out, uint32_t* reindexed_out, int32_t* node_count, int32_t* max_depth) {
    if ((n_tris > 0 && (!positions || !indices)) || !nodes_out || !reindexed_out || !node_count)
        return fail(PT_E_INVALID, "pt_mesh_bvh_build: null argument");
    if (n_tris < 0 || n_vertices < 0) return fail(PT_E_INVALID, "pt_mesh_bvh_build: negative count");
    for (int64_t i = 0; i < 3 * (int64_t)n_tris; i++)
        if (indices[i] >= (uint32_t)n_vertices) return fail(PT_E_INVALID, "pt_mesh_bvh_build: index " + std::to_string(indices[i]) + " out of range");
    FlatBVH b;
    std::vector<uint32_t> re;
    build_mesh_bvh(positions, indices, n_tris, b, re);
    std::memcpy(nodes_out, b.nodes.data(), b.nodes.size() * sizeof(float));
    if (n_tris > 0) std::memcpy(reindexed_out, re.data(), re.size() * sizeof(uint32_t));
    *node_count = b.node_count;
    if (max_depth) *max_depth = b.max_depth;
    return PT_OK;
}

// planDataLayout (components/data/ledger.ts:122-184): always-built regions first (mesh geometry,
// batch placements / attrs / TLASes, mesh-light bakes), then the optional ones in a fixed order
// (CWBVH regions, the scene table, the light tree, the region-material ids), so an optional
// structure never moves an always-built base.
pt_status pt_plan_data_layout(const pt_data_tenants* t, pt_data_layout* out) {
    if (!t || !out) return fail(PT_E_INVALID, "pt_plan_data_layout: null argument");
    if (t->n_meshes < 0 || t->n_batches < 0 || t->n_mesh_lights < 0) return fail(PT_E_INVALID, "pt_plan_data_layout: negative count");
    if ((t->n_meshes && (!t->meshes || !out->meshes)) || (t->n_batches && (!t->batches || !out->batches)) ||
        (t->n_mesh_lights && (!t->mesh_lights || !out->mesh_lights)))
        return fail(PT_E_INVALID, "pt_plan_data_layout: missing tenant or slot array");
    int64_t v = 0, tr = 0, n = 0, r = 0, nq = 0;
    for (int i = 0; i < t->n_meshes; i++) {
        const pt_mesh_tenant& m = t->meshes[i];
        if (m.vertex_count < 0 || m.tri_count < 0) return fail(PT_E_INVALID, "pt_plan_data_layout: negative mesh count");
        out->meshes[i] = {(int32_t)v, (int32_t)tr, (int32_t)n};
        v += m.vertex_count;
        tr += m.tri_count;
        n += node_texel_bound(m.tri_count);
    }
    const int64_t mesh_vertices = v;
    for (int i = 0; i < t->n_batches; i++) {
        const pt_batch_tenant& b = t->batches[i];
        if (b.instance_count < 0 || (b.placement_texels != 1 && b.placement_texels != 2) || b.attr_texels < 0 ||
            b.cwbvh_node_texels < 0 || b.cwbvh_record_texels < 0)
            return fail(PT_E_INVALID, "pt_plan_data_layout: batch " + std::to_string(i) + " has a bad count or placement stride");
        pt_batch_slot& s = out->batches[i];
        s.placements_base = (int32_t)r; r += (int64_t)b.placement_texels * b.instance_count;
        s.attrs_base = b.attr_texels > 0 ? (int32_t)r : -1; r += b.attr_texels;
        s.tlas_base = (int32_t)n; n += node_texel_bound(b.instance_count);
        s.cwbvh_nodes_base = -1; s.cwbvh_records_base = -1;
    }
    for (int i = 0; i < t->n_mesh_lights; i++) {
        const pt_mesh_light_tenant& l = t->mesh_lights[i];
        if (l.vertex_count < 0 || l.tri_count < 0) return fail(PT_E_INVALID, "pt_plan_data_layout: negative mesh-light count");
        out->mesh_lights[i] = {l.mesh_ordinal, (int32_t)v, (int32_t)r};
        v += l.vertex_count;
        r += l.tri_count;
    }
    for (int i = 0; i < t->n_batches; i++) {
        const pt_batch_tenant& b = t->batches[i];
        if (b.cwbvh_record_texels > 0) { out->batches[i].cwbvh_records_base = (int32_t)r; r += b.cwbvh_record_texels; }
        if (b.cwbvh_node_texels > 0) { out->batches[i].cwbvh_nodes_base = (int32_t)nq; nq += b.cwbvh_node_texels; }
    }
    out->table_leaf_list_base = out->table_analytic_base = out->table_tlas_base = -1;
    if (t->has_scene_table && t->table_leaf_count > 0) {
        out->table_leaf_list_base = (int32_t)r; r += t->table_leaf_count;
        out->table_analytic_base = (int32_t)r; r += t->table_analytic_texels;
        out->table_tlas_base = (int32_t)n; n += node_texel_bound(t->table_leaf_count);
    }
    out->light_table_base = out->light_trails_base = out->light_tree_base = -1;
    if (t->has_light_tree && t->light_count > 0) {
        out->light_table_base = (int32_t)r; r += t->light_table_texels;
        out->light_trails_base = (int32_t)r; r += t->light_count;
        out->light_tree_base = (int32_t)n; n += node_texel_bound(t->light_count);
    }
    out->region_materials_base = -1;
    if (t->has_region_materials && t->region_count > 0) {
        out->region_materials_base = (int32_t)r;
        r += (t->region_count + 3) / 4;
    }
    if (!fits32(v) || !fits32(tr) || !fits32(n) || !fits32(r) || !fits32(nq))
        return fail(PT_E_INVALID, "pt_plan_data_layout: a channel exceeds 2^31 texels");
    out->total_vertices = (int32_t)v;
    out->total_normals = out->total_uvs = (int32_t)mesh_vertices;   // mesh geometry only (no light bakes)
    out->total_indices = (int32_t)tr;
    out->total_nodes = (int32_t)n;
    out->total_records = (int32_t)r;
    out->total_nodesq = (int32_t)nq;
    return PT_OK;
}

// packSceneData's geometry pass (app/sceneData.ts:36-47) with the texel writers of
// components/data/pack.ts (items at base + i, unused components left as the caller's zeros).
pt_status pt_pack_mesh_channels(const pt_data_layout* layout, int32_t n_meshes, const pt_mesh_view* meshes,
                                float* vertices, float* normals, float* uvs, float* indices, float* nodes, float* root_boxes) {
    if (!layout || (n_meshes > 0 && (!meshes || !layout->meshes))) return fail(PT_E_INVALID, "pt_pack_mesh_channels: null argument");
    if (n_meshes > 0 && (!vertices || !indices || !nodes)) return fail(PT_E_INVALID, "pt_pack_mesh_channels: missing channel");
    // a mesh object shared by several slots builds its BLAS once (the pack cache)
    std::map<std::pair<const float*, const uint32_t*>, std::pair<FlatBVH, std::vector<uint32_t>>> cache;
    for (int s = 0; s < n_meshes; s++) {
        const pt_mesh_view& m = meshes[s];
        const pt_mesh_slot& slot = layout->meshes[s];
        if (m.n_vertices < 0 || m.n_tris < 0 || (m.n_tris > 0 && (!m.positions || !m.indices)))
            return fail(PT_E_INVALID, "pt_pack_mesh_channels: mesh " + std::to_string(s) + " is malformed");
        for (int64_t i = 0; i < 3 * (int64_t)m.n_tris; i++)
            if (m.indices[i] >= (uint32_t)m.n_vertices)
                return fail(PT_E_INVALID, "pt_pack_mesh_channels: mesh " + std::to_string(s) + " index out of range");
        auto key = std::make_pair(m.positions, m.indices);
        auto it = cache.find(key);
        if (it == cache.end()) {
            std::pair<FlatBVH, std::vector<uint32_t>> built;
            build_mesh_bvh(m.positions, m.indices, m.n_tris, built.first, built.second);
            it = cache.emplace(key, std::move(built)).first;
        }
        const FlatBVH& b = it->second.first;
        const std::vector<uint32_t>& re = it->second.second;
        const int64_t region = node_texel_bound(m.n_tris);
        if ((int64_t)b.node_count * 2 > region)     // assertFits('mesh BLAS nodes', ...)
            return fail(PT_E_INVALID, "data rail: mesh BLAS nodes packed " + std::to_string(b.node_count * 2) + " texels into a " +
                                          std::to_string(region) + "-texel region");
        for (int i = 0; i < m.n_vertices; i++) {
            float* d = vertices + 4 * ((size_t)slot.vbase + i);
            for (int c = 0; c < 3; c++) d[c] = m.positions[3 * (size_t)i + c];
            if (m.normals && normals) { float* q = normals + 4 * ((size_t)slot.vbase + i); for (int c = 0; c < 3; c++) q[c] = m.normals[3 * (size_t)i + c]; }
            if (m.uvs && uvs) { float* q = uvs + 4 * ((size_t)slot.vbase + i); for (int c = 0; c < 2; c++) q[c] = m.uvs[2 * (size_t)i + c]; }
        }
        for (int i = 0; i < m.n_tris; i++) {
            float* d = indices + 4 * ((size_t)slot.tbase + i);
            for (int c = 0; c < 3; c++) d[c] = (float)re[3 * (size_t)i + c];
        }
        std::memcpy(nodes + 4 * (size_t)slot.nbase, b.nodes.data(), b.nodes.size() * sizeof(float));
        if (root_boxes) {
            float* rb = root_boxes + 6 * (size_t)s;
            if (b.node_count == 0) { rb[0] = rb[1] = rb[2] = INFINITY; rb[3] = rb[4] = rb[5] = -INFINITY; }
            else { rb[0] = b.nodes[0]; rb[1] = b.nodes[1]; rb[2] = b.nodes[2]; rb[3] = b.nodes[4]; rb[4] = b.nodes[5]; rb[5] = b.nodes[6]; }
        }
    }
    return PT_OK;
}

}  // extern "C"
