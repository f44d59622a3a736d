// builder.cpp — wide BVH, binary BVH and env-table builders (host, C++17).
#include "builder.h"
#include "rail.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "../../include/ptmi_math.h"
#include "sky.h"

// Collapse of the binary SAH tree into 8-wide nodes: 2 = the SAH-optimal dynamic programme
// (Ylitie, Karras, Laine 2017, section 3.2; the reference's accel/cwbvh/cwbvh.ts builds the same
// way), 1 = greedy largest-area opening (non-leaf subtrees first), 0 = greedy by area only.
#ifndef PTMI_COLLAPSE_RULE
#define PTMI_COLLAPSE_RULE 2
#endif
// Bins per axis of the binary SAH split search that feeds the collapse.
#ifndef PTMI_SAH_BINS
#define PTMI_SAH_BINS 16
#endif

namespace ptmi {

namespace {

struct Box {
    float lo[3], hi[3];
    void reset() { lo[0] = lo[1] = lo[2] = INFINITY; hi[0] = hi[1] = hi[2] = -INFINITY; }
    void grow(const Box& b) {
        for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], b.lo[a]); hi[a] = std::max(hi[a], b.hi[a]); }
    }
    void grow(const float* p) {
        for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], p[a]); hi[a] = std::max(hi[a], p[a]); }
    }
    double area() const {
        double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

// ---------------------------------------------------------------- binary SAH (leaf = 1 tri)
struct BNode {
    Box box;
    int left = -1, right = -1;   // internal
    int first = 0, count = 0;    // subtree triangle range in `order`
};

struct SahBuilder {
    const Box* tri_box;
    std::vector<float> cen;      // 3 per item
    std::vector<uint32_t> order;
    std::vector<BNode> nodes;

    int build(int s, int e) {
        int id = (int)nodes.size();
        nodes.emplace_back();
        Box b, cb; b.reset(); cb.reset();
        for (int i = s; i < e; i++) {
            b.grow(tri_box[order[i]]);
            cb.grow(&cen[3 * (size_t)order[i]]);
        }
        nodes[id].box = b;
        nodes[id].first = s;
        nodes[id].count = e - s;
        if (e - s <= 1) return id;
        constexpr int NB = PTMI_SAH_BINS;
        int best_axis = -1, best_split = -1;
        double best_cost = INFINITY;
        for (int a = 0; a < 3; a++) {
            float ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 0.0f)) continue;
            double scale = NB / (double)ext;
            Box bb[NB]; int bc[NB];
            for (int k = 0; k < NB; k++) { bb[k].reset(); bc[k] = 0; }
            for (int i = s; i < e; i++) {
                uint32_t t = order[i];
                int k = (int)((cen[3 * (size_t)t + a] - cb.lo[a]) * scale);
                k = std::min(std::max(k, 0), NB - 1);
                bb[k].grow(tri_box[t]); bc[k]++;
            }
            double la[NB - 1]; int lc[NB - 1];
            Box acc; acc.reset(); int n = 0;
            for (int k = 0; k < NB - 1; k++) { acc.grow(bb[k]); n += bc[k]; la[k] = acc.area(); lc[k] = n; }
            acc.reset(); n = 0;
            for (int k = NB - 1; k > 0; k--) {
                acc.grow(bb[k]); n += bc[k];
                if (lc[k - 1] == 0 || n == 0) continue;
                double cost = la[k - 1] * lc[k - 1] + acc.area() * n;
                if (cost < best_cost) { best_cost = cost; best_axis = a; best_split = k - 1; }
            }
        }
        int mid = -1;
        if (best_axis >= 0) {
            float lo = cb.lo[best_axis];
            double scale = NB / (double)(cb.hi[best_axis] - lo);
            auto it = std::partition(order.begin() + s, order.begin() + e, [&](uint32_t t) {
                int k = (int)((cen[3 * (size_t)t + best_axis] - lo) * scale);
                k = std::min(std::max(k, 0), NB - 1);
                return k <= best_split;
            });
            mid = (int)(it - order.begin());
            if (mid <= s || mid >= e) mid = -1;
        }
        if (mid < 0) mid = (s + e) / 2;
        int l = build(s, mid);
        int r = build(mid, e);
        nodes[id].left = l;
        nodes[id].right = r;
        return id;
    }
};

// The collapse cost model (Ylitie et al. 2017, section 3.2): a wide node visit costs C_NODE per
// unit of its box area, a leaf triangle test C_PRIM; a leaf slot holds at most 3 triangles. The
// paper's C_PRIM is 0.3; measured on the 1M-triangle icosphere cloud at 1080p, C_PRIM = 1 cuts
// the triangle tests per closest-hit query from 5.42 to 3.86 at equal node visits and k_extend
// from 121.1 to 108.8 ms per frame (0.6: 110.8 ms; 1.5 and 2.5: 108.8 ms). The reference names
// 1.0 as its second build setting (accel/cwbvh/cwbvh.ts:44-48).
constexpr double COLLAPSE_C_NODE = 1.0;
#ifndef PTMI_COLLAPSE_CPRIM
#define PTMI_COLLAPSE_CPRIM 1.0
#endif
constexpr double COLLAPSE_C_PRIM = PTMI_COLLAPSE_CPRIM;

// cost[8 n + i] (i = 1..7) is the cheapest SAH cost of representing binary subtree n as at most
// i slot subtrees of its wide parent; pick[8 n + j] is the left share k of the best split of n
// over j slots (j = 2..8), or 0 where j - 1 slots cost no more (j <= 7). Bottom-up over the
// binary nodes, which SahBuilder emits parents first.
struct CollapseDP {
    const std::vector<BNode>& bn;
    std::vector<float> cost;
    std::vector<uint8_t> pick;
    explicit CollapseDP(const std::vector<BNode>& nodes) : bn(nodes) {
        const size_t n = bn.size();
        cost.assign(8 * n, 0.0f);
        pick.assign(9 * n, 0);
        for (size_t q = n; q-- > 0;) {
            const BNode& b = bn[q];
            const double area = b.box.area();
            const double leaf = b.count <= 3 ? area * b.count * COLLAPSE_C_PRIM : INFINITY;
            float* c = &cost[8 * q];
            if (b.left < 0) { for (int i = 1; i < 8; i++) c[i] = (float)leaf; continue; }
            const float* cl = &cost[8 * (size_t)b.left];
            const float* cr = &cost[8 * (size_t)b.right];
            double dist[9];
            uint8_t* pk = &pick[9 * q];
            for (int j = 2; j <= 8; j++) {
                double best = INFINITY; int kb = 1;
                for (int k = std::max(1, j - 7); k <= std::min(7, j - 1); k++) {
                    double v = (double)cl[k] + (double)cr[j - k];
                    if (v < best) { best = v; kb = k; }
                }
                dist[j] = best;
                pk[j] = (uint8_t)kb;
            }
            c[1] = (float)std::min(leaf, dist[8] + area * COLLAPSE_C_NODE);
            for (int i = 2; i < 8; i++) {
                if (dist[i] < (double)c[i - 1]) c[i] = (float)dist[i];
                else { c[i] = c[i - 1]; pk[i] = 0; }
            }
        }
    }
    // The slot subtrees of binary node m given i slots.
    void expand(int m, int i, std::vector<int>& out) const {
        const BNode& b = bn[m];
        while (i > 1 && b.left >= 0 && pick[9 * (size_t)m + i] == 0) i--;
        if (i == 1 || b.left < 0) { out.push_back(m); return; }
        int k = pick[9 * (size_t)m + i];
        expand(b.left, k, out);
        expand(b.right, i - k, out);
    }
    // The children of the wide node made from binary node m (internal): its best 8-slot split.
    void children(int m, std::vector<int>& out) const {
        const BNode& b = bn[m];
        int k = pick[9 * (size_t)m + 8];
        expand(b.left, k, out);
        expand(b.right, 8 - k, out);
    }
};

inline float decode(float p, int q, float s) { return p + (float)q * s; }

// Quantize [clo, chi] on one axis against origin p and scale s; false if 255 is exceeded.
bool quantize_axis(float p, float s, float clo, float chi, int& qlo, int& qhi) {
    qlo = (int)std::floor((clo - p) / s);
    qhi = (int)std::ceil((chi - p) / s);
    qlo = std::max(qlo, 0);
    while (qlo > 0 && decode(p, qlo, s) > clo) qlo--;
    while (qlo < 255 && decode(p, qlo + 1, s) <= clo) qlo++;
    if (decode(p, qlo, s) > clo) return false;
    qhi = std::max(qhi, qlo);
    while (qhi <= 255 && decode(p, qhi, s) < chi) qhi++;
    while (qhi > qlo && decode(p, qhi - 1, s) >= chi) qhi--;
    return qhi <= 255;
}

}  // namespace

void build_wide_bvh(const float* tv, int n_tris, WideBVH& out) {
    out = WideBVH();
    if (n_tris <= 0) return;
    std::vector<Box> tri_box((size_t)n_tris);
    SahBuilder sb;
    sb.tri_box = tri_box.data();
    sb.cen.resize(3 * (size_t)n_tris);
    sb.order.resize((size_t)n_tris);
    for (int t = 0; t < n_tris; t++) {
        Box b; b.reset();
        for (int v = 0; v < 3; v++) b.grow(tv + 9 * (size_t)t + 3 * v);
        tri_box[t] = b;
        for (int a = 0; a < 3; a++) sb.cen[3 * (size_t)t + a] = 0.5f * (b.lo[a] + b.hi[a]);
        sb.order[t] = (uint32_t)t;
    }
    sb.nodes.reserve(2 * (size_t)n_tris);
    sb.build(0, n_tris);
    const std::vector<BNode>& bn = sb.nodes;

    // Conservative child box: the subtree box widened by 4e-6 of its largest extent, so the
    // Moller-Trumbore edge tolerance (1e-6 barycentric) cannot place a hit outside it.
    auto padded = [&](int b) {
        Box x = bn[b].box;
        float ext = std::max(x.hi[0] - x.lo[0], std::max(x.hi[1] - x.lo[1], x.hi[2] - x.lo[2]));
        float m = 4e-6f * ext + 1e-30f;
        for (int a = 0; a < 3; a++) {
            x.lo[a] = x.lo[a] - std::max(m, 4e-7f * std::fabs(x.lo[a]));
            x.hi[a] = x.hi[a] + std::max(m, 4e-7f * std::fabs(x.hi[a]));
        }
        return x;
    };

#if PTMI_COLLAPSE_RULE == 2
    const CollapseDP dp(bn);
#endif
    struct Job { int bnode; uint32_t wide; uint32_t depth; };
    std::vector<Job> stack;
    out.nodes.assign(20, 0u);
    out.n_nodes = 1;
    stack.push_back({0, 0, 1});
    out.tris.reserve(12 * (size_t)n_tris);
    while (!stack.empty()) {
        Job job = stack.back(); stack.pop_back();
        out.max_depth = std::max(out.max_depth, job.depth);
        std::vector<int> ch;
#if PTMI_COLLAPSE_RULE == 2
        if (bn[job.bnode].left < 0) ch.push_back(job.bnode);
        else dp.children(job.bnode, ch);
#else
        // collect up to 8 children by greedy largest-area expansion
        if (bn[job.bnode].left < 0) ch.push_back(job.bnode);
        else { ch.push_back(bn[job.bnode].left); ch.push_back(bn[job.bnode].right); }
        // Fill the 8 slots: open the largest subtree that cannot be a leaf (> 3 triangles)
        // first, so big subtrees do not cost a node visit; only then spend spare slots on
        // splitting small leaves (tighter boxes, fewer triangle tests; slot tests are paid
        // per node visit anyway).
        while (ch.size() < 8) {
            int best = -1; double ba = -1.0;
            for (int pass = 0; pass < 2 && best < 0; pass++) {
                for (size_t i = 0; i < ch.size(); i++) {
                    const BNode& c = bn[ch[i]];
                    if (c.left < 0) continue;
#if PTMI_COLLAPSE_RULE == 0
                    if (pass == 1) continue;
#else
                    if ((pass == 0) != (c.count > 3)) continue;
#endif
                    double a = c.box.area();
                    if (a > ba) { ba = a; best = (int)i; }
                }
            }
            if (best < 0) break;
            int b = ch[best];
            ch[best] = bn[b].left;
            ch.push_back(bn[b].right);
        }
#endif
        // slot assignment by centroid octant
        Box nb; nb.reset();
        std::vector<Box> cbox(ch.size());
        for (size_t i = 0; i < ch.size(); i++) { cbox[i] = padded(ch[i]); nb.grow(cbox[i]); }
        float pc[3];
        for (int a = 0; a < 3; a++) pc[a] = 0.5f * (nb.lo[a] + nb.hi[a]);
        struct Cand { float score; int c, s; };
        std::vector<Cand> cand;
        for (size_t i = 0; i < ch.size(); i++)
            for (int s = 0; s < 8; s++) {
                float sc = 0.0f;
                for (int a = 0; a < 3; a++) {
                    float d = 0.5f * (cbox[i].lo[a] + cbox[i].hi[a]) - pc[a];
                    sc += ((s >> a) & 1) ? -d : d;
                }
                cand.push_back({sc, (int)i, s});
            }
        std::stable_sort(cand.begin(), cand.end(), [](const Cand& x, const Cand& y) { return x.score > y.score; });
        int slot_of[8]; int child_in_slot[8];
        for (int i = 0; i < 8; i++) { slot_of[i] = -1; child_in_slot[i] = -1; }
        for (const Cand& c : cand) {
            if (slot_of[c.c] >= 0 || child_in_slot[c.s] >= 0) continue;
            slot_of[c.c] = c.s; child_in_slot[c.s] = c.c;
        }
        // quantization grid
        float p[3]; int e[3]; float sc[3];
        for (int a = 0; a < 3; a++) {
            p[a] = nb.lo[a];
            float ext = nb.hi[a] - nb.lo[a];
            int ex = -126;
            if (ext > 0.0f) {
                int k; std::frexp(ext / 255.0f, &k);   // ext/255 in [2^(k-1), 2^k)
                ex = std::max(k, -126);
            }
            e[a] = ex;
        }
        uint32_t* W = &out.nodes[20 * (size_t)job.wide];
        uint8_t meta[8] = {0}, qlo[3][8] = {{0}}, qhi[3][8] = {{0}};
        uint32_t imask = 0;
        // internal children first (contiguous at child_base in slot order), then leaf triangles
        uint32_t child_base = out.n_nodes, tri_base = (uint32_t)(out.tris.size() / 12);
        int n_internal = 0, n_leaf_tris = 0;
        for (int s = 0; s < 8; s++) {
            int i = child_in_slot[s];
            if (i < 0) continue;
            const BNode& c = bn[ch[i]];
            const Box& b = cbox[i];
            for (int a = 0; a < 3; a++) {
                int lo, hi;
                for (;;) {
                    sc[a] = ptm_u2f((uint32_t)(e[a] + 127) << 23);
                    if (quantize_axis(p[a], sc[a], b.lo[a], b.hi[a], lo, hi)) break;
                    e[a]++;
                    // restart all earlier slots of this axis at the new scale
                    for (int s2 = 0; s2 < s; s2++) {
                        int i2 = child_in_slot[s2];
                        if (i2 < 0) continue;
                        int l2, h2;
                        sc[a] = ptm_u2f((uint32_t)(e[a] + 127) << 23);
                        if (!quantize_axis(p[a], sc[a], cbox[i2].lo[a], cbox[i2].hi[a], l2, h2)) throw std::runtime_error("wide BVH quantization failed");
                        qlo[a][s2] = (uint8_t)l2; qhi[a][s2] = (uint8_t)h2;
                    }
                }
                qlo[a][s] = (uint8_t)lo; qhi[a][s] = (uint8_t)hi;
            }
            if (c.count > 3) {
                imask |= 1u << s;
                meta[s] = (uint8_t)(0x20 | s);
                n_internal++;
            } else {
                static const uint8_t unary[4] = {0, 0x20, 0x60, 0xE0};
                meta[s] = (uint8_t)(unary[c.count] | n_leaf_tris);
                n_leaf_tris += c.count;
            }
        }
        // reserve internal children and emit leaf triangles in slot order
        out.nodes.resize(20 * (size_t)(out.n_nodes + n_internal), 0u);
        W = &out.nodes[20 * (size_t)job.wide];
        uint32_t next = child_base;
        for (int s = 0; s < 8; s++) {
            int i = child_in_slot[s];
            if (i < 0) continue;
            const BNode& c = bn[ch[i]];
            if (imask & (1u << s)) {
                stack.push_back({ch[i], next++, job.depth + 1});
            } else {
                for (int k = 0; k < c.count; k++) {
                    uint32_t t = sb.order[c.first + k];
                    const float* v = tv + 9 * (size_t)t;
                    float rec[12] = {v[0], v[1], v[2], ptm_u2f(t),
                                     v[3] - v[0], v[4] - v[1], v[5] - v[2], 0.0f,
                                     v[6] - v[0], v[7] - v[1], v[8] - v[2], 0.0f};
                    out.tris.insert(out.tris.end(), rec, rec + 12);
                }
            }
        }
        out.n_nodes += (uint32_t)n_internal;
        W[0] = ptm_f2u(p[0]); W[1] = ptm_f2u(p[1]); W[2] = ptm_f2u(p[2]);
        W[3] = (uint32_t)(e[0] + 127) | ((uint32_t)(e[1] + 127) << 8) | ((uint32_t)(e[2] + 127) << 16) | (imask << 24);
        W[4] = child_base; W[5] = tri_base;
        auto pack4 = [](const uint8_t* b) { return (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24); };
        W[6] = pack4(meta); W[7] = pack4(meta + 4);
        W[8] = pack4(qlo[0]); W[9] = pack4(qlo[0] + 4); W[10] = pack4(qlo[1]); W[11] = pack4(qlo[1] + 4);
        W[12] = pack4(qlo[2]); W[13] = pack4(qlo[2] + 4); W[14] = pack4(qhi[0]); W[15] = pack4(qhi[0] + 4);
        W[16] = pack4(qhi[1]); W[17] = pack4(qhi[1] + 4); W[18] = pack4(qhi[2]); W[19] = pack4(qhi[2] + 4);
    }
}

// ---------------------------------------------------------------- binary BVH (containment)
// The reference's own builder (buildBVHNodesFlat, median fallback; rail.cpp), byte for byte.
void build_binary_bvh(const double* boxes, int n, int leaf_size, BinaryBVH& out) {
    FlatBVH b;
    build_bvh_flat(boxes, n, leaf_size, false, false, b);
    out.nodes = std::move(b.nodes);
    out.order = std::move(b.order);
}

// ---------------------------------------------------------------- env importance table
namespace {
struct F3 { float x, y, z; };
}  // namespace

void build_env_table(const pt_env& env, std::vector<float>& cond, std::vector<float>& marg, double& total) {
    const bool oct = env.chart == PT_CHART_OCTAHEDRAL;          // N x N, N = table_h (EnvironmentBake.ts:20-27)
    const int H = env.table_h > 0 ? env.table_h : 256, W = oct ? H : (env.table_w > 0 ? env.table_w : 512);
    std::vector<float> Y((size_t)W * H);
    for (int j = 0; j < H; j++) {
        for (int i = 0; i < W; i++) {
            float u = ((float)i + 0.5f) / (float)W, v = ((float)j + 0.5f) / (float)H;
            // env_chart_dir with the bake's rotation 0 (EnvironmentBake.ts: const u_envRotation / u_envRotCS)
            F3 d;
            if (oct) ptm_oct_dir(u, v, 1.0f, 0.0f, &d.x, &d.y, &d.z);
            else {
                float phi = (u - 0.5f) * PTM_TWO_PI - 0.0f;
                float theta = v * PTM_PI;
                float st = ptm_sin(theta);
                d = {ptm_cos(phi) * st, ptm_cos(theta), ptm_sin(phi) * st};
            }
            float c[3];
            pt_sky_body(env.body, d.x, d.y, d.z, c);
            Y[(size_t)j * W + i] = (float)(0.2126 * (double)c[0] + 0.7152 * (double)c[1] + 0.0722 * (double)c[2]);
        }
    }
    env_table_from_luma(Y, W, H, oct, env.compensated != 0, cond, marg, total);
}

// The importance chain of components/env/importance.ts:44-127 over a W x H luminance table:
// 3 x 3 tent blur (u wraps on the equirect chart), MIS compensation, row and marginal CDFs.
void env_table_from_luma(const std::vector<float>& Y, int W, int H, bool oct, bool compensated, std::vector<float>& cond,
                         std::vector<float>& marg, double& total) {
    std::vector<float> Bl((size_t)W * H);
    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++) {
            double sum = 0, wsum = 0;
            for (int dj = -1; dj <= 1; dj++) {
                int jj = std::min(H - 1, std::max(0, j + dj));
                for (int di = -1; di <= 1; di++) {
                    int ii = oct ? std::min(W - 1, std::max(0, i + di)) : (i + di + W) % W;   // u wraps on equirect
                    double w = (dj == 0 ? 2 : 1) * (di == 0 ? 2 : 1);
                    sum += w * (double)Y[(size_t)jj * W + ii];
                    wsum += w;
                }
            }
            Bl[(size_t)j * W + i] = (float)(sum / wsum);
        }
    if (compensated) {
        double num = 0, den = 0;
        for (int j = 0; j < H; j++) {
            double dO = oct ? (4 * M_PI) / (W * H) : (2 * M_PI / W) * (M_PI / H) * std::sin(M_PI * (j + 0.5) / H);
            for (int i = 0; i < W; i++) { num += (double)Bl[(size_t)j * W + i] * dO; den += dO; }
        }
        double mean = den > 0 ? num / den : 0;
        for (float& y : Bl) { double t = (double)y - mean; y = (float)(t > 0 ? t : 0); }
    }
    cond.assign((size_t)W * H, 0.0f);
    marg.assign((size_t)H, 0.0f);
    std::vector<double> row((size_t)H);
    total = 0.0;
    for (int j = 0; j < H; j++) {
        double sinT = oct ? 1.0 : std::sin(M_PI * (j + 0.5) / H);   // octahedral texels are equal-area
        double sum = 0;
        for (int i = 0; i < W; i++) sum += std::max(0.0, (double)Bl[(size_t)j * W + i]) * sinT;
        row[j] = sum;
        double c = 0;
        for (int i = 0; i < W; i++) {
            c += std::max(0.0, (double)Bl[(size_t)j * W + i]) * sinT;
            cond[(size_t)j * W + i] = (float)(sum > 0 ? c / sum : 0);
        }
        total += sum;
    }
    double acc = 0;
    for (int j = 0; j < H; j++) { acc += row[j]; marg[j] = (float)(total > 0 ? acc / total : 0); }
}

// equalAreaSquareToSphere (env/octahedral/octahedral.ts:10-26), f64
static void eq_area_dir(double u, double v, double* d) {
    const double up = 2 * u - 1, vp = 2 * v - 1;
    const double ua = std::fabs(up), va = std::fabs(vp);
    const double sd = 1 - (ua + va);
    const double r = 1 - std::fabs(sd);
    const double phi = ((r == 0 ? 1 : (va - ua) / r) + 1) * M_PI / 4;
    const double y = std::copysign(1 - r * r, sd);
    const double cp = std::copysign(std::cos(phi), up), sp = std::copysign(std::sin(phi), vp);
    const double s = r * std::sqrt(std::max(0.0, 2 - r * r));
    d[0] = cp * s; d[1] = y; d[2] = sp * s;
}

void resample_equirect_to_octahedral(const float* rgb, int W, int H, int N, std::vector<float>& out) {
    out.assign((size_t)3 * N * N, 0.0f);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
            double d[3];
            eq_area_dir((i + 0.5) / N, (j + 0.5) / N, d);
            const double phi = std::atan2(d[2], d[0]);
            const double theta = std::acos(std::min(1.0, std::max(-1.0, d[1])));
            const double u = phi / (2 * M_PI) + 0.5, v = theta / M_PI;
            const double fx = u * W - 0.5, fy = v * H - 0.5;
            const double x0 = std::floor(fx), y0 = std::floor(fy);
            const double tx = fx - x0, ty = fy - y0;
            auto wrap = [&](double x) { return (int)std::fmod(std::fmod(x, (double)W) + W, (double)W); };
            auto cy = [&](double y) { return (int)std::min((double)(H - 1), std::max(0.0, y)); };
            for (int c = 0; c < 3; c++) {
                auto at = [&](double x, double y) { return (double)rgb[3 * ((size_t)cy(y) * W + wrap(x)) + c]; };
                out[3 * ((size_t)j * N + i) + c] = (float)((1 - tx) * (1 - ty) * at(x0, y0) + tx * (1 - ty) * at(x0 + 1, y0) +
                                                           (1 - tx) * ty * at(x0, y0 + 1) + tx * ty * at(x0 + 1, y0 + 1));
            }
        }
}

void build_env_table_image(const pt_env& env, const float* rgb, int w, int h, std::vector<float>& cond,
                           std::vector<float>& marg, double& total) {
    const bool oct = env.chart == PT_CHART_OCTAHEDRAL;
    std::vector<float> resampled;
    int W = w, H = h;
    if (oct) {
        W = H = env.table_h > 0 ? env.table_h : 256;
        resample_equirect_to_octahedral(rgb, w, h, W, resampled);
        rgb = resampled.data();
    }
    std::vector<float> Y((size_t)W * H);
    for (size_t k = 0; k < Y.size(); k++)
        Y[k] = (float)(0.2126 * (double)rgb[3 * k] + 0.7152 * (double)rgb[3 * k + 1] + 0.0722 * (double)rgb[3 * k + 2]);
    env_table_from_luma(Y, W, H, oct, env.compensated != 0, cond, marg, total);
}

// ---------------------------------------------------------------- containment probe map
void probe_basis(float basis[6]) {
    const double d[3] = {PROBE_DIR[0], PROBE_DIR[1], PROBE_DIR[2]};
    const double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double n[3] = {d[0] / dl, d[1] / dl, d[2] / dl};
    // a = normalize(n x z), b = n x a: both across the direction (its z component is the smallest)
    double a[3] = {n[1], -n[0], 0.0};
    const double al = std::sqrt(a[0] * a[0] + a[1] * a[1]);
    a[0] /= al; a[1] /= al;
    const double b[3] = {n[1] * a[2] - n[2] * a[1], n[2] * a[0] - n[0] * a[2], n[0] * a[1] - n[1] * a[0]};
    for (int k = 0; k < 3; k++) { basis[k] = (float)a[k]; basis[3 + k] = (float)b[k]; }
}

// fp32 neighbours of x on the outside: the largest float <= x, the smallest float >= x.
static float round_down(double x) { float f = (float)x; return (double)f > x ? std::nextafter(f, -INFINITY) : f; }
static float round_up(double x) { float f = (float)x; return (double)f < x ? std::nextafter(f, INFINITY) : f; }

void build_probe_map(const float* tris12, int n, int first, ProbeMap& m) {
    probe_basis(m.basis);
    const double A[3] = {m.basis[0], m.basis[1], m.basis[2]}, B[3] = {m.basis[3], m.basis[4], m.basis[5]};
    // the triangles' projected boxes and the object's 3D box (f64; b = a + e1, c = a + e2)
    std::vector<double> fb(4 * (size_t)n);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < n; k++) {
        const float* t = tris12 + 12 * (size_t)k;
        double v[3][3];
        for (int a = 0; a < 3; a++) {
            v[0][a] = t[a];
            v[1][a] = (double)t[a] + (double)t[4 + a];
            v[2][a] = (double)t[a] + (double)t[8 + a];
        }
        double umin = INFINITY, umax = -INFINITY, vmin = INFINITY, vmax = -INFINITY;
        for (int c = 0; c < 3; c++) {
            const double u = A[0] * v[c][0] + A[1] * v[c][1] + A[2] * v[c][2];
            const double w = B[0] * v[c][0] + B[1] * v[c][1] + B[2] * v[c][2];
            umin = std::min(umin, u); umax = std::max(umax, u);
            vmin = std::min(vmin, w); vmax = std::max(vmax, w);
            for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], v[c][a]); hi[a] = std::max(hi[a], v[c][a]); }
        }
        fb[4 * k] = umin; fb[4 * k + 1] = umax; fb[4 * k + 2] = vmin; fb[4 * k + 3] = vmax;
    }
    // Margin: the Moller-Trumbore test accepts rays that miss the exact triangle by its 1e-6
    // barycentric tolerance plus rounding (both below 1e-6 of the distance from the ray origin
    // to the triangle, the origin lying in the object's box), and the device's fp32 projection
    // and cell index err by under 1e-6 of the point's magnitude: 1e-4 of the box diagonal plus
    // 1e-5 of the box's largest corner norm covers both a hundredfold.
    double diag = 0.0, rmax = 0.0;
    for (int a = 0; a < 3; a++) diag += (hi[a] - lo[a]) * (hi[a] - lo[a]);
    diag = std::sqrt(diag);
    for (int c = 0; c < 8; c++) {
        double r = 0.0;
        for (int a = 0; a < 3; a++) { const double x = (c >> a) & 1 ? hi[a] : lo[a]; r += x * x; }
        rmax = std::max(rmax, std::sqrt(r));
    }
    const double margin = 1e-4 * diag + 1e-5 * rmax + 1e-30;
    double U0 = INFINITY, U1 = -INFINITY, V0 = INFINITY, V1 = -INFINITY;
    for (int k = 0; k < n; k++) {
        U0 = std::min(U0, fb[4 * k] - margin); U1 = std::max(U1, fb[4 * k + 1] + margin);
        V0 = std::min(V0, fb[4 * k + 2] - margin); V1 = std::max(V1, fb[4 * k + 3] + margin);
    }
    // about one cell per triangle (two per triangle test fewer candidates in the same time, with
    // 60 % more items), square cells, at most 2048 a side
    const double W = U1 - U0, H = V1 - V0;
    const double cell = std::sqrt(std::max(W * H, 1e-60) / std::max(1.0 * n, 1.0));
    m.gu = (int)std::min(2048.0, std::max(1.0, std::ceil(W / cell)));
    m.gv = (int)std::min(2048.0, std::max(1.0, std::ceil(H / cell)));
    // the device indexes with fp32: keep the grid's origin and scale as the fp32 values it uses
    m.u0 = (float)U0; m.v0 = (float)V0;
    m.inv_u = (float)(m.gu / std::max(W, 1e-30));
    m.inv_v = (float)(m.gv / std::max(H, 1e-30));
    auto cell_u = [&](double x) { return (int)std::min<double>(m.gu - 1, std::max(0.0, std::floor((x - (double)m.u0) * (double)m.inv_u))); };
    auto cell_v = [&](double x) { return (int)std::min<double>(m.gv - 1, std::max(0.0, std::floor((x - (double)m.v0) * (double)m.inv_v))); };
    std::vector<int> count((size_t)m.gu * m.gv + 1, 0);
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) {
            m.cell_start.assign((size_t)m.gu * m.gv + 1, 0);
            for (size_t c = 0; c < (size_t)m.gu * m.gv; c++) m.cell_start[c + 1] = m.cell_start[c] + count[c];
            m.items.assign((size_t)m.cell_start.back(), 0);
            m.boxes.assign(4 * (size_t)m.cell_start.back(), 0.0f);
            std::fill(count.begin(), count.end(), 0);
        }
        for (int k = 0; k < n; k++) {
            const int cu0 = cell_u(fb[4 * k] - margin), cu1 = cell_u(fb[4 * k + 1] + margin);
            const int cv0 = cell_v(fb[4 * k + 2] - margin), cv1 = cell_v(fb[4 * k + 3] + margin);
            for (int cv = cv0; cv <= cv1; cv++)
                for (int cu = cu0; cu <= cu1; cu++) {
                    const size_t c = (size_t)cv * m.gu + cu;
                    if (pass == 1) {
                        const size_t it = (size_t)m.cell_start[c] + count[c];
                        m.items[it] = first + k;
                        m.boxes[4 * it] = round_down(fb[4 * k] - margin);
                        m.boxes[4 * it + 1] = round_up(fb[4 * k + 1] + margin);
                        m.boxes[4 * it + 2] = round_down(fb[4 * k + 2] - margin);
                        m.boxes[4 * it + 3] = round_up(fb[4 * k + 3] + margin);
                    }
                    count[c]++;
                }
        }
    }
}

}  // namespace ptmi
