// traversal.h — closest-hit / any-hit queries over the wide BVH + analytic primitives.
//
// One ray per lane. A node visit loads the 80-byte node as five 16-byte loads, tests the
// (up to) eight quantized child boxes against the ray and yields two bit groups: internal
// children in front-to-back octant order (bits 0..7 of `hits`, position = slot ^ (oct ^ 7))
// and leaf triangles (24-bit mask relative to tri_base). Pending node groups live on a
// per-lane stack in LDS (bank-conflict free: entry k of lane t at k*BLOCK + t).
//
// Results equal the reference traversal (intersection/mesh/mesh.glsl:41-191 walking
// accel/bvh/bvh.glsl:28-36 boxes) bit for bit:
//   * the triangle test is Moller-Trumbore in the reference's operation order, with
//     e1 = b - a, e2 = c - a precomputed in fp32 (the same roundings the shader does);
//   * child boxes are decoded with the fp32 ops the builder verified against, and they
//     contain the exact triangle boxes, so the slab test (the reference's form, far plane
//     padded by BVH_TFAR_PAD) never rejects a box the exact-box walk would enter;
//   * bit-equal t ties resolve to the smaller global primitive id (order independent).
#pragma once
#include "dev_common.h"

namespace ptmi {

// Moller-Trumbore (mesh.glsl:41-59). Returns true with t when u, v, w >= -1e-6.
PT_DEV bool tri_test(f3 ro, f3 rd, f3 a, f3 e1, f3 e2, float& t_out) {
    f3 g = cross(e1, e2);
    float det = -dot(rd, g);
    if (fabsf(det) < 1e-12f) return false;
    float invdet = 1.0f / det;
    f3 ao = ro - a;
    f3 dao = cross(ao, rd);
    float u = dot(e2, dao) * invdet;
    float v = -dot(e1, dao) * invdet;
    float t = dot(ao, g) * invdet;
    float w = 1.0f - u - v;
    t_out = t;
    const float E = 1e-6f;
    return u >= -E && v >= -E && w >= -E;
}

// Analytic sphere / quad (geometry/sphere/sphere.glsl:28-46, geometry/quad/quad.glsl:9-22).
PT_DEV bool sphere_intersect(f3 ro, f3 rd, f3 c, float r, float& t) {
    f3 oc = ro - c;
    float b = dot(oc, rd);
    f3 perp = oc - rd * b;
    float disc = r * r - dot(perp, perp);
    if (disc < 0.0f) return false;
    float cc = dot(oc, oc) - r * r;
    float q = (b >= 0.0f) ? -b - sqrtf(disc) : -b + sqrtf(disc);
    if (q == 0.0f) return false;
    float ta = q, tb = cc / q;
    t = (cc < 0.0f) ? gmax(ta, tb) : gmin(ta, tb);
    return t > 0.0f;
}
PT_DEV bool quad_intersect(f3 ro, f3 rd, const float* p, float& t) {
    f3 corner = ld3(p), e1 = ld3(p + 3), e2 = ld3(p + 6), n = ld3(p + 9);
    float denom = dot(rd, n);
    if (fabsf(denom) < 1e-8f) return false;
    t = dot(corner - ro, n) / denom;
    if (t <= 0.0f) return false;
    f3 local = (ro + rd * t) - corner;
    float e11 = dot(e1, e1), e22 = dot(e2, e2), e12 = dot(e1, e2);
    float d1 = dot(local, e1), d2 = dot(local, e2);
    float det = e11 * e22 - e12 * e12;
    float u = (d1 * e22 - d2 * e12) / det;
    float v = (d2 * e11 - d1 * e12) / det;
    return u >= 0.0f && u <= 1.0f && v >= 0.0f && v <= 1.0f;
}

struct TraceCounters { uint32_t nodes, tris; };

#ifndef PTMI_SLAB_IEEE_MINMAX
#define PTMI_SLAB_IEEE_MINMAX 1
#endif
#if PTMI_SLAB_IEEE_MINMAX
#define SMIN(a, b) __builtin_fminf(a, b)
#define SMAX(a, b) __builtin_fmaxf(a, b)
#else
#define SMIN(a, b) gmin(a, b)
#define SMAX(a, b) gmax(a, b)
#endif

PT_DEV float byte_of(uint32_t w, int k) { return (float)((w >> (8 * k)) & 0xFFu); }

// The slab bound: 6 u relative to |oq| + 255 |dq| (u = 2^-24), and an absolute floor for the
// subnormal range of dq.
constexpr float SLAB_REL = 6.0f * 5.9604644775390625e-8f;
constexpr float SLAB_ABS = 1e-30f;


// Per-lane traversal state of the wide-BVH walk, so a persistent kernel can interleave
// rays: a lane that finishes picks up the next ray while its wave-mates keep walking.
struct TravState {
    f3 o, d;
    float ix, iy, iz, tmax;
    uint32_t oct_inv, g_base, g_bits;
    uint32_t t_base, t_bits;    // pending leaf-triangle group (24-bit mask relative to t_base)
    int prim, sp;
};

PT_DEV void trace_init(TravState& st, f3 o, f3 d, float tmax, uint32_t root = 0) {
    st.o = o; st.d = d; st.tmax = tmax; st.prim = -1; st.sp = 0;
    // inverse direction with the reference CWBVH's zero guard (|d| < 1e-20 -> +-1e-20)
    float dx = fabsf(d.x) < 1e-20f ? copysignf(1e-20f, d.x) : d.x;
    float dy = fabsf(d.y) < 1e-20f ? copysignf(1e-20f, d.y) : d.y;
    float dz = fabsf(d.z) < 1e-20f ? copysignf(1e-20f, d.z) : d.z;
    st.ix = 1.0f / dx; st.iy = 1.0f / dy; st.iz = 1.0f / dz;
    uint32_t oct = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
    st.oct_inv = oct ^ 7u;
    st.g_base = root;
    st.g_bits = (1u << 8) | (1u << (0u ^ st.oct_inv));   // root = slot 0 of a virtual parent
    st.t_base = 0; st.t_bits = 0;
}

// The node at the head of the pending group (the nearest child not yet visited).
PT_DEV uint32_t head_node(const TravState& st) {
    const uint32_t k = __builtin_ctz(st.g_bits & 0xFFu);
    const uint32_t slot = k ^ st.oct_inv;
    const uint32_t imask = st.g_bits >> 8;
    return st.g_base + __popc(imask & ((1u << slot) - 1u));
}

// One step of a lane's walk: with no triangle pending, visit the nearest pending node; then
// test one pending triangle. At most one triangle per lane per step keeps a wave's lanes in
// step: a node that exposes several triangles does not hold the wave in a triangle loop while
// the other lanes wait (a wave's walk costs the sum of its steps, each step the slowest lane's
// work). The closest hit is the (t, id) minimum over every triangle whose box the ray enters,
// so the order of tests does not change it.
// Returns true when the ray is finished: nothing is pending, or (ANY) a hit with t < tmax was
// found, in which case st.prim is set to its id.
template <bool ANY, int BLOCK, bool COUNT>
PT_DEV bool trace_step(const uint4* __restrict__ nodes, const float4* __restrict__ tris, TravState& st, uint2* stk,
                       TraceCounters& cnt) {
    const int lane_slot = threadIdx.x;
    if (st.t_bits == 0u && (st.g_bits & 0xFFu) != 0u) {
        const uint4* np = nodes + 5 * (size_t)head_node(st);
        const uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3], w4 = np[4];
        st.g_bits &= ~(1u << __builtin_ctz(st.g_bits & 0xFFu));
        if ((st.g_bits & 0xFFu) != 0u) {
            stk[st.sp * BLOCK + lane_slot] = make_uint2(st.g_base, st.g_bits);
            st.sp++;
        }
        if (COUNT) cnt.nodes++;
        const float px = __uint_as_float(w0.x), py = __uint_as_float(w0.y), pz = __uint_as_float(w0.z);
        const float sx = __uint_as_float((w0.w & 0xFFu) << 23);
        const float sy = __uint_as_float(((w0.w >> 8) & 0xFFu) << 23);
        const float sz = __uint_as_float(((w0.w >> 16) & 0xFFu) << 23);
        const uint32_t nimask = w0.w >> 24;
        const f3 o = st.o;
        // The ray in the node's quantized frame, once per node: a child plane q (0..255) lies at
        // t = q * dq + oq with dq = 2^e * inv (exact: a power-of-two scale) and oq = (p - o) * inv,
        // one fused multiply-add per plane. Against the exact slab distance (p + q 2^e - o) / d
        // these roundings err by at most 5 u (|oq| + 255 |dq|) (u = 2^-24: the subtraction, the
        // two products, the rounded reciprocal, the final rounding); near planes move down and far
        // planes up by slab_err >= 6 u (|oq| + 255 |dq|), so a child is never culled where the
        // exact slab test of its quantized box would keep it. The quantized boxes contain the
        // triangles' boxes padded by the Moller-Trumbore tolerance, so every walk returns the
        // brute-force closest hit; the margin also covers the hit distance's own rounding
        // against the current tmax.
        const float dqx = sx * st.ix, dqy = sy * st.iy, dqz = sz * st.iz;
        const float oqx = (px - o.x) * st.ix, oqy = (py - o.y) * st.iy, oqz = (pz - o.z) * st.iz;
        const float ex = __builtin_fmaf(255.0f, fabsf(dqx), fabsf(oqx)) * SLAB_REL + SLAB_ABS;
        const float ey = __builtin_fmaf(255.0f, fabsf(dqy), fabsf(oqy)) * SLAB_REL + SLAB_ABS;
        const float ez = __builtin_fmaf(255.0f, fabsf(dqz), fabsf(oqz)) * SLAB_REL + SLAB_ABS;
        const float onx = oqx - ex, ofx = oqx + ex, ony = oqy - ey, ofy = oqy + ey, onz = oqz - ez, ofz = oqz + ez;
        // Near / far planes chosen once per node by the ray's octant: for inv > 0 the near plane of
        // an axis is the low one (the high one for inv < 0).
        const bool nxs = st.ix < 0.0f, nys = st.iy < 0.0f, nzs = st.iz < 0.0f;
        const uint32_t nx0 = nxs ? w3.z : w2.x, nx1 = nxs ? w3.w : w2.y, fx0 = nxs ? w2.x : w3.z, fx1 = nxs ? w2.y : w3.w;
        const uint32_t ny0 = nys ? w4.x : w2.z, ny1 = nys ? w4.y : w2.w, fy0 = nys ? w2.z : w4.x, fy1 = nys ? w2.w : w4.y;
        const uint32_t nz0 = nzs ? w4.z : w3.x, nz1 = nzs ? w4.w : w3.y, fz0 = nzs ? w3.x : w4.z, fz1 = nzs ? w3.y : w4.w;
        // The children whose boxes the ray enters before tmax, as a slot mask. An empty slot needs
        // no test of its own: its imask bit is 0 and its meta byte 0 (no triangles), so it adds
        // nothing below.
        uint32_t hit = 0, tri_hits = 0;
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int b = s & 3;
            const float tnx = __builtin_fmaf(byte_of(s < 4 ? nx0 : nx1, b), dqx, onx);
            const float tny = __builtin_fmaf(byte_of(s < 4 ? ny0 : ny1, b), dqy, ony);
            const float tnz = __builtin_fmaf(byte_of(s < 4 ? nz0 : nz1, b), dqz, onz);
            const float tfx = __builtin_fmaf(byte_of(s < 4 ? fx0 : fx1, b), dqx, ofx);
            const float tfy = __builtin_fmaf(byte_of(s < 4 ? fy0 : fy1, b), dqy, ofy);
            const float tfz = __builtin_fmaf(byte_of(s < 4 ? fz0 : fz1, b), dqz, ofz);
            const float te = SMAX(SMAX(SMAX(tnx, tny), tnz), 0.0f);
            const float tf = SMIN(SMIN(SMIN(tfx, tfy), tfz), st.tmax);
            hit |= te <= tf ? (1u << s) : 0u;
        }
        // leaf children: their triangles as bits relative to the node's triangle base
        const uint32_t leaf = hit & ~nimask;
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const uint32_t meta = ((s < 4 ? w1.z : w1.w) >> (8 * (s & 3))) & 0xFFu;
            tri_hits |= ((leaf >> s) & 1u) ? ((meta >> 5) & 7u) << (meta & 31u) : 0u;
        }
        // internal children, reordered front to back: slot s moves to bit s ^ (octant ^ 7)
        uint32_t inner = hit & nimask;
        const uint32_t c = st.oct_inv;
        if (c & 1u) inner = ((inner & 0x55u) << 1) | ((inner >> 1) & 0x55u);
        if (c & 2u) inner = ((inner & 0x33u) << 2) | ((inner >> 2) & 0x33u);
        if (c & 4u) inner = ((inner & 0x0Fu) << 4) | ((inner >> 4) & 0x0Fu);
        const uint32_t nhits = inner;
        (void)ANY;
        st.g_base = w1.x;
        st.g_bits = (nimask << 8) | nhits;
        st.t_base = w1.y;
        st.t_bits = tri_hits;
    }
    // this step's triangle: its loads go out first, then the next group is settled (a pop when
    // nothing else is pending), so the stack read overlaps the triangle fetch
    const bool tri = st.t_bits != 0u;
    uint32_t tk = 0;
    if (tri) { tk = __builtin_ctz(st.t_bits); st.t_bits &= st.t_bits - 1u; }
    float4 ta, te1, te2;
    if (tri) {
        const float4* tp = tris + 3 * (size_t)(st.t_base + tk);
        ta = tp[0]; te1 = tp[1]; te2 = tp[2];
    }
    bool finished = false;
    if ((st.g_bits & 0xFFu) == 0u && st.t_bits == 0u) {
        if (st.sp == 0) {
            finished = true;
        } else {
            st.sp--;
            const uint2 e = stk[st.sp * BLOCK + lane_slot];
            st.g_base = e.x; st.g_bits = e.y;
        }
    }
    if (tri) {
        if (COUNT) cnt.tris++;
        float t;
        if (tri_test(st.o, st.d, xyz(ta), xyz(te1), xyz(te2), t) && t > 0.0f) {
            const int id = (int)__float_as_uint(ta.w);
            if (ANY) {
                if (t < st.tmax) { st.prim = id; return true; }
            } else if (t < st.tmax || (t == st.tmax && id < st.prim)) {
                st.tmax = t; st.prim = id;
            }
        }
    }
    return finished;
}

// Whole-ray wide-BVH walk (one ray per lane, no interleaving).
template <bool ANY, int STACK, int BLOCK, bool COUNT>
PT_DEV bool trace_wide(const DevScene& S, f3 o, f3 d, float& tmax, int& prim, uint2* stk, TraceCounters& cnt) {
    if (S.n_cw_nodes == 0) return false;
    TravState st;
    trace_init(st, o, d, tmax);
    st.prim = prim;
    while (!trace_step<ANY, BLOCK, COUNT>(S.cw_nodes, S.cw_tris, st, stk, cnt)) {}
    (void)STACK;
    if (ANY) return st.prim >= 0 && st.prim != prim;
    bool found = st.prim != prim;
    tmax = st.tmax; prim = st.prim;
    return found;
}

// Placements (glsl/core/placement.glsl:19-41): rotate by a unit quaternion.
PT_DEV f3 qrot(float4 q, f3 v) {
    const f3 qv = mk3(q.x, q.y, q.z);
    const f3 c = cross(qv, v) + v * q.w;
    return v + cross(qv, c) * 2.0f;
}
// bvh_aabb_hit (accel/bvh/bvh.glsl:28-36) on a binary node (min, A) (max, B).
PT_DEV bool bin_aabb_hit(float4 n0, float4 n1, f3 o, f3 inv, float tmax) {
    const float t0x = (n0.x - o.x) * inv.x, t1x = (n1.x - o.x) * inv.x;
    const float t0y = (n0.y - o.y) * inv.y, t1y = (n1.y - o.y) * inv.y;
    const float t0z = (n0.z - o.z) * inv.z, t1z = (n1.z - o.z) * inv.z;
    const float tn = gmax(gmax(gmin(t0x, t1x), gmin(t0y, t1y)), gmin(t0z, t1z));
    const float tf = gmin(gmin(gmax(t0x, t1x), gmax(t0y, t1y)), gmax(t0z, t1z)) * BVH_TFAR_PAD;
    const float te = gmax(tn, 0.0f);
    return tf >= te && te <= tmax;
}
// Instance batches after the analytic objects (intersection/index.ts:83-91 tlas walk +
// features/intersection.ts:981-1066 instanceLeafItem): a binary TLAS over the placements'
// world boxes; per placement the ray goes to the prototype's frame (t is preserved) and walks
// the batch's wide BVH on the same LDS stack. Instance ids (inst_prim0 + base + local) exceed
// every other id; ties across placements keep the (t, id) rule: a placement whose ids are all
// below the current best accepts an equal t.
template <bool ANY, int BLOCK, bool COUNT>
PT_DEV bool trace_instances(const DevScene& S, f3 o, f3 d, float& tmax, int& prim, uint2* stk, TraceCounters& cnt) {
    if (S.n_inst_nodes == 0) return false;
    const f3 inv = mk3(1.0f / (fabsf(d.x) < 1e-20f ? copysignf(1e-20f, d.x) : d.x),
                       1.0f / (fabsf(d.y) < 1e-20f ? copysignf(1e-20f, d.y) : d.y),
                       1.0f / (fabsf(d.z) < 1e-20f ? copysignf(1e-20f, d.z) : d.z));
    int stack[64]; int sp = 0; stack[0] = 0;
    bool found = false;
    while (sp >= 0) {
        const int ni = stack[sp--];
        const float4 n0 = S.inst_nodes[2 * ni], n1 = S.inst_nodes[2 * ni + 1];
        if (!bin_aabb_hit(n0, n1, o, inv, tmax)) continue;
        if (n0.w >= 0.0f) {
            const int off = (int)n1.w, count = (int)n0.w;
            for (int j = off; j < off + count; j++) {
                const int i = S.inst_order[j];
                const int batch = S.inst_batch[i];
                const int root = S.ow_root[batch];
                if (root < 0) continue;
                const float4 q = S.inst_q[i], ts = S.inst_ts[i];
                const f3 r = qrot(q, o) + xyz(ts), rd = qrot(q, d);
                const f3 lo = mk3(r.x / ts.w, r.y / ts.w, r.z / ts.w), ld = mk3(rd.x / ts.w, rd.y / ts.w, rd.z / ts.w);
                const int pbase = S.inst_prim0 + S.inst_base[i];
                const int shade_first = S.n_tris + S.objects[batch].first_tri;
                TravState st;
                trace_init(st, lo, ld, tmax, (uint32_t)root);
                const int init = (!ANY && prim >= 0 && prim > pbase) ? 0x7fffffff : -1;
                st.prim = init;
                while (!trace_step<ANY, BLOCK, COUNT>(S.ow_nodes, S.ow_tris, st, stk, cnt)) {}
                if (ANY) {
                    if (st.prim >= 0 && st.prim != init) { prim = pbase + (st.prim - shade_first); return true; }
                } else if (st.prim != init) {
                    tmax = st.tmax; prim = pbase + (st.prim - shade_first); found = true;
                }
            }
        } else if (sp + 2 < 64) {
            stack[++sp] = ni + 1; stack[++sp] = (int)n1.w;
        }
    }
    return found;
}

// Analytic objects after the meshes; their ids (n_tris + object) exceed every triangle id,
// so a strict < keeps the (t, id) rule.
template <bool ANY>
PT_DEV bool trace_analytic(const DevScene& S, f3 o, f3 d, float& tmax, int& prim) {
    bool found = false;
    for (int k = 0; k < S.n_analytic; k++) {
        int oid = S.analytic[k];
        const pt_object& ob = S.objects[oid];
        float t; bool ok;
        if (ob.kind == PT_OBJ_QUAD) ok = quad_intersect(o, d, ob.p, t);
        else ok = sphere_intersect(o, d, ld3(ob.p), ob.p[3], t);
        if (ok && t < tmax) {
            if (ANY) return true;
            tmax = t; prim = S.n_tris + oid; found = true;
        }
    }
    return found;
}

}  // namespace ptmi
