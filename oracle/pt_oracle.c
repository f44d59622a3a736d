/*
 * pt_oracle.c — CPU restatement of the reference's generated transport loop.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: only tests/, the smoke check and the
 * CPU-baseline leg of bench.py load it (liboracle.so via ctypes). The product (libptmi.so)
 * never links, calls or includes it.
 *
 * It follows the GLSL the reference emits for a mesh scene + quad/env lights + pt/pt-nee/
 * pt-mis, statement by statement, in IEEE fp32 with no FMA contraction and with the
 * transcendentals of include/ptmi_math.h (the declared arithmetic). Each function cites the
 * reference source it restates. Declared deviations (also listed in the header of each):
 *   - closest-hit ties at bit-equal t resolve to the smaller global primitive id (the
 *     reference keeps the first one its BVH order visits, intersection/mesh/mesh.glsl:94),
 *     and the node prune therefore admits tenter == tmax (accel/bvh/bvh.glsl:28-36 uses <);
 *   - analytic primitives are tested after the meshes, with the same (t, id) rule;
 *   - slab tests take the inverse direction with the CWBVH zero guard (see guard_inv);
 *   - BLAS triangle boxes are widened by the Moller-Trumbore tolerance, so every walk equals
 *     the brute-force closest hit (the reference's bvh.test.ts asserts BVH == brute force).
 * Host-side builders (binned-SAH BVH over f64 boxes, env CDF) follow the TS sources.
 */
#define _GNU_SOURCE
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#include "../include/ptmi.h"
#include "../include/ptmi_math.h"

/* Planted mistakes (the liboracle_plants variant only): a test switches one on with
 * or_set_plant(k) to show that its check fails when the oracle computes a plausible wrong
 * thing. The default builds have no such code path. k = 1: the MIS light pdf takes light 0's
 * selection pdf for every light; 2: the light-selection random is reused for the light's own
 * sample without the cdf_rescale; 3: d log hg / dg with the sign of its (g - c) term flipped;
 * 4: the adjoint's in-scattering term -1/sigma_s instead of +1/sigma_s; 5: the adjoint's
 * sigma_a transmittance term dropped; 6: Rayleigh's Cardano root with its second cube root added
 * instead of subtracted; 7: a Gaussian band's exponent with one factor of (lambda - center). */
#ifdef ORACLE_PLANTS
static int g_plant = 0;
void or_set_plant(int k) { g_plant = k; }
#define PLANT(k) (g_plant == (k))
#else
#define PLANT(k) 0
#endif

#ifdef ORACLE_LIBM
/* libm variant: GLSL builtins mapped to the C library instead of the declared arithmetic.
 * Used only to show the declared arithmetic does not move the estimate. */
#define SIN sinf
#define COS cosf
#define EXP expf
#define LOG logf
#define ATAN2 atan2f
#define POW powf
#define TAN tanf
#else
#define SIN ptm_sin
#define COS ptm_cos
#define EXP ptm_exp
#define LOG ptm_log
#define ATAN2 ptm_atan2
#define POW ptm_pow
#define TAN ptm_tan
#endif

/* ---------------- vectors (GLSL semantics, left-to-right evaluation) ---------------- */
typedef struct { float x, y, z; } v3;
static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 muls(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 divs(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
static inline v3 neg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float length3(v3 a) { return sqrtf(dot(a, a)); }
static inline v3 normalize(v3 a) { return muls(a, 1.0f / sqrtf(dot(a, a))); }
static inline float fmaxf_(float a, float b) { return a > b ? a : b; }   /* GLSL max(a,b) */
static inline float fminf_(float a, float b) { return a < b ? a : b; }
static inline float clampf(float x, float lo, float hi) { return fminf_(fmaxf_(x, lo), hi); }
static inline v3 ld3(const float* p) { return V(p[0], p[1], p[2]); }
static inline float signf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

#define MAX_DIST 1000.0f
#define EPS_INTERFACE 0.001f
#define MARCH_CLEARANCE 0.001f
#define SHADOW_BACKOFF 0.002f
#define BVH_TFAR_PAD 1.00000024f
#define INV_PI (1.0f / PTM_PI)

/* ---------------- RNG: pcg4d (components/sampler/pcg4d/pcg4d.glsl:12-40) ---------------- */
typedef struct { uint32_t base, dim; } rng_t;
/* the path's wavelengths: lam[0] is the hero (spectral programs) */
typedef struct { int spec; int uniform; float lam[4]; } wl_t;
static inline void pcg4d(uint32_t v[4]) {
    for (int i = 0; i < 4; i++) v[i] = v[i] * 1664525u + 1013904223u;
    v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
    for (int i = 0; i < 4; i++) v[i] ^= v[i] >> 16u;
    v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
}
static inline void rng_init(rng_t* r, uint32_t px, uint32_t py, uint32_t sample, uint32_t salt) {
    uint32_t v[4] = {px, py, sample, salt};
    pcg4d(v);
    r->base = v[0];
    r->dim = 0;
}
static inline uint32_t rng_u32(rng_t* r) {
    uint32_t v[4] = {r->base, r->dim++, 0u, 0u};
    pcg4d(v);
    return v[0];
}
static inline float rnd(rng_t* r) { return (float)(rng_u32(r) >> 8) * (1.0f / 16777216.0f); }

void or_pcg4d(const uint32_t* in, uint32_t* out, int n) {
    for (int i = 0; i < n; i++) {
        uint32_t v[4] = {in[4 * i], in[4 * i + 1], in[4 * i + 2], in[4 * i + 3]};
        pcg4d(v);
        memcpy(out + 4 * i, v, 16);
    }
}
void or_random_stream(uint32_t px, uint32_t py, uint32_t sample, uint32_t salt, float* out, int n) {
    rng_t r; rng_init(&r, px, py, sample, salt);
    for (int i = 0; i < n; i++) out[i] = rnd(&r);
}

/* ---------------- binned-SAH binary BVH (accel/bvh/bvh.ts:132-334) ----------------
 * f64 boxes, f32 centroids and f32 SAH areas, 12 bins over all three axes, leaf <= 2,
 * strict < cost ties, median fallback. Node = (min.xyz, A, max.xyz, B):
 * A >= 0 leaf (count A at offset B of order[]), A < 0 internal (axis -A-1, right B). */
#define BVH_BINS 12
typedef struct { float* nodes; int n_nodes; int cap; uint32_t* order; int n; } bvh_t;

typedef struct {
    const double* boxes; float *cx, *cy, *cz; uint32_t* order; bvh_t* out; int leaf_size;
} bvh_build_t;

static int bvh_push_node(bvh_t* b) {
    if (b->n_nodes == b->cap) { b->cap = b->cap ? 2 * b->cap : 64; b->nodes = (float*)realloc(b->nodes, sizeof(float) * 8 * (size_t)b->cap); }
    return b->n_nodes++;
}
static double box_area(const double* b) {
    double dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0;
    return 2 * (dx * dy + dy * dz + dz * dx);
}
static void box_reset(double* b) { b[0] = b[1] = b[2] = INFINITY; b[3] = b[4] = b[5] = -INFINITY; }
static void box_grow(double* b, const double* s) {
    for (int a = 0; a < 3; a++) { if (s[a] < b[a]) b[a] = s[a]; if (s[3 + a] > b[3 + a]) b[3 + a] = s[3 + a]; }
}

static int bvh_emit(bvh_build_t* B, int start, int end) {
    bvh_t* out = B->out;
    int node = bvh_push_node(out);
    double bounds[6], cb[6];
    box_reset(bounds); box_reset(cb);
    for (int i = start; i < end; i++) {
        uint32_t t = B->order[i]; const double* bx = B->boxes + 6 * (size_t)t;
        for (int a = 0; a < 3; a++) { if (bx[a] < bounds[a]) bounds[a] = bx[a]; if (bx[3 + a] > bounds[3 + a]) bounds[3 + a] = bx[3 + a]; }
        double c[3] = {B->cx[t], B->cy[t], B->cz[t]};
        for (int a = 0; a < 3; a++) { if (c[a] < cb[a]) cb[a] = c[a]; if (c[a] > cb[3 + a]) cb[3 + a] = c[a]; }
    }
    int count = end - start;
#define SET_NODE(A_, B_) do { float* nd = out->nodes + 8 * (size_t)node; \
        nd[0] = (float)bounds[0]; nd[1] = (float)bounds[1]; nd[2] = (float)bounds[2]; nd[3] = (float)(A_); \
        nd[4] = (float)bounds[3]; nd[5] = (float)bounds[4]; nd[6] = (float)bounds[5]; nd[7] = (float)(B_); } while (0)
    if (count <= B->leaf_size) { SET_NODE(count, start); return node; }
    int bestAxis = -1, bestSplit = -1; double bestCost = INFINITY;
    int v[3] = {cb[3] - cb[0] >= 1e-12, cb[4] - cb[1] >= 1e-12, cb[5] - cb[2] >= 1e-12};
    if (v[0] || v[1] || v[2]) {
        double s[3], lo[3] = {cb[0], cb[1], cb[2]};
        for (int a = 0; a < 3; a++) s[a] = v[a] ? BVH_BINS / (cb[3 + a] - cb[a]) : 0;
        double binBox[3 * BVH_BINS][6]; int binCnt[3 * BVH_BINS];
        for (int b = 0; b < 3 * BVH_BINS; b++) { box_reset(binBox[b]); binCnt[b] = 0; }
        for (int i = start; i < end; i++) {
            uint32_t t = B->order[i]; const double* bx = B->boxes + 6 * (size_t)t;
            float c[3] = {B->cx[t], B->cy[t], B->cz[t]};
            for (int a = 0; a < 3; a++) {
                if (!v[a]) continue;
                int b = (int)floor(((double)c[a] - lo[a]) * s[a]);
                if (b < 0) b = 0; if (b >= BVH_BINS) b = BVH_BINS - 1;
                binCnt[a * BVH_BINS + b]++;
                box_grow(binBox[a * BVH_BINS + b], bx);
            }
        }
        for (int a = 0; a < 3; a++) {
            if (!v[a]) continue;
            float leftArea[BVH_BINS - 1], rightArea[BVH_BINS - 1];
            int leftCnt[BVH_BINS - 1], rightCnt[BVH_BINS - 1];
            double acc[6]; int accCnt = 0;
            box_reset(acc);
            for (int b = 0; b < BVH_BINS - 1; b++) {
                box_grow(acc, binBox[a * BVH_BINS + b]); accCnt += binCnt[a * BVH_BINS + b];
                leftArea[b] = (float)box_area(acc); leftCnt[b] = accCnt;
            }
            box_reset(acc); accCnt = 0;
            for (int b = BVH_BINS - 1; b > 0; b--) {
                box_grow(acc, binBox[a * BVH_BINS + b]); accCnt += binCnt[a * BVH_BINS + b];
                rightArea[b - 1] = (float)box_area(acc); rightCnt[b - 1] = accCnt;
            }
            for (int b = 0; b < BVH_BINS - 1; b++) {
                if (leftCnt[b] == 0 || rightCnt[b] == 0) continue;
                double cost = (double)leftArea[b] * (double)leftCnt[b] + (double)rightArea[b] * (double)rightCnt[b];
                if (cost < bestCost) { bestCost = cost; bestAxis = a; bestSplit = b; }
            }
        }
    }
    double leafCost = box_area(bounds) * count;
    if (bestAxis == -1 || bestCost >= leafCost) { SET_NODE(count, start); return node; }
    int mid = -1;
    {
        double plo = cb[bestAxis], scale = BVH_BINS / (cb[3 + bestAxis] - cb[bestAxis]);
        const float* cen = bestAxis == 0 ? B->cx : bestAxis == 1 ? B->cy : B->cz;
        int m = start;
        for (int i = start; i < end; i++) {
            uint32_t t = B->order[i];
            int b = (int)floor(((double)cen[t] - plo) * scale);
            if (b < 0) b = 0; if (b >= BVH_BINS) b = BVH_BINS - 1;
            if (b <= bestSplit) { uint32_t tmp = B->order[i]; B->order[i] = B->order[m]; B->order[m] = tmp; m++; }
        }
        if (m > start && m < end) mid = m;
    }
    if (mid == -1) mid = (start + end) >> 1;
    double keep[6]; memcpy(keep, bounds, sizeof keep);
    bvh_emit(B, start, mid);
    int right = bvh_emit(B, mid, end);
    memcpy(bounds, keep, sizeof keep);
    SET_NODE(-(bestAxis + 1), right);
#undef SET_NODE
    return node;
}

/* boxes: n x 6 f64 (min, max). */
static void bvh_build(bvh_t* out, const double* boxes, int n, int leaf_size) {
    memset(out, 0, sizeof *out);
    out->n = n;
    out->order = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    bvh_build_t B;
    B.boxes = boxes; B.out = out; B.leaf_size = leaf_size; B.order = out->order;
    B.cx = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    B.cy = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    B.cz = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    for (int t = 0; t < n; t++) {
        B.cx[t] = (float)((boxes[6 * t] + boxes[6 * t + 3]) * 0.5);
        B.cy[t] = (float)((boxes[6 * t + 1] + boxes[6 * t + 4]) * 0.5);
        B.cz[t] = (float)((boxes[6 * t + 2] + boxes[6 * t + 5]) * 0.5);
        out->order[t] = (uint32_t)t;
    }
    if (n > 0) bvh_emit(&B, 0, n);
    free(B.cx); free(B.cy); free(B.cz);
}
static void bvh_free(bvh_t* b) { free(b->nodes); free(b->order); memset(b, 0, sizeof *b); }

/* The restated builder on its own (test infrastructure: the library's byte-faithful builder is
 * pinned against it): nodes (8 f32 each, capacity 8 (2n - 1)) and the item order. Returns the
 * node count. */
int or_bvh_build(const double* boxes, int n, int leaf_size, float* nodes, uint32_t* order) {
    bvh_t b;
    bvh_build(&b, boxes, n, leaf_size);
    if (b.n_nodes > 0) memcpy(nodes, b.nodes, sizeof(float) * 8 * (size_t)b.n_nodes);
    if (n > 0) memcpy(order, b.order, sizeof(uint32_t) * (size_t)n);
    int nn = b.n_nodes;
    bvh_free(&b);
    return nn;
}

/* ---------------- the light tree (accel/light_tree/light_tree.ts:46-77) ----------------
 * buildBVHCore (accel/bvh/bvh.ts:132-320) with the light tree's parameters: the split cost
 * A_L Phi_L + A_R Phi_R, one light per leaf, a split always (the sorted-median fallback where
 * SAH has no plane), balanced splits once depth + ceil(log2 count) reaches 48, and per light
 * the root-to-leaf bits (0 left, 1 right) LSB first, levels 0-23 in lo and 24-47 in hi.
 * Node = (min.xyz, Phi) (max.xyz, link): link >= 0 right child (left = node + 1), link < 0
 * the leaf of light -link - 1. */
#define LT_MAX_DEPTH 48
typedef struct {
    const double* boxes; const double* pw; float *cx, *cy, *cz; uint32_t* order;
    float* nodes; float* trails; int n_nodes; int max_depth;
} lt_build_t;

static int lt_ceil_log2(int c) { int k = 0; while ((1 << k) < c) k++; return k; }

static int lt_emit(lt_build_t* B, int start, int end, int depth, uint32_t lo, uint32_t hi) {
    if (depth > B->max_depth) B->max_depth = depth;
    int node = B->n_nodes++;
    double bounds[6], cb[6], w = 0.0;
    box_reset(bounds); box_reset(cb);
    for (int i = start; i < end; i++) {
        uint32_t t = B->order[i]; const double* bx = B->boxes + 6 * (size_t)t;
        for (int a = 0; a < 3; a++) { if (bx[a] < bounds[a]) bounds[a] = bx[a]; if (bx[3 + a] > bounds[3 + a]) bounds[3 + a] = bx[3 + a]; }
        double c[3] = {B->cx[t], B->cy[t], B->cz[t]};
        for (int a = 0; a < 3; a++) { if (c[a] < cb[a]) cb[a] = c[a]; if (c[a] > cb[3 + a]) cb[3 + a] = c[a]; }
        w += B->pw[t];
    }
    float* nd = B->nodes + 8 * (size_t)node;
    nd[0] = (float)bounds[0]; nd[1] = (float)bounds[1]; nd[2] = (float)bounds[2]; nd[3] = (float)w;
    nd[4] = (float)bounds[3]; nd[5] = (float)bounds[4]; nd[6] = (float)bounds[5];
    int count = end - start;
    if (count == 1) {
        uint32_t light = B->order[start];
        nd[7] = -(float)light - 1.0f;
        B->trails[4 * (size_t)light] = (float)lo; B->trails[4 * (size_t)light + 1] = (float)hi;
        return node;
    }
    int balanced = depth + lt_ceil_log2(count) >= LT_MAX_DEPTH;
    int bestAxis = -1, bestSplit = -1; double bestCost = INFINITY;
    int v[3] = {cb[3] - cb[0] >= 1e-12, cb[4] - cb[1] >= 1e-12, cb[5] - cb[2] >= 1e-12};
    if (!balanced && (v[0] || v[1] || v[2])) {
        double s[3];
        for (int a = 0; a < 3; a++) s[a] = v[a] ? BVH_BINS / (cb[3 + a] - cb[a]) : 0;
        double binBox[3 * BVH_BINS][6], binW[3 * BVH_BINS]; int binCnt[3 * BVH_BINS];
        for (int b = 0; b < 3 * BVH_BINS; b++) { box_reset(binBox[b]); binCnt[b] = 0; binW[b] = 0.0; }
        for (int i = start; i < end; i++) {
            uint32_t t = B->order[i]; const double* bx = B->boxes + 6 * (size_t)t;
            float c[3] = {B->cx[t], B->cy[t], B->cz[t]};
            for (int a = 0; a < 3; a++) {
                if (!v[a]) continue;
                int b = (int)floor(((double)c[a] - cb[a]) * s[a]);
                if (b < 0) b = 0; if (b >= BVH_BINS) b = BVH_BINS - 1;
                binCnt[a * BVH_BINS + b]++; binW[a * BVH_BINS + b] += B->pw[t];
                box_grow(binBox[a * BVH_BINS + b], bx);
            }
        }
        for (int a = 0; a < 3; a++) {
            if (!v[a]) continue;
            float la[BVH_BINS - 1], ra[BVH_BINS - 1]; int lc[BVH_BINS - 1], rc[BVH_BINS - 1];
            double lw[BVH_BINS - 1], rw[BVH_BINS - 1], acc[6], aw = 0.0; int an = 0;
            box_reset(acc);
            for (int b = 0; b < BVH_BINS - 1; b++) {
                box_grow(acc, binBox[a * BVH_BINS + b]); an += binCnt[a * BVH_BINS + b]; aw += binW[a * BVH_BINS + b];
                la[b] = (float)box_area(acc); lc[b] = an; lw[b] = aw;
            }
            box_reset(acc); an = 0; aw = 0.0;
            for (int b = BVH_BINS - 1; b > 0; b--) {
                box_grow(acc, binBox[a * BVH_BINS + b]); an += binCnt[a * BVH_BINS + b]; aw += binW[a * BVH_BINS + b];
                ra[b - 1] = (float)box_area(acc); rc[b - 1] = an; rw[b - 1] = aw;
            }
            for (int b = 0; b < BVH_BINS - 1; b++) {
                if (lc[b] == 0 || rc[b] == 0) continue;
                double cost = (double)la[b] * lw[b] + (double)ra[b] * rw[b];
                if (cost < bestCost) { bestCost = cost; bestAxis = a; bestSplit = b; }
            }
        }
    }
    int mid = -1;
    if (bestAxis != -1) {
        double plo = cb[bestAxis], scale = BVH_BINS / (cb[3 + bestAxis] - cb[bestAxis]);
        const float* cen = bestAxis == 0 ? B->cx : bestAxis == 1 ? B->cy : B->cz;
        int m = start;
        for (int i = start; i < end; i++) {
            uint32_t t = B->order[i];
            int b = (int)floor(((double)cen[t] - plo) * scale);
            if (b < 0) b = 0; if (b >= BVH_BINS) b = BVH_BINS - 1;
            if (b <= bestSplit) { uint32_t tmp = B->order[i]; B->order[i] = B->order[m]; B->order[m] = tmp; m++; }
        }
        if (m > start && m < end) mid = m;
    }
    if (mid == -1) {   /* sorted median: a stable insertion sort by centroid on the widest axis, halved */
        double e0 = cb[3] - cb[0], e1 = cb[4] - cb[1], e2 = cb[5] - cb[2];
        const float* cen = (e0 >= e1 && e0 >= e2) ? B->cx : (e1 >= e2 ? B->cy : B->cz);
        for (int i = start + 1; i < end; i++) {
            uint32_t t = B->order[i]; int j = i - 1;
            while (j >= start && cen[B->order[j]] > cen[t]) { B->order[j + 1] = B->order[j]; j--; }
            B->order[j + 1] = t;
        }
        mid = (start + end) >> 1;
    }
    uint32_t rlo = depth < 24 ? lo | (1u << depth) : lo;
    uint32_t rhi = depth < 24 ? hi : hi | (1u << (depth - 24));
    lt_emit(B, start, mid, depth + 1, lo, hi);
    int right = lt_emit(B, mid, end, depth + 1, rlo, rhi);
    B->nodes[8 * (size_t)node + 7] = (float)right;
    return node;
}
/* nodes: 8 (2n - 1) floats, trails 4 n. Returns the node count; *max_depth the deepest leaf. */
int or_light_tree_build(const double* boxes, const double* powers, int n, float* nodes, float* trails, int* max_depth) {
    if (n <= 0) { if (max_depth) *max_depth = 0; return 0; }
    lt_build_t B;
    memset(&B, 0, sizeof B);
    B.boxes = boxes; B.pw = powers; B.nodes = nodes; B.trails = trails;
    B.order = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    B.cx = (float*)malloc(sizeof(float) * (size_t)n); B.cy = (float*)malloc(sizeof(float) * (size_t)n);
    B.cz = (float*)malloc(sizeof(float) * (size_t)n);
    for (int t = 0; t < n; t++) {
        B.cx[t] = (float)((boxes[6 * t] + boxes[6 * t + 3]) * 0.5);
        B.cy[t] = (float)((boxes[6 * t + 1] + boxes[6 * t + 4]) * 0.5);
        B.cz[t] = (float)((boxes[6 * t + 2] + boxes[6 * t + 5]) * 0.5);
        B.order[t] = (uint32_t)t;
    }
    memset(trails, 0, sizeof(float) * 4 * (size_t)n);
    lt_emit(&B, 0, n, 0, 0u, 0u);
    free(B.order); free(B.cx); free(B.cy); free(B.cz);
    if (max_depth) *max_depth = B.max_depth;
    return B.n_nodes;
}

/* Inverse ray direction for the slab tests. The BLAS walk (mesh.glsl:167) takes 1/rd, which
 * for an exactly axis-aligned ray gives 0*inf = NaN in the slab test (undefined in GLSL);
 * declared rule: the CWBVH walk's guard |d| < 1e-20 -> +-1e-20 (accel/cwbvh, cwbvh.ts). */
static inline float guard_inv(float d) { return 1.0f / (fabsf(d) < 1e-20f ? copysignf(1e-20f, d) : d); }
static inline v3 inv_dir(v3 d) { return V(guard_inv(d.x), guard_inv(d.y), guard_inv(d.z)); }

/* bvh_aabb_hit (accel/bvh/bvh.glsl:28-36), prune admitting tenter == tmax (tie rule). */
static inline int aabb_hit(const float* n, v3 ro, v3 inv, float tmax) {
    v3 t0 = mul(sub(V(n[0], n[1], n[2]), ro), inv);
    v3 t1 = mul(sub(V(n[4], n[5], n[6]), ro), inv);
    v3 tsm = V(fminf_(t0.x, t1.x), fminf_(t0.y, t1.y), fminf_(t0.z, t1.z));
    v3 tbg = V(fmaxf_(t0.x, t1.x), fmaxf_(t0.y, t1.y), fmaxf_(t0.z, t1.z));
    float tn = fmaxf_(fmaxf_(tsm.x, tsm.y), tsm.z);
    float tf = fminf_(fminf_(tbg.x, tbg.y), tbg.z) * BVH_TFAR_PAD;
    float tenter = fmaxf_(tn, 0.0f);
    return tf >= tenter && tenter <= tmax;
}

/* ---------------- scene ---------------- */
typedef struct {
    int n_tris; float* tri;   /* [n][9]: world triangles, then prototypes */
    float* nrm;               /* [n][9] per-corner shading normals (smooth objects), or NULL */
    int32_t* tri_object;
    int n_objects; pt_object* objects;
    int n_materials; pt_material* materials;
    int n_lights; pt_light* lights;
    pt_env env;
    bvh_t* blas;              /* per object (meshes only) */
    bvh_t tlas;               /* over mesh objects */
    int* tlas_obj;            /* tlas item -> object */
    int n_mesh_objects;
    /* env table */
    float* env_cond; float* env_marg; int env_w, env_h; double env_total;
    float env_rot_cs[2];
    int has_transmission;
    int emissive_solid;       /* some closed/solid region is emissive */
    int ambient_material;     /* material of region -1 (ambientMedium), -1 = none */
    int has_media, has_null;
    int weighted_absorption;  /* some scattering medium weights absorption (Planner.ts:545-547) */
    /* instance batches: prototype triangles follow the world triangles in tri[]; each batch
     * object's BLAS is over its prototype range; a TLAS over the placements' world boxes */
    int n_proto;
    int* obj_first;           /* absolute first triangle of every object's range in tri[] */
    int n_inst;
    float* inst;              /* [n_inst][8]: q xyzw, t_rigid xyz, s */
    int* inst_batch;          /* owning batch object */
    int* inst_base;           /* first global primitive id offset of each placement */
    int inst_prim0;           /* n_tris + n_objects: the first instance primitive id */
    bvh_t itlas;
    float spec_curves[95 * 4];/* CIE X, Y, Z, D65 at 360 + 5 i nm (spectral programs) */
    float xyz_rgb[9], cie_y;
    int has_spectral;
    float* atlas;             /* the coefficient atlas [3 * 64^3][4] (null: none) and its z nodes */
    float atlas_z[64];
    float* lt_nodes;          /* the light tree (8 floats per node) and the lights' trails (4 each), or null */
    float* lt_trails;
    float* img; int img_w, img_h;   /* PT_ENV_IMAGE: the equirect picture, RGB */
} or_scene;

/* Möller–Trumbore (intersection/mesh/mesh.glsl:41-59). */
static inline int tri_test(v3 ro, v3 rd, v3 a, v3 b, v3 c, float* t_out, v3* g_out) {
    v3 e1 = sub(b, a), e2 = sub(c, a);
    v3 g = cross(e1, e2);
    float det = -dot(rd, g);
    if (fabsf(det) < 1e-12f) return 0;
    float invdet = 1.0f / det;
    v3 ao = sub(ro, a);
    v3 dao = cross(ao, rd);
    float u = dot(e2, dao) * invdet;
    float v = -dot(e1, dao) * invdet;
    float t = dot(ao, g) * invdet;
    float w = 1.0f - u - v;
    *t_out = t; *g_out = g;
    const float E = 1e-6f;
    return u >= -E && v >= -E && w >= -E;
}
static inline void tri_load(const or_scene* S, int i, v3* a, v3* b, v3* c) {
    const float* p = S->tri + 9 * (size_t)i;
    *a = V(p[0], p[1], p[2]); *b = V(p[3], p[4], p[5]); *c = V(p[6], p[7], p[8]);
}
/* mesh_test_range's smooth shading normal (mesh.glsl:84-109): the hit's barycentrics from
 * mesh_tri_test, bary.x n_a + bary.y n_b + bary.z n_c, turned to the geometric normal's ray
 * side. ro, rd: the ray in the mesh's frame. */
static v3 smooth_normal(const or_scene* S, int i, v3 ro, v3 rd, v3 a, v3 b, v3 c) {
    v3 e1 = sub(b, a), e2 = sub(c, a);
    v3 g = cross(e1, e2);
    float invdet = 1.0f / -dot(rd, g);
    v3 ao = sub(ro, a);
    v3 dao = cross(ao, rd);
    float u = dot(e2, dao) * invdet;
    float v = -dot(e1, dao) * invdet;
    float w = 1.0f - u - v;
    const float* n = S->nrm + 9 * (size_t)i;
    v3 ns = add(add(muls(V(n[0], n[1], n[2]), w), muls(V(n[3], n[4], n[5]), u)), muls(V(n[6], n[7], n[8]), v));
    if (dot(rd, ns) * dot(rd, g) < 0.0f) ns = neg(ns);
    return ns;
}

/* Nearest over one mesh BLAS (mesh.glsl:162-191) with the (t, id) tie rule. */
static void blas_nearest(const or_scene* S, int obj, v3 ro, v3 rd, float* tmax, int* best) {
    const bvh_t* B = &S->blas[obj];
    if (B->n_nodes == 0) return;
    int first = S->obj_first[obj];
    v3 inv = inv_dir(rd);
    int stack[64]; int ptr = 0; stack[0] = 0;
    while (ptr >= 0) {
        int ni = stack[ptr--];
        const float* n = B->nodes + 8 * (size_t)ni;
        if (!aabb_hit(n, ro, inv, *tmax)) continue;
        if (n[3] >= 0.0f) {
            int off = (int)n[7], cnt = (int)n[3];
            for (int i = off; i < off + cnt; i++) {
                int tri = first + (int)B->order[i];
                v3 a, b, c, g; float t;
                tri_load(S, tri, &a, &b, &c);
                if (tri_test(ro, rd, a, b, c, &t, &g) && t > 0.0f && (t < *tmax || (t == *tmax && tri < *best))) {
                    *tmax = t; *best = tri;
                }
            }
        } else {
            int axis = (int)(-n[3] - 1.0f);
            int L = ni + 1, R = (int)n[7];
            float rda = axis == 0 ? rd.x : axis == 1 ? rd.y : rd.z;
            int nearFirst = rda >= 0.0f;
            if (ptr + 2 < 64) { stack[++ptr] = nearFirst ? R : L; stack[++ptr] = nearFirst ? L : R; }
        }
    }
}
static int blas_any(const or_scene* S, int obj, v3 ro, v3 rd, float maxDist) {
    const bvh_t* B = &S->blas[obj];
    if (B->n_nodes == 0) return 0;
    int first = S->obj_first[obj];
    v3 inv = inv_dir(rd);
    int stack[64]; int ptr = 0; stack[0] = 0;
    while (ptr >= 0) {
        int ni = stack[ptr--];
        const float* n = B->nodes + 8 * (size_t)ni;
        if (!aabb_hit(n, ro, inv, maxDist)) continue;
        if (n[3] >= 0.0f) {
            int off = (int)n[7], cnt = (int)n[3];
            for (int i = off; i < off + cnt; i++) {
                v3 a, b, c, g; float t;
                tri_load(S, first + (int)B->order[i], &a, &b, &c);
                if (tri_test(ro, rd, a, b, c, &t, &g) && t > 0.0f && t < maxDist) return 1;
            }
        } else if (ptr + 2 < 64) {
            stack[++ptr] = ni + 1; stack[++ptr] = (int)n[7];
        }
    }
    return 0;
}

/* Analytic primitives (geometry/sphere/sphere.glsl:28-46, geometry/quad/quad.glsl:9-22). */
static int sphere_intersect(v3 ro, v3 rd, v3 c, float r, float* t) {
    v3 oc = sub(ro, c);
    float b = dot(oc, rd);
    v3 perp = sub(oc, muls(rd, b));
    float disc = r * r - dot(perp, perp);
    if (disc < 0.0f) return 0;
    float cc = dot(oc, oc) - r * r;
    float q = (b >= 0.0f) ? -b - sqrtf(disc) : -b + sqrtf(disc);
    if (q == 0.0f) return 0;
    float ta = q, tb = cc / q;
    *t = (cc < 0.0f) ? fmaxf_(ta, tb) : fminf_(ta, tb);
    return *t > 0.0f;
}
static int quad_intersect(v3 ro, v3 rd, const float* p, float* t) {
    v3 corner = ld3(p), e1 = ld3(p + 3), e2 = ld3(p + 6), n = ld3(p + 9);
    float denom = dot(rd, n);
    if (fabsf(denom) < 1e-8f) return 0;
    *t = dot(sub(corner, ro), n) / denom;
    if (*t <= 0.0f) return 0;
    v3 local = sub(add(ro, muls(rd, *t)), corner);
    float e11 = dot(e1, e1), e22 = dot(e2, e2), e12 = dot(e1, e2);
    float d1 = dot(local, e1), d2 = dot(local, e2);
    float det = e11 * e22 - e12 * e12;
    float u = (d1 * e22 - d2 * e12) / det;
    float v = (d2 * e11 - d1 * e12) / det;
    return u >= 0.0f && u <= 1.0f && v >= 0.0f && v <= 1.0f;
}

/* Placements (glsl/core/placement.glsl:19-41): v + 2 q.xyz x (q.xyz x v + q.w v). */
static inline v3 qrot(const float* q, v3 v) {
    v3 qv = V(q[0], q[1], q[2]);
    v3 c = add(cross(qv, v), muls(v, q[3]));
    return add(v, muls(cross(qv, c), 2.0f));
}
static inline void inst_ray(const float* I, v3 ro, v3 rd, v3* lo, v3* ld) {
    v3 r = add(qrot(I, ro), V(I[4], I[5], I[6]));
    v3 d = qrot(I, rd);
    *lo = V(r.x / I[7], r.y / I[7], r.z / I[7]);
    *ld = V(d.x / I[7], d.y / I[7], d.z / I[7]);
}
static inline v3 inst_normal(const float* I, v3 n) {
    float qc[4] = {-I[0], -I[1], -I[2], I[3]};
    return qrot(qc, n);
}
/* The prototype BLAS in the placement's frame; candidates compete by GLOBAL id. */
static void instance_nearest(const or_scene* S, int i, v3 ro, v3 rd, float* tmax, int* best) {
    const float* I = S->inst + 8 * (size_t)i;
    int obj = S->inst_batch[i];
    const bvh_t* B = &S->blas[obj];
    if (B->n_nodes == 0) return;
    v3 lo, ld; inst_ray(I, ro, rd, &lo, &ld);
    int first = S->obj_first[obj];
    int pbase = S->inst_prim0 + S->inst_base[i];
    v3 inv = inv_dir(ld);
    int stack[64]; int ptr = 0; stack[0] = 0;
    while (ptr >= 0) {
        int ni = stack[ptr--];
        const float* n = B->nodes + 8 * (size_t)ni;
        if (!aabb_hit(n, lo, inv, *tmax)) continue;
        if (n[3] >= 0.0f) {
            int off = (int)n[7], cnt = (int)n[3];
            for (int k = off; k < off + cnt; k++) {
                int local = (int)B->order[k];
                v3 a, b, c, g; float t;
                tri_load(S, first + local, &a, &b, &c);
                int prim = pbase + local;
                if (tri_test(lo, ld, a, b, c, &t, &g) && t > 0.0f && (t < *tmax || (t == *tmax && prim < *best))) {
                    *tmax = t; *best = prim;
                }
            }
        } else if (ptr + 2 < 64) {
            stack[++ptr] = ni + 1; stack[++ptr] = (int)n[7];
        }
    }
}
static int instance_any(const or_scene* S, int i, v3 ro, v3 rd, float maxDist) {
    const float* I = S->inst + 8 * (size_t)i;
    int obj = S->inst_batch[i];
    const bvh_t* B = &S->blas[obj];
    if (B->n_nodes == 0) return 0;
    v3 lo, ld; inst_ray(I, ro, rd, &lo, &ld);
    int first = S->obj_first[obj];
    v3 inv = inv_dir(ld);
    int stack[64]; int ptr = 0; stack[0] = 0;
    while (ptr >= 0) {
        int ni = stack[ptr--];
        const float* n = B->nodes + 8 * (size_t)ni;
        if (!aabb_hit(n, lo, inv, maxDist)) continue;
        if (n[3] >= 0.0f) {
            int off = (int)n[7], cnt = (int)n[3];
            for (int k = off; k < off + cnt; k++) {
                v3 a, b, c, g; float t;
                tri_load(S, first + (int)B->order[k], &a, &b, &c);
                if (tri_test(lo, ld, a, b, c, &t, &g) && t > 0.0f && t < maxDist) return 1;
            }
        } else if (ptr + 2 < 64) {
            stack[++ptr] = ni + 1; stack[++ptr] = (int)n[7];
        }
    }
    return 0;
}
/* Instance TLAS over the placements' world boxes (intersection/index.ts:83-91). */
static void instances_nearest(const or_scene* S, v3 ro, v3 rd, float* tmax, int* best) {
    if (S->itlas.n_nodes == 0) return;
    v3 inv = inv_dir(rd);
    int stack[64]; int ptr = 0; stack[0] = 0;
    while (ptr >= 0) {
        int ni = stack[ptr--];
        const float* n = S->itlas.nodes + 8 * (size_t)ni;
        if (!aabb_hit(n, ro, inv, *tmax)) continue;
        if (n[3] >= 0.0f) {
            int off = (int)n[7], cnt = (int)n[3];
            for (int j = 0; j < cnt; j++) instance_nearest(S, (int)S->itlas.order[off + j], ro, rd, tmax, best);
        } else if (ptr + 2 < 64) {
            stack[++ptr] = ni + 1; stack[++ptr] = (int)n[7];
        }
    }
}
static int instances_any(const or_scene* S, v3 ro, v3 rd, float maxDist) {
    if (S->itlas.n_nodes == 0) return 0;
    v3 inv = inv_dir(rd);
    int stack[64]; int ptr = 0; stack[0] = 0;
    while (ptr >= 0) {
        int ni = stack[ptr--];
        const float* n = S->itlas.nodes + 8 * (size_t)ni;
        if (!aabb_hit(n, ro, inv, maxDist)) continue;
        if (n[3] >= 0.0f) {
            int off = (int)n[7], cnt = (int)n[3];
            for (int j = 0; j < cnt; j++) if (instance_any(S, (int)S->itlas.order[off + j], ro, rd, maxDist)) return 1;
        } else if (ptr + 2 < 64) {
            stack[++ptr] = ni + 1; stack[++ptr] = (int)n[7];
        }
    }
    return 0;
}
/* The placement of an instance primitive id. */
static inline int instance_of(const or_scene* S, int prim) {
    int rel = prim - S->inst_prim0, lo = 0, hi = S->n_inst - 1;
    while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (S->inst_base[mid] <= rel) lo = mid; else hi = mid - 1; }
    return lo;
}

/* Closest hit over the whole scene: scene TLAS over mesh objects (leaf size 1,
 * app/sceneData.ts:88 + features/intersection.ts:1411-1468), then the analytic objects.
 * Returns the global primitive id (-1 = none) and shrinks *tmax. */
static int scene_closest(const or_scene* S, v3 ro, v3 rd, float* tmax) {
    int best = -1;
    if (S->tlas.n_nodes > 0) {
        v3 inv = inv_dir(rd);
        int stack[64]; int ptr = 0; stack[0] = 0;
        while (ptr >= 0) {
            int ni = stack[ptr--];
            const float* n = S->tlas.nodes + 8 * (size_t)ni;
            if (!aabb_hit(n, ro, inv, *tmax)) continue;
            if (n[3] >= 0.0f) {
                int off = (int)n[7], cnt = (int)n[3];
                for (int j = 0; j < cnt; j++) blas_nearest(S, S->tlas_obj[S->tlas.order[off + j]], ro, rd, tmax, &best);
            } else {
                int axis = (int)(-n[3] - 1.0f);
                int L = ni + 1, R = (int)n[7];
                float rda = axis == 0 ? rd.x : axis == 1 ? rd.y : rd.z;
                int nearFirst = rda >= 0.0f;
                if (ptr + 2 < 64) { stack[++ptr] = nearFirst ? R : L; stack[++ptr] = nearFirst ? L : R; }
            }
        }
    }
    for (int o = 0; o < S->n_objects; o++) {
        const pt_object* ob = &S->objects[o];
        float t; int ok = 0;
        if (ob->kind == PT_OBJ_QUAD) ok = quad_intersect(ro, rd, ob->p, &t);
        else if (ob->kind == PT_OBJ_SPHERE) ok = sphere_intersect(ro, rd, ld3(ob->p), ob->p[3], &t);
        if (ok && t < *tmax) { *tmax = t; best = S->n_tris + o; }
    }
    instances_nearest(S, ro, rd, tmax, &best);
    return best;
}
static int scene_any(const or_scene* S, v3 ro, v3 rd, float maxDist) {
    if (instances_any(S, ro, rd, maxDist)) return 1;
    for (int o = 0; o < S->n_objects; o++) {
        const pt_object* ob = &S->objects[o];
        float t;
        if (ob->kind == PT_OBJ_QUAD && quad_intersect(ro, rd, ob->p, &t) && t < maxDist) return 1;
        if (ob->kind == PT_OBJ_SPHERE && sphere_intersect(ro, rd, ld3(ob->p), ob->p[3], &t) && t < maxDist) return 1;
    }
    if (S->tlas.n_nodes == 0) return 0;
    v3 inv = inv_dir(rd);
    int stack[64]; int ptr = 0; stack[0] = 0;
    while (ptr >= 0) {
        int ni = stack[ptr--];
        const float* n = S->tlas.nodes + 8 * (size_t)ni;
        if (!aabb_hit(n, ro, inv, maxDist)) continue;
        if (n[3] >= 0.0f) {
            int off = (int)n[7], cnt = (int)n[3];
            for (int j = 0; j < cnt; j++) if (blas_any(S, S->tlas_obj[S->tlas.order[off + j]], ro, rd, maxDist)) return 1;
        } else if (ptr + 2 < 64) {
            stack[++ptr] = ni + 1; stack[++ptr] = (int)n[7];
        }
    }
    return 0;
}

/* ---------------- containment: scene_region_at (features/intersection.ts:1121-1265) ---- */
static const v3 MESH_INSIDE_DIR = {0.5320544f, 0.7396997f, 0.4114743f};

/* mesh_inside_bvh (mesh.glsl:202-225): first-hit facing along a fixed direction. */
static int mesh_inside(const or_scene* S, int obj, v3 p) {
    const bvh_t* B = &S->blas[obj];
    if (B->n_nodes == 0) return 0;
    int first = S->obj_first[obj];
    float tmax = 1.0e20f, sideDot = 0.0f; int best = -1;
    v3 dir = MESH_INSIDE_DIR;
    v3 inv = V(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
    int stack[64]; int ptr = 0; stack[0] = 0;
    while (ptr >= 0) {
        int ni = stack[ptr--];
        const float* n = B->nodes + 8 * (size_t)ni;
        if (!aabb_hit(n, p, inv, tmax)) continue;
        if (n[3] >= 0.0f) {
            int off = (int)n[7], cnt = (int)n[3];
            for (int i = off; i < off + cnt; i++) {
                int tri = first + (int)B->order[i];
                v3 a, b, c, g; float t;
                tri_load(S, tri, &a, &b, &c);
                if (tri_test(p, dir, a, b, c, &t, &g) && t > 0.0f && (t < tmax || (t == tmax && tri < best))) {
                    tmax = t; best = tri; sideDot = dot(dir, g);
                }
            }
        } else {
            int axis = (int)(-n[3] - 1.0f);
            int L = ni + 1, R = (int)n[7];
            float da = axis == 0 ? dir.x : axis == 1 ? dir.y : dir.z;
            int nearFirst = da >= 0.0f;
            if (ptr + 2 < 64) { stack[++ptr] = nearFirst ? R : L; stack[++ptr] = nearFirst ? L : R; }
        }
    }
    return sideDot > 0.0f;
}
/* Ericson closest point on triangle, squared (mesh.glsl:236-262). */
static float point_tri_dist2(v3 p, v3 a, v3 b, v3 c) {
    v3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) { v3 q = sub(p, a); return dot(q, q); }
    v3 bp = sub(p, b);
    float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) { v3 q = sub(p, b); return dot(q, q); }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) { v3 q = sub(ap, muls(ab, d1 / (d1 - d3))); return dot(q, q); }
    v3 cp = sub(p, c);
    float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) { v3 q = sub(p, c); return dot(q, q); }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) { v3 q = sub(ap, muls(ac, d2 / (d2 - d6))); return dot(q, q); }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        v3 q = sub(bp, muls(sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
        return dot(q, q);
    }
    float denom = 1.0f / (va + vb + vc);
    v3 closest = add(add(a, muls(ab, vb * denom)), muls(ac, vc * denom));
    v3 q = sub(p, closest);
    return dot(q, q);
}
static float box_dist2(const float* n, v3 p) {
    v3 d = V(fmaxf_(fmaxf_(n[0] - p.x, 0.0f), p.x - n[4]), fmaxf_(fmaxf_(n[1] - p.y, 0.0f), p.y - n[5]),
             fmaxf_(fmaxf_(n[2] - p.z, 0.0f), p.z - n[6]));
    return dot(d, d);
}
/* mesh_closest_bvh (mesh.glsl:266-296): the minimum is order independent. */
static float mesh_closest(const or_scene* S, int obj, v3 p) {
    const bvh_t* B = &S->blas[obj];
    if (B->n_nodes == 0) return 1.0e20f;
    int first = S->obj_first[obj];
    float best2 = 1.0e30f;
    int stack[64]; int ptr = 0; stack[0] = 0;
    while (ptr >= 0) {
        int ni = stack[ptr--];
        const float* n = B->nodes + 8 * (size_t)ni;
        if (box_dist2(n, p) >= best2) continue;
        if (n[3] >= 0.0f) {
            int off = (int)n[7], cnt = (int)n[3];
            for (int i = off; i < off + cnt; i++) {
                v3 a, b, c; tri_load(S, first + (int)B->order[i], &a, &b, &c);
                best2 = fminf_(best2, point_tri_dist2(p, a, b, c));
            }
        } else {
            int L = ni + 1, R = (int)n[7];
            int leftNear = box_dist2(B->nodes + 8 * (size_t)L, p) <= box_dist2(B->nodes + 8 * (size_t)R, p);
            if (ptr + 2 < 64) { stack[++ptr] = leftNear ? R : L; stack[++ptr] = leftNear ? L : R; }
        }
    }
    return sqrtf(best2);
}
/* Instanced batches claiming an interior (features/intersection.ts:1196-1215): point descent
 * of the instance TLAS; per placement, the prototype's frame (÷ s), the local box, the
 * first-hit-facing inside test, and the world distance s x closest-triangle distance. */
static int region_instances(const or_scene* S, v3 p, int region, float best) {
    if (S->itlas.n_nodes == 0) return region;
    int stack[64]; int ptr = 0; stack[0] = 0;
    while (ptr >= 0) {
        int ni = stack[ptr--];
        const float* n = S->itlas.nodes + 8 * (size_t)ni;
        if (!(p.x >= n[0] && p.y >= n[1] && p.z >= n[2] && p.x <= n[4] && p.y <= n[5] && p.z <= n[6])) continue;
        if (n[3] >= 0.0f) {
            int off = (int)n[7], cnt = (int)n[3];
            for (int j = 0; j < cnt; j++) {
                int i = (int)S->itlas.order[off + j];
                int o = S->inst_batch[i];
                const pt_object* ob = &S->objects[o];
                if (!ob->closed) continue;
                const float* I = S->inst + 8 * (size_t)i;
                v3 r = add(qrot(I, p), V(I[4], I[5], I[6]));
                v3 lp = V(r.x / I[7], r.y / I[7], r.z / I[7]);
                float d = 1.0e20f;
                if (lp.x >= ob->box_min[0] && lp.y >= ob->box_min[1] && lp.z >= ob->box_min[2] &&
                    lp.x <= ob->box_max[0] && lp.y <= ob->box_max[1] && lp.z <= ob->box_max[2] && mesh_inside(S, o, lp))
                    d = -I[7] * mesh_closest(S, o, lp);
                if (d < 0.0f && d > best) { best = d; region = o; }
            }
        } else if (ptr + 2 < 64) {
            stack[++ptr] = ni + 1; stack[++ptr] = (int)n[7];
        }
    }
    return region;
}
static int region_at(const or_scene* S, v3 p) {
    int region = -1; float best = -1.0e20f;
    for (int o = 0; o < S->n_objects; o++) {
        const pt_object* ob = &S->objects[o];
        if (ob->kind != PT_OBJ_SPHERE) continue;
        float d = length3(sub(p, ld3(ob->p))) - ob->p[3];
        if (d < 0.0f && d > best) { best = d; region = o; }
    }
    /* closed meshes: point descent of the scene TLAS (only leaves whose box holds p) */
    if (S->tlas.n_nodes == 0) return region_instances(S, p, region, best);
    int stack[64]; int ptr = 0; stack[0] = 0;
    while (ptr >= 0) {
        int ni = stack[ptr--];
        const float* n = S->tlas.nodes + 8 * (size_t)ni;
        if (!(p.x >= n[0] && p.y >= n[1] && p.z >= n[2] && p.x <= n[4] && p.y <= n[5] && p.z <= n[6])) continue;
        if (n[3] >= 0.0f) {
            int off = (int)n[7], cnt = (int)n[3];
            for (int j = 0; j < cnt; j++) {
                int o = S->tlas_obj[S->tlas.order[off + j]];
                const pt_object* ob = &S->objects[o];
                if (!ob->closed) continue;
                float d = 1.0e20f;
                if (p.x >= ob->box_min[0] && p.y >= ob->box_min[1] && p.z >= ob->box_min[2] &&
                    p.x <= ob->box_max[0] && p.y <= ob->box_max[1] && p.z <= ob->box_max[2] && mesh_inside(S, o, p))
                    d = -1.0f * mesh_closest(S, o, p);
                if (d < 0.0f && d > best) { best = d; region = o; }
            }
        } else if (ptr + 2 < 64) {
            stack[++ptr] = ni + 1; stack[++ptr] = (int)n[7];
        }
    }
    return region_instances(S, p, region, best);
}
static int object_thin(const or_scene* S, int o) {
    const pt_object* ob = &S->objects[o];
    return ob->kind == PT_OBJ_QUAD || ((ob->kind == PT_OBJ_MESH || ob->kind == PT_OBJ_INSTANCES) && !ob->closed);
}

/* ---------------- hit (structs.glsl Hit; features/intersection.ts:889-948,1665-1728) ---- */
typedef struct { v3 p, t, b, n; } frame_t;
typedef struct {
    float t; v3 p; frame_t frame; v3 ng;
    int region_from, region_to, region_owner; float eps; int prim;
} hit_t;

static inline float fp_uncertainty(v3 p) {
    return fmaxf_(1.52587890625e-5f, 3.0517578125e-5f * fmaxf_(fabsf(p.x), fmaxf_(fabsf(p.y), fabsf(p.z))));
}
/* ambient_frame (components/ambient/euclidean/euclidean.glsl:13-19) */
static inline frame_t ambient_frame(v3 p, v3 n) {
    frame_t f;
    v3 nn = normalize(n);
    v3 t = fabsf(nn.x) < 0.9f ? V(1.0f, 0.0f, 0.0f) : V(0.0f, 1.0f, 0.0f);
    t = normalize(sub(t, muls(nn, dot(t, nn))));
    f.p = p; f.t = t; f.b = cross(nn, t); f.n = nn;
    return f;
}

static int scene_intersect(const or_scene* S, v3 ro, v3 rd, hit_t* hit) {
    float tmax = MAX_DIST;
    int prim = scene_closest(S, ro, rd, &tmax);
    if (prim < 0) return 0;
    hit->t = tmax; hit->prim = prim;
    hit->p = add(ro, muls(rd, tmax));
    hit->eps = fp_uncertainty(hit->p);
    if (prim < S->n_tris) {
        v3 a, b, c; tri_load(S, prim, &a, &b, &c);
        v3 g = cross(sub(b, a), sub(c, a));
        hit->frame = ambient_frame(hit->p, normalize(S->objects[S->tri_object[prim]].smooth ? smooth_normal(S, prim, ro, rd, a, b, c) : g));
        hit->ng = normalize(g);
        hit->region_owner = S->tri_object[prim];
    } else if (prim >= S->inst_prim0) {   /* instanceLeafItem mesh arm: normals by the forward rotation */
        int i = instance_of(S, prim);
        int o = S->inst_batch[i];
        v3 a, b, c; tri_load(S, S->obj_first[o] + (prim - S->inst_prim0 - S->inst_base[i]), &a, &b, &c);
        v3 gl = cross(sub(b, a), sub(c, a));
        v3 ng = normalize(inst_normal(S->inst + 8 * (size_t)i, gl));
        v3 ns = ng;
        if (S->objects[o].smooth) {
            v3 lo, ld; inst_ray(S->inst + 8 * (size_t)i, ro, rd, &lo, &ld);
            ns = normalize(inst_normal(S->inst + 8 * (size_t)i,
                                       smooth_normal(S, S->obj_first[o] + (prim - S->inst_prim0 - S->inst_base[i]), lo, ld, a, b, c)));
        }
        hit->frame = ambient_frame(hit->p, ns);
        hit->ng = ng;
        hit->region_owner = o;
    } else {
        int o = prim - S->n_tris;
        const pt_object* ob = &S->objects[o];
        v3 n = ob->kind == PT_OBJ_QUAD ? ld3(ob->p + 9) : normalize(sub(hit->p, ld3(ob->p)));
        hit->frame = ambient_frame(hit->p, n);
        hit->ng = hit->frame.n;
        hit->region_owner = o;
    }
    int outside = region_at(S, add(hit->p, muls(hit->ng, EPS_INTERFACE)));
    if (dot(rd, hit->ng) < 0.0f) {
        hit->region_from = outside;
        hit->region_to = hit->region_owner;
    } else {
        hit->region_from = object_thin(S, hit->region_owner)
            ? region_at(S, add(hit->p, muls(hit->ng, -EPS_INTERFACE))) : hit->region_owner;
        hit->region_to = outside;
        hit->frame = ambient_frame(hit->p, neg(hit->frame.n));
        hit->ng = neg(hit->ng);
    }
    return 1;
}

/* material_of (generated): the ambient region -1 maps to the ambientMedium material. */
static inline int material_of(const or_scene* S, int region) { return region < 0 ? S->ambient_material : S->objects[region].material; }
static inline int is_null_interface(const or_scene* S, int mat) { return mat >= 0 && S->materials[mat].model == PT_MAT_NULL; }
static inline int material_has_medium(const or_scene* S, int mat) { return mat >= 0 && S->materials[mat].has_medium; }
/* cauchy_index / sellmeier_index (components/color/spectra/cauchy/cauchy.glsl,
 * sellmeier/sellmeier.glsl): coefficients in micrometres, the carrier in nm. */
static inline float cauchy_index(const float* c, float lam) {
    float inv2 = 1e6f / (lam * lam);
    return c[0] + inv2 * (c[1] + inv2 * c[2]);
}
static inline float sellmeier_index(const float* c, float lam) {
    float l2 = (lam * 0.001f) * (lam * 0.001f);
    float t0 = c[0] * l2 / (l2 - c[3]), t1 = c[1] * l2 / (l2 - c[4]), t2 = c[2] * l2 / (l2 - c[5]);
    return sqrtf(1.0f + t0 + t1 + t2);
}
/* Regions with a row in the generated ior_of table (features/intersection.ts:1609-1649): every
 * primitive object, closed meshes and closed instance batches; open meshes are thin (1.0). */
static inline int region_has_index_row(const pt_object* ob) {
    return !((ob->kind == PT_OBJ_MESH || ob->kind == PT_OBJ_INSTANCES) && !ob->closed);
}
/* The models that cross their interface and read the region's index (the descriptors'
 * 'transmission' capability: dielectric, rough_dielectric). */
static inline int transmissive_model(int model) { return model == PT_MAT_DIELECTRIC || model == PT_MAT_ROUGH_DIELECTRIC; }
/* ior_of: the transmissive material's index, at the hero wavelength wavelengths_hero
 * (spectral.glsl:12-15) for a dispersive formula under the spectral colour model. */
static inline float ior_of(const or_scene* S, int region, const wl_t* wl) {
    if (region < 0) return 1.0f;
    const pt_object* ob = &S->objects[region];
    if (!region_has_index_row(ob)) return 1.0f;
    const pt_material* m = &S->materials[ob->material];
    if (!transmissive_model(m->model)) return 1.0f;
    if (wl->spec && m->ior_kind == PT_IOR_CAUCHY) return cauchy_index(m->ior_coef, wl->lam[0]);
    if (wl->spec && m->ior_kind == PT_IOR_SELLMEIER) return sellmeier_index(m->ior_coef, wl->lam[0]);
    return m->ior;
}
/* A region of the program's dispersion record (Planner.ts:1038-1053 planDispersion). */
static inline int dispersive_region(const or_scene* S, int region) {
    if (region < 0) return 0;
    const pt_object* ob = &S->objects[region];
    const pt_material* m = &S->materials[ob->material];
    return region_has_index_row(ob) && transmissive_model(m->model) && m->ior_kind != PT_IOR_CONSTANT;
}

static inline v3 lv(const float* a) { return V(a[0], a[1], a[2]); }

/* ---------------- spectra (components/color/spectrum.glsl, components/color/spectral) ----
 * Four lanes. An RGB program uses lanes 0-2 and reduces over 3; a spectral program carries
 * the path's four hero wavelengths in wl_t and reduces over 4. */
typedef struct { float x, y, z, w; } sp;
static inline sp SP(float x, float y, float z, float w) { sp r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
static inline sp sgrey(float v) { return SP(v, v, v, v); }
static inline sp sadd(sp a, sp b) { return SP(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
static inline sp smul(sp a, sp b) { return SP(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
static inline sp smuls(sp a, float s) { return SP(a.x * s, a.y * s, a.z * s, a.w * s); }
static inline sp sdivs(sp a, float s) { return SP(a.x / s, a.y / s, a.z / s, a.w / s); }
static inline sp sneg(sp a) { return SP(-a.x, -a.y, -a.z, -a.w); }
static inline sp srgb(const float* p) { return SP(p[0], p[1], p[2], 0.0f); }
static inline sp srgbv(v3 v) { return SP(v.x, v.y, v.z, 0.0f); }
static inline float s_lane(sp s, int c) { return c == 0 ? s.x : (c == 1 ? s.y : (c == 2 ? s.z : s.w)); }
static inline float s_avg(const wl_t* wl, sp s) {   /* spectrum_average */
    float sum = s.x; sum += s.y; sum += s.z;
    if (wl->spec) { sum += s.w; return sum * 0.25f; }
    return sum * (1.0f / 3.0f);
}
static inline float s_max(const wl_t* wl, sp s) {   /* spectrum_max */
    float m = wl->spec ? s.w : s.z;
    if (wl->spec) m = fmaxf_(s.z, m);
    m = fmaxf_(s.y, m); m = fmaxf_(s.x, m);
    return m;
}
static inline int s_black(const wl_t* wl, sp s) { return s.x <= 0.0f && s.y <= 0.0f && s.z <= 0.0f && (!wl->spec || s.w <= 0.0f); }
static inline sp s_exp(sp s) { return SP(EXP(s.x), EXP(s.y), EXP(s.z), EXP(s.w)); }
static inline float glsl_mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
/* spectral_curves (lookup.glsl:2-7) */
static inline void spectral_curves(const or_scene* S, float lam, float out[4]) {
    float x = (lam - 360.0f) / 5.0f;
    int i = (int)x; if (i > 93) i = 93;
    float f = x - (float)i;
    for (int c = 0; c < 4; c++) out[c] = glsl_mix(S->spec_curves[4 * i + c], S->spec_curves[4 * (i + 1) + c], f);
}
/* spectral_bounded (lookup.glsl:40-48) for one lane */
static inline float spectral_bounded_lane(const pt_spectrum* r, float lam) {
    float t = (lam - 360.0f) / 470.0f;
    float p = (r->c[0] * t + r->c[1]) * t + r->c[2];
    return clampf(0.5f + 0.5f * p / sqrtf(1.0f + p * p), 0.0f, 1.0f);
}
/* blackbody_emission (color/spectra/blackbody/blackbody.glsl:4-10) */
static inline float blackbody_lane(float kelvin, float normalization, float lam) {
    float q = 14387768.775039338f / kelvin;
    float ratio = 550.0f / lam;
    float ratio2 = ratio * ratio;
    return normalization * (EXP(q / 550.0f) - 1.0f) * ratio2 * ratio2 * ratio / (EXP(q / lam) - 1.0f);
}
/* color_lift_{reflectance,coefficient,emission} (lookup.glsl:50-60), scale * blackbody, or a
 * Gaussian reflectance band (color/spectra/gaussian/gaussian.ts emit) */
static sp spectrum_eval(const or_scene* S, const pt_spectrum* r, const wl_t* wl) {
    float v[4];
    for (int i = 0; i < 4; i++) {
        float lam = wl->lam[i], x;
        if (r->kind == PT_SPEC_BLACKBODY) x = r->scale * blackbody_lane(r->c[0], r->c[1], lam);
        else if (r->kind == PT_SPEC_GAUSSIAN) {   /* gaussian.ts emit: floor + (peak - floor) exp(B s s) */
            float sh = lam - r->c[0];
            x = r->value + r->c[2] * EXP(PLANT(7) ? r->c[1] * sh : (r->c[1] * sh) * sh);
        } else {
            x = r->kind == PT_SPEC_ZERO ? 0.0f : (r->kind == PT_SPEC_GREY ? r->value : spectral_bounded_lane(r, lam));
            if (r->scaled) x = r->scale * x;
            if (r->d65) { float c[4]; spectral_curves(S, lam, c); x = x * c[3]; }
        }
        v[i] = x;
    }
    return SP(v[0], v[1], v[2], v[3]);
}
static inline sp lift(const or_scene* S, const float* rgb, const pt_spectrum* r, const wl_t* wl) {
    return wl->spec ? spectrum_eval(S, r, wl) : srgb(rgb);
}
/* wavelength_from_cdf / wavelength_pdf (color/spectral/visible.glsl:3-13, uniform.glsl:1-2) */
static inline float wl_visible_a(void) { return ptm_tanh(0.0072f * (360.0f - 538.0f)); }
static inline float wl_visible_b(void) { return ptm_tanh(0.0072f * (830.0f - 538.0f)); }
static inline float wavelength_from_cdf(int uniform, float u) {
    if (uniform) return 360.0f + 470.0f * u;
    return 538.0f + ptm_atanh(glsl_mix(wl_visible_a(), wl_visible_b(), u)) / 0.0072f;
}
static inline float wavelength_pdf(int uniform, float lam) {
    if (uniform) return 1.0f / 470.0f;
    float c = ptm_cosh(0.0072f * (lam - 538.0f));
    return 0.0072f / ((wl_visible_b() - wl_visible_a()) * c * c);
}
/* sensor_response (sensor/cie1931/cie1931.glsl:3-5) */
static inline v3 sensor_response(const or_scene* S, float lam) {
    float c[4]; spectral_curves(S, lam, c);
    const float* m = S->xyz_rgb;
    v3 r = V((m[0] * c[0] + m[1] * c[1]) + m[2] * c[2], (m[3] * c[0] + m[4] * c[1]) + m[5] * c[2],
             (m[6] * c[0] + m[7] * c[1]) + m[8] * c[2]);
    return divs(r, S->cie_y);
}

/* ---------------- materials ---------------- */
enum { LOBE_REFLECTION = 1, LOBE_TRANSMISSION = 2, LOBE_DELTA = 4 };
typedef struct { v3 wi; sp weight; float pdf; int flags; } isample_t;

static inline v3 to_local(frame_t f, v3 v) { return V(dot(v, f.t), dot(v, f.b), dot(v, f.n)); }
static inline v3 from_local(frame_t f, v3 v) { return normalize(add(add(muls(f.t, v.x), muls(f.b, v.y)), muls(f.n, v.z))); }
/* microfacet.glsl:29-54 */
static inline float mf_D(v3 h, float a) { float t = h.z * h.z * (a * a - 1.0f) + 1.0f; return a * a / (PTM_PI * t * t); }
static inline float mf_G1(v3 v, float a) { float c = fabsf(v.z); return 2.0f * c / (c + sqrtf(a * a + (1.0f - a * a) * c * c)); }
static v3 mf_vndf(v3 wol, float a, float ux, float uy) {
    v3 vh = normalize(V(a * wol.x, a * wol.y, wol.z));
    float lensq = vh.x * vh.x + vh.y * vh.y;
    v3 T1 = lensq > 0.0f ? muls(V(-vh.y, vh.x, 0.0f), 1.0f / sqrtf(lensq)) : V(1.0f, 0.0f, 0.0f);
    v3 T2 = cross(vh, T1);
    float rr = sqrtf(ux);
    float phi = PTM_TWO_PI * uy;
    float t1 = rr * COS(phi);
    float t2 = rr * SIN(phi);
    float s_ = 0.5f * (1.0f + vh.z);
    t2 = (1.0f - s_) * sqrtf(fmaxf_(0.0f, 1.0f - t1 * t1)) + s_ * t2;
    v3 nh = add(add(muls(T1, t1), muls(T2, t2)), muls(vh, sqrtf(fmaxf_(0.0f, 1.0f - t1 * t1 - t2 * t2))));
    return normalize(V(a * nh.x, a * nh.y, fmaxf_(1e-6f, nh.z)));
}
static inline sp schlick(sp f0, float c) {
    float p = POW(1.0f - c, 5.0f);
    return SP(f0.x + (1.0f - f0.x) * p, f0.y + (1.0f - f0.y) * p, f0.z + (1.0f - f0.z) * p, f0.w + (1.0f - f0.w) * p);
}
/* math.glsl:95-102 */
static inline float dielectric_fresnel(float cos_i, float eta) {
    float sin2_t = eta * eta * (1.0f - cos_i * cos_i);
    if (sin2_t >= 1.0f) return 1.0f;
    float cos_t = sqrtf(1.0f - sin2_t);
    float r_par = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    float r_perp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    return 0.5f * (r_par * r_par + r_perp * r_perp);
}
/* rough_dielectric.glsl:37-46: the generalized half vector of (wi, wo), faced up; 0 for a
 * degenerate pair or a back-facing microfacet. */
static int rough_half(v3 wil, v3 wol, float eta, v3* m) {
    float etap = wil.z > 0.0f ? 1.0f : 1.0f / eta;
    v3 h = add(muls(wil, etap), wol);
    if (dot(h, h) <= 0.0f) return 0;
    h = normalize(h);
    if (h.z < 0.0f) h = neg(h);
    *m = h;
    return dot(h, wil) * wil.z >= 0.0f && dot(h, wol) * wol.z >= 0.0f;
}

/* Hit.uv (glsl/core/structs.glsl:40-43): quad_uv (quad.glsl:33-41), sphere_uv (sphere.glsl:84-89),
 * else the planar placeholder (x, z) * 0.1 */
static void hit_uv(const or_scene* S, const hit_t* h, float* u, float* v) {
    const pt_object* o = &S->objects[h->region_owner];
    if (o->kind == PT_OBJ_QUAD) {
        v3 local = sub(h->p, V(o->p[0], o->p[1], o->p[2]));
        v3 e1 = V(o->p[3], o->p[4], o->p[5]), e2 = V(o->p[6], o->p[7], o->p[8]);
        float e11 = dot(e1, e1), e22 = dot(e2, e2), e12 = dot(e1, e2);
        float d1 = dot(local, e1), d2 = dot(local, e2);
        float det = e11 * e22 - e12 * e12;
        *u = (d1 * e22 - d2 * e12) / det; *v = (d2 * e11 - d1 * e12) / det;
    } else if (o->kind == PT_OBJ_SPHERE) {
        v3 d = normalize(sub(h->p, V(o->p[0], o->p[1], o->p[2])));
        *u = ATAN2(d.z, d.x) / PTM_TWO_PI + 0.5f;
        *v = ptm_acos(clampf(d.y, -1.0f, 1.0f)) / PTM_PI;
    } else {
        *u = h->p.x * 0.1f; *v = h->p.z * 0.1f;
    }
}
/* the Lambert lobe's albedo: the material's, or the checker cell's (checker.glsl:13-17) */
static sp lambert_albedo(const or_scene* S, const pt_material* m, const hit_t* hit, const wl_t* wl) {
    sp a = lift(S, m->albedo, &m->s_albedo, wl);
    if (m->model != PT_MAT_CHECKER) return a;
    float u, v;
    hit_uv(S, hit, &u, &v);
    float x = floorf(u * m->pad1) + floorf(v * m->pad1);
    float k = x - 2.0f * floorf(x / 2.0f);
    if (PLANT(13)) k = 1.0f - k;
    return sadd(smuls(a, 1.0f - k), smuls(lift(S, m->f0, &m->s_f0, wl), k));
}
static isample_t surface_sample(const or_scene* S, const pt_material* m, v3 wo, const hit_t* hit, float uc, float ux, float uy, const wl_t* wl) {
    isample_t s;
    frame_t f = hit->frame;
    if (m->model == PT_MAT_LAMBERT || m->model == PT_MAT_CHECKER) {   /* materials/lambert/lambert.glsl:17-28 */
        v3 n = dot(wo, f.n) < 0.0f ? neg(f.n) : f.n;
        float cos_theta = sqrtf(uy);
        float sin_theta = sqrtf(fmaxf_(0.0f, 1.0f - uy));
        float phi = PTM_TWO_PI * ux;
        s.wi = normalize(add(add(muls(f.t, sin_theta * COS(phi)), muls(f.b, sin_theta * SIN(phi))), muls(n, cos_theta)));
        s.weight = lambert_albedo(S, m, hit, wl);
        s.pdf = cos_theta * INV_PI;
        s.flags = LOBE_REFLECTION;
    } else if (m->model == PT_MAT_DIELECTRIC) {     /* materials/dielectric/dielectric.glsl:14-52 */
        float n_i = ior_of(S, hit->region_from, wl);
        float n_t = ior_of(S, hit->region_to, wl);
        float eta = n_i / n_t;
        v3 n = f.n;
        float cos_i = clampf(dot(wo, n), 1e-6f, 1.0f);
        float F = dielectric_fresnel(cos_i, eta);
        if (uc < F) {
            s.wi = normalize(sub(muls(n, 2.0f * cos_i), wo));
            s.weight = sgrey(1.0f);
            s.pdf = 0.0f;
            s.flags = LOBE_REFLECTION | LOBE_DELTA;
        } else {
            float sin2_t = eta * eta * (1.0f - cos_i * cos_i);
            float cos_t = sqrtf(fmaxf_(0.0f, 1.0f - sin2_t));
            s.wi = normalize(add(muls(wo, -eta), muls(n, eta * cos_i - cos_t)));
            float radiance_scale = (n_i * n_i) / (n_t * n_t);
            s.weight = smuls(lift(S, m->transmittance, &m->s_transmittance, wl), radiance_scale);
            s.pdf = 0.0f;
            s.flags = LOBE_TRANSMISSION | LOBE_DELTA;
        }
    } else if (m->model == PT_MAT_MIRROR) {        /* materials/mirror/mirror.glsl:21-33 */
        v3 n = f.n;
        float cos_i = clampf(dot(wo, n), 1e-6f, 1.0f);
        s.wi = normalize(sub(muls(n, 2.0f * cos_i), wo));
        s.weight = schlick(lift(S, m->f0, &m->s_f0, wl), cos_i);
        s.pdf = 0.0f;
        s.flags = LOBE_REFLECTION | LOBE_DELTA;
    } else if (m->model == PT_MAT_ROUGH_DIELECTRIC) {   /* rough_dielectric.glsl:74-133 */
        v3 wol = to_local(f, wo);
        s.wi = wo; s.weight = sgrey(0.0f); s.pdf = 0.0f; s.flags = LOBE_REFLECTION;
        if (wol.z <= 0.0f) return s;
        float a = m->alpha;
        float n_i = ior_of(S, hit->region_from, wl), n_t = ior_of(S, hit->region_to, wl);
        float eta = n_i / n_t;
        if (n_i == n_t) {                          /* index-matched: the delta pass-through */
            s.wi = neg(wo);
            s.weight = lift(S, m->transmittance, &m->s_transmittance, wl);
            s.flags = LOBE_TRANSMISSION | LOBE_DELTA;
            return s;
        }
        v3 mh = mf_vndf(wol, a, ux, uy);
        float cos_om = clampf(dot(wol, mh), 1e-6f, 1.0f);
        float F = dielectric_fresnel(cos_om, eta);
        v3 wil;
        if (uc < F) {
            v3 I = neg(wol);
            wil = sub(I, muls(mh, 2.0f * dot(mh, I)));
            if (wil.z <= 0.0f) return s;
            s.weight = sgrey(mf_G1(wil, a));
            s.pdf = F * mf_G1(wol, a) * mf_D(mh, a) / (4.0f * wol.z);
            s.flags = LOBE_REFLECTION;
        } else {
            float sin2_t = eta * eta * (1.0f - cos_om * cos_om);
            float cos_t = sqrtf(fmaxf_(0.0f, 1.0f - sin2_t));
            wil = normalize(add(muls(wol, -eta), muls(mh, eta * cos_om - cos_t)));
            if (wil.z >= 0.0f) return s;
            float denom = dot(wil, mh) + dot(wol, mh) * eta;
            denom *= denom;
            s.weight = smuls(lift(S, m->transmittance, &m->s_transmittance, wl), mf_G1(wil, a) * eta * eta);
            s.pdf = (1.0f - F) * mf_G1(wol, a) * mf_D(mh, a) * cos_om / wol.z * fabsf(dot(wil, mh)) / denom;
            s.flags = LOBE_TRANSMISSION;
        }
        s.wi = from_local(f, wil);
    } else {                                        /* materials/ggx/ggx.glsl:27-46 */
        v3 wol = to_local(f, wo);
        float side = wol.z < 0.0f ? -1.0f : 1.0f;
        wol = muls(wol, side);
        float a = m->alpha;
        v3 h = mf_vndf(wol, a, ux, uy);
        v3 I = neg(wol);
        v3 wil = sub(I, muls(h, 2.0f * dot(h, I)));
        if (wil.z <= 0.0f) {
            s.wi = wo; s.weight = sgrey(0.0f); s.pdf = 0.0f; s.flags = LOBE_REFLECTION;
            return s;
        }
        sp F = schlick(lift(S, m->f0, &m->s_f0, wl), fabsf(dot(wol, h)));
        s.weight = smuls(F, mf_G1(wil, a));
        s.pdf = mf_G1(wol, a) * mf_D(h, a) / (4.0f * fabsf(wol.z));
        s.flags = LOBE_REFLECTION;
        s.wi = from_local(f, muls(wil, side));
    }
    return s;
}
static sp surface_eval(const or_scene* S, const pt_material* m, v3 wi, v3 wo, const hit_t* hit, const wl_t* wl) {
    frame_t f = hit->frame;
    if (m->model == PT_MAT_LAMBERT || m->model == PT_MAT_CHECKER) {
        if (dot(wi, f.n) * dot(wo, f.n) <= 0.0f) return sgrey(0.0f);
        return smuls(lambert_albedo(S, m, hit, wl), INV_PI);
    } else if (m->model == PT_MAT_GGX) {
        v3 wil = to_local(f, wi), wol = to_local(f, wo);
        if (wil.z * wol.z <= 0.0f) return sgrey(0.0f);
        float a = m->alpha;
        v3 h = normalize(add(wil, wol));
        sp F = schlick(lift(S, m->f0, &m->s_f0, wl), fabsf(dot(wol, h)));
        return smuls(F, mf_D(h, a) * mf_G1(wil, a) * mf_G1(wol, a) / (4.0f * fabsf(wil.z) * fabsf(wol.z)));
    } else if (m->model == PT_MAT_ROUGH_DIELECTRIC) {   /* rough_dielectric.glsl:48-72 */
        v3 wil = to_local(f, wi), wol = to_local(f, wo);
        if (wol.z <= 0.0f || wil.z == 0.0f) return sgrey(0.0f);
        float n_i = ior_of(S, hit->region_from, wl), n_t = ior_of(S, hit->region_to, wl);
        if (n_i == n_t) return sgrey(0.0f);
        float eta = n_i / n_t;
        v3 mh;
        if (!rough_half(wil, wol, eta, &mh)) return sgrey(0.0f);
        float a = m->alpha;
        float cos_om = clampf(fabsf(dot(wol, mh)), 1e-6f, 1.0f);
        float F = dielectric_fresnel(cos_om, eta);
        float DG = mf_D(mh, a) * mf_G1(wol, a) * mf_G1(wil, a);
        if (wil.z > 0.0f) return sgrey(DG * F / (4.0f * wil.z * wol.z));
        float denom = dot(wil, mh) + dot(wol, mh) * eta;
        denom *= denom;
        float ft = DG * (1.0f - F) * fabsf(dot(wil, mh) * dot(wol, mh) / (wil.z * wol.z * denom));
        return smuls(lift(S, m->transmittance, &m->s_transmittance, wl), ft * eta * eta);
    }
    return sgrey(0.0f);
}
static float surface_pdf(const or_scene* S, const pt_material* m, v3 wi, v3 wo, const hit_t* hit, const wl_t* wl) {
    frame_t f = hit->frame;
    if (m->model == PT_MAT_LAMBERT || m->model == PT_MAT_CHECKER) {
        float c = dot(wi, f.n) * signf(dot(wo, f.n));
        return fmaxf_(0.0f, c) * INV_PI;
    } else if (m->model == PT_MAT_GGX) {
        v3 wil = to_local(f, wi), wol = to_local(f, wo);
        if (wil.z * wol.z <= 0.0f) return 0.0f;
        float a = m->alpha;
        v3 h = normalize(add(wil, wol));
        return mf_G1(wol, a) * mf_D(h, a) / (4.0f * fabsf(wol.z));
    } else if (m->model == PT_MAT_ROUGH_DIELECTRIC) {   /* rough_dielectric.glsl:135-155 */
        v3 wil = to_local(f, wi), wol = to_local(f, wo);
        if (wol.z <= 0.0f || wil.z == 0.0f) return 0.0f;
        float n_i = ior_of(S, hit->region_from, wl), n_t = ior_of(S, hit->region_to, wl);
        if (n_i == n_t) return 0.0f;
        float eta = n_i / n_t;
        v3 mh;
        if (!rough_half(wil, wol, eta, &mh)) return 0.0f;
        float a = m->alpha;
        float cos_om = clampf(fabsf(dot(wol, mh)), 1e-6f, 1.0f);
        float F = dielectric_fresnel(cos_om, eta);
        if (wil.z > 0.0f) return F * mf_G1(wol, a) * mf_D(mh, a) / (4.0f * wol.z);
        float denom = dot(wil, mh) + dot(wol, mh) * eta;
        denom *= denom;
        return (1.0f - F) * mf_G1(wol, a) * mf_D(mh, a) * cos_om / wol.z * fabsf(dot(wil, mh)) / denom;
    }
    return 0.0f;
}
static inline int has_nondelta(const pt_material* m) { return m->model != PT_MAT_DIELECTRIC && m->model != PT_MAT_MIRROR; }

/* ---------------- environment (components/env, features/environment.ts:191-231) ---------------- */
/* The procedural sky bodies (environment.glsl.source) of the reference's scenes, at a unit
 * direction in the table's azimuth frame:
 *   PT_SKY_SUN      proc-sky (generated-glsl.snapshot.test.ts.snap:63758)
 *   PT_SKY_BIG_SUN  proc-sky-rotated (tests/witnesses/scenes/estimatorAgreementWitness.ts:61-72)
 *   PT_SKY_CAP      sky-missed-cap, vec3(step(0.9, dir.y)) (witnesses/scenes/spectralSkies.ts:39-41)
 *   PT_SKY_WHITE    spectral-comp-empty, vec3(1.0) (spectralSkies.ts:48) */
static void equirect_uv(const or_scene* S, v3 dir, float* u, float* v);
static v3 sky_radiance_raw(int body, v3 dir) {
    v3 sun = normalize(V(0.4f, 0.35f, 0.2f));
    if (body == PT_SKY_SUN) {
        float s = 0.5f * (dir.y + 1.0f);
        v3 lo = V(0.06f, 0.08f, 0.12f), hi = V(0.35f, 0.45f, 0.65f);
        v3 base = add(muls(lo, 1.0f - s), muls(hi, s));
        float k = ptm_smoothstep(0.9994f, 0.9999f, dot(dir, sun));
        return add(base, muls(V(18.0f, 16.0f, 13.0f), k));
    }
    if (body == PT_SKY_BIG_SUN) {
        float k = ptm_smoothstep(0.985f, 0.99f, dot(dir, sun));
        return add(V(0.02f, 0.02f, 0.02f), muls(V(40.0f, 40.0f, 40.0f), k));
    }
    if (body == PT_SKY_CAP) {   /* GLSL step(edge, x) = x < edge ? 0 : 1 */
        float v = PLANT(9) ? (0.9f < dir.y ? 0.0f : 1.0f) : (dir.y < 0.9f ? 0.0f : 1.0f);
        return V(v, v, v);
    }
    if (body == PT_SKY_WHITE) return V(1.0f, 1.0f, 1.0f);
    return V(0.0f, 0.0f, 0.0f);
}
/* spectral_coefficients (lookup.glsl:16-38): trilinear lookup of the coefficient atlas for a
 * bounded rgb, channel = the largest (first on ties), x and y = the other two (channel + 1,
 * channel + 2) over it, the z cell by six halvings over the 64 z nodes. */
static void atlas_coefficients(const or_scene* S, v3 rgb, float out[3]) {
    float v[3] = {rgb.x, rgb.y, rgb.z};
    int ch = rgb.x >= rgb.y && rgb.x >= rgb.z ? 0 : (rgb.y >= rgb.z ? 1 : 2);
    float z = v[ch];
    int ax = PLANT(8) ? (ch + 2) % 3 : (ch + 1) % 3, ay = PLANT(8) ? (ch + 1) % 3 : (ch + 2) % 3;
    float x = (v[ax] / z) * 63.0f, y = (v[ay] / z) * 63.0f;
    int ix = (int)x < 62 ? (int)x : 62, iy = (int)y < 62 ? (int)y : 62;
    float fx = x - (float)ix, fy = y - (float)iy;
    int lo = 0, hi = 63;
    for (int k = 0; k < 6; k++) {
        int mid = (lo + hi) / 2;
        if (S->atlas_z[mid] <= z) lo = mid; else hi = mid;
    }
    float fz = (z - S->atlas_z[lo]) / (S->atlas_z[lo + 1] - S->atlas_z[lo]);
    for (int c = 0; c < 3; c++) {
        float p[2];
        for (int dz = 0; dz < 2; dz++) {
#define NODE(xx, yy) S->atlas[4 * ((((size_t)ch * 64 + (size_t)(lo + dz)) * 64 + (size_t)(yy)) * 64 + (size_t)(xx)) + c]
            p[dz] = glsl_mix(glsl_mix(NODE(ix, iy), NODE(ix + 1, iy), fx), glsl_mix(NODE(ix, iy + 1), NODE(ix + 1, iy + 1), fx), fy);
#undef NODE
        }
        out[c] = glsl_mix(p[0], p[1], fz);
    }
}
/* color_lift_emission of an RGB computed at run time (lookup.glsl:50-60) */
static sp lift_emission(const or_scene* S, v3 rgb, const wl_t* wl) {
    pt_spectrum r;
    memset(&r, 0, sizeof r);
    r.scaled = 1; r.d65 = 1;
    r.scale = 2.0f * fmaxf_(rgb.x, fmaxf_(rgb.y, rgb.z));
    if (r.scale == 0.0f) r.kind = PT_SPEC_ZERO;
    else {
        v3 b = V(rgb.x / r.scale, rgb.y / r.scale, rgb.z / r.scale);
        if (b.x == b.y && b.y == b.z) { r.kind = PT_SPEC_GREY; r.value = b.x; }
        else { r.kind = PT_SPEC_SIGMOID; atlas_coefficients(S, b, r.c); }
    }
    return spectrum_eval(S, &r, wl);
}
static sp env_radiance(const or_scene* S, v3 dir, const wl_t* wl) {
    if (S->env.type == PT_ENV_CONSTANT)   /* features/environment.ts:119-133: color * u_envIntensity */
        return smuls(lift(S, S->env.color, &S->env.s_color, wl), S->env.intensity);
    if (S->env.type == PT_ENV_IMAGE) {   /* texture(u_envMap, equirect_uv(dir)): NEAREST, clamped edges */
        float u, v;
        equirect_uv(S, dir, &u, &v);
        int i = (int)(u * (float)S->img_w), j = (int)(v * (float)S->img_h);
        if (i < 0) i = 0; if (i > S->img_w - 1) i = S->img_w - 1;
        if (j < 0) j = 0; if (j > S->img_h - 1) j = S->img_h - 1;
        const float* t = S->img + 3 * ((size_t)j * S->img_w + i);
        v3 c = V(t[0], t[1], t[2]);
        if (!wl->spec) return srgbv(muls(c, S->env.intensity));
        return smuls(lift_emission(S, c, wl), S->env.intensity);
    }
    if (S->env.type != PT_ENV_SKY) return sgrey(0.0f);
    v3 d = normalize(dir);
    float cx = S->env_rot_cs[0], sy = S->env_rot_cs[1];
    d = V(cx * d.x - sy * d.z, d.y, sy * d.x + cx * d.z);
    v3 c = sky_radiance_raw(S->env.body, d);
    if (!wl->spec) return srgbv(muls(c, S->env.intensity));
    return smuls(lift_emission(S, c, wl), S->env.intensity);   /* skyRadiance: lift, then * u_envIntensity */
}
/* equirect chart (env/equirect/equirect_map.glsl, equirect.glsl:19-48) */
static void equirect_uv(const or_scene* S, v3 dir, float* u, float* v) {
    v3 n = normalize(dir);
    float phi = (n.x == 0.0f && n.z == 0.0f ? 0.0f : ATAN2(n.z, n.x)) + S->env.rotation;
    float theta = ATAN2(sqrtf(n.x * n.x + n.z * n.z), n.y);
    *u = PLANT(10) ? clampf(phi * (1.0f / PTM_TWO_PI) + 0.5f, 0.0f, 1.0f) : ptm_fract(phi * (1.0f / PTM_TWO_PI) + 0.5f);
    *v = theta * (1.0f / PTM_PI);
}
static v3 env_chart_dir(float rot, float u, float v) {
    float phi = (u - 0.5f) * PTM_TWO_PI - rot;
    float theta = v * PTM_PI;
    float s = SIN(theta);
    return V(COS(phi) * s, COS(theta), SIN(phi) * s);
}
static float env_texel_dOmega(const or_scene* S, int j, int W, int H) {
    if (S->env.chart == PT_CHART_OCTAHEDRAL) return (4.0f * PTM_PI) / ((float)W * (float)H);   /* octahedral.glsl:21-23 */
    float theta = PTM_PI * ((float)j + 0.5f) / (float)H;
    return (PTM_TWO_PI / (float)W) * 2.0f * SIN(theta) * SIN(0.5f * PTM_PI / (float)H);
}
static v3 env_chart_sample(const or_scene* S, int ci, int cj, float ux, float uy) {
    int W = S->env_w, H = S->env_h;
    if (S->env.chart == PT_CHART_OCTAHEDRAL) {   /* octahedral.glsl:67-69 */
        v3 d;
        ptm_oct_dir(((float)ci + ux) / (float)W, ((float)cj + uy) / (float)H, S->env_rot_cs[0], S->env_rot_cs[1], &d.x, &d.y, &d.z);
        return d;
    }
    float phi = (((float)ci + ux) / (float)W - 0.5f) * PTM_TWO_PI - S->env.rotation;
    int north = 2 * cj < H;
    float row = north ? (float)cj : (float)(H - cj - 1);
    float a = SIN(0.5f * PTM_PI * row / (float)H);
    float b = SIN(0.5f * PTM_PI * (row + 1.0f) / (float)H);
    float tt = north ? uy : 1.0f - uy;
    float h = a * a * (1.0f - tt) + b * b * tt;
    float y = (north ? 1.0f : -1.0f) * (1.0f - 2.0f * h);
    float r = 2.0f * sqrtf(h * (1.0f - h));
    return V(r * COS(phi), y, r * SIN(phi));
}
static inline float env_marg(const or_scene* S, int j) { return S->env_marg[j]; }
static inline float env_cond(const or_scene* S, int j, int i) { return S->env_cond[(size_t)j * S->env_w + i]; }
/* sampler_cdf.glsl:13-84 */
static float env_table_pdf(const or_scene* S, v3 dir) {
    int W = S->env_w, H = S->env_h;
    float u, v;
    if (S->env.chart == PT_CHART_OCTAHEDRAL) ptm_oct_uv(dir.x, dir.y, dir.z, S->env_rot_cs[0], S->env_rot_cs[1], &u, &v);   /* octahedral.glsl:44-65 */
    else equirect_uv(S, dir, &u, &v);
    int i = (int)(u * (float)W); i = i < 0 ? 0 : (i > W - 1 ? W - 1 : i);
    int j = (int)(v * (float)H); j = j < 0 ? 0 : (j > H - 1 ? H - 1 : j);
    float dMarg = env_marg(S, j) - (j > 0 ? env_marg(S, j - 1) : 0.0f);
    float dCond = env_cond(S, j, i) - (i > 0 ? env_cond(S, j, i - 1) : 0.0f);
    return (dCond * dMarg) / env_texel_dOmega(S, j, W, H);
}
static v3 env_sample_table(const or_scene* S, float x0, float x1) {
    int W = S->env_w, H = S->env_h;
    int lo = 0, hi = H - 1;
    while (lo < hi) { int mid = (lo + hi) >> 1; if (x0 >= env_marg(S, mid)) lo = mid + 1; else hi = mid; }
    int j = lo;
    float mLo = (j > 0) ? env_marg(S, j - 1) : 0.0f;
    float dMarg = env_marg(S, j) - mLo;
    float vFrac = (dMarg > 0.0f) ? clampf((x0 - mLo) / dMarg, 0.0f, 0.9999999f) : 0.5f;
    lo = 0; hi = W - 1;
    while (lo < hi) { int mid = (lo + hi) >> 1; if (x1 >= env_cond(S, j, mid)) lo = mid + 1; else hi = mid; }
    int i = lo;
    float cLo = (i > 0) ? env_cond(S, j, i - 1) : 0.0f;
    float dCond = env_cond(S, j, i) - cLo;
    float uFrac = (dCond > 0.0f) ? clampf((x1 - cLo) / dCond, 0.0f, 0.9999999f) : 0.5f;
    return env_chart_sample(S, i, j, uFrac, vFrac);
}
static float env_uniform_fraction(const or_scene* S) {
    if (S->env.compensated) return 0.0f;
    return env_marg(S, S->env_h - 1) > 0.0f ? 0.0625f : 1.0f;
}
static float environment_pdf(const or_scene* S, v3 dir) {
    if (S->env.type == PT_ENV_CONSTANT) return 1.0f / (4.0f * PTM_PI);   /* CONSTANT_PDF */
    float a = env_uniform_fraction(S);
    return (1.0f - a) * env_table_pdf(S, dir) + a / (4.0f * PTM_PI);
}

/* Env bake (compiler/EnvironmentBake.ts:62-67) + importance table (env/importance.ts:16-140). */
/* The importance-table chain after the luminance pass (components/env/importance.ts:47-127): 3 x 3
 * tent blur, MIS compensation when asked, conditional CDF per row and the marginal CDF over rows.
 * Y (W x H luminance) is consumed; the table replaces S's. */
static void env_chain(or_scene* S, float* Y, int W, int H, int oct, int compensated) {
    S->env_w = W; S->env_h = H;
    float* Bl = (float*)malloc(sizeof(float) * (size_t)W * H);
    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++) {
            double sum = 0, wsum = 0;
            for (int dj = -1; dj <= 1; dj++) {
                int jj = j + dj; jj = jj < 0 ? 0 : (jj > H - 1 ? H - 1 : jj);
                for (int di = -1; di <= 1; di++) {
                    /* importance.ts:53-63: equirect wraps u, octahedral clamps both */
                    int ii = oct ? (i + di < 0 ? 0 : (i + di > W - 1 ? W - 1 : i + di)) : (i + di + W) % W;
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
            double dO = oct ? (4 * M_PI) / (W * H) : (2 * M_PI / W) * (M_PI / H) * sin(M_PI * (j + 0.5) / H);   /* importance.ts:77-80 */
            for (int i = 0; i < W; i++) { num += (double)Bl[(size_t)j * W + i] * dO; den += dO; }
        }
        double mean = den > 0 ? num / den : 0;
        for (size_t k = 0; k < (size_t)W * H; k++) { double y = (double)Bl[k] - mean; Bl[k] = (float)(y > 0 ? y : 0); }
    }
    free(S->env_cond); free(S->env_marg);
    S->env_cond = (float*)malloc(sizeof(float) * (size_t)W * H);
    S->env_marg = (float*)malloc(sizeof(float) * (size_t)H);
    double* rowSum = (double*)malloc(sizeof(double) * (size_t)H);
    double total = 0;
    for (int j = 0; j < H; j++) {
        double sinT = oct ? 1.0 : sin(M_PI * (j + 0.5) / H);   /* importance.ts:107-109 */
        double sum = 0;
        for (int i = 0; i < W; i++) { double y = Bl[(size_t)j * W + i]; sum += (y > 0 ? y : 0) * sinT; }
        rowSum[j] = sum;
        double c = 0;
        for (int i = 0; i < W; i++) {
            double y = Bl[(size_t)j * W + i]; c += (y > 0 ? y : 0) * sinT;
            S->env_cond[(size_t)j * W + i] = (float)(sum > 0 ? c / sum : 0);
        }
        total += sum;
    }
    double acc = 0;
    for (int j = 0; j < H; j++) { acc += rowSum[j]; S->env_marg[j] = (float)(total > 0 ? acc / total : 0); }
    S->env_total = total;
    free(Y); free(Bl); free(rowSum);
}

/* The procedural sky's table: its luminance at every texel centre of the chart (the bake pass,
 * compiler/EnvironmentBake.ts:29-80, at rotation 0), then the chain. */
static void env_build(or_scene* S) {
    /* equirect W x H; octahedral N x N with N = table_h (EnvironmentBake.ts:21-27) */
    int oct = S->env.chart == PT_CHART_OCTAHEDRAL;
    int H = S->env.table_h > 0 ? S->env.table_h : 256, W = oct ? H : (S->env.table_w > 0 ? S->env.table_w : 512);
    float* Y = (float*)malloc(sizeof(float) * (size_t)W * H);
    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++) {
            float u = ((float)i + 0.5f) / (float)W, v = ((float)j + 0.5f) / (float)H;
            v3 d;
            if (oct) ptm_oct_dir(u, v, 1.0f, 0.0f, &d.x, &d.y, &d.z);   /* octahedral.glsl:25-42, rotation 0 */
            else d = env_chart_dir(0.0f, u, v);
            v3 c = sky_radiance_raw(S->env.body, d);
            Y[(size_t)j * W + i] = (float)(0.2126 * (double)c.x + 0.7152 * (double)c.y + 0.0722 * (double)c.z);
        }
    env_chain(S, Y, W, H, oct, S->env.compensated);
}

/* resampleEquirectToOctahedral (env/octahedral/octahedral.ts:72-100) at the equalAreaSquareToSphere
 * (:10-26) texel centres, in f64 */
static void or_resample_oct(const float* rgb, int W, int H, int N, float* out) {
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
            double u0 = (i + 0.5) / N, v0 = (j + 0.5) / N;
            double up = 2 * u0 - 1, vp = 2 * v0 - 1, ua = fabs(up), va = fabs(vp);
            double sd = 1 - (ua + va), r = 1 - fabs(sd);
            double ph = ((r == 0 ? 1 : (va - ua) / r) + 1) * M_PI / 4;
            double y = copysign(1 - r * r, sd);
            double s = r * sqrt(2 - r * r > 0 ? 2 - r * r : 0);
            double x = copysign(cos(ph), up) * s, z = copysign(sin(ph), vp) * s;
            double phi = atan2(z, x), theta = acos(y > 1 ? 1 : (y < -1 ? -1 : y));
            double fx = (phi / (2 * M_PI) + 0.5) * W - 0.5, fy = (theta / M_PI) * H - 0.5;
            double x0 = floor(fx), y0 = floor(fy), tx = fx - x0, ty = fy - y0;
            int xa = (int)fmod(fmod(x0, W) + W, W), xb = (int)fmod(fmod(x0 + 1, W) + W, W);
            int ya = (int)(y0 < 0 ? 0 : (y0 > H - 1 ? H - 1 : y0)), yb = (int)(y0 + 1 < 0 ? 0 : (y0 + 1 > H - 1 ? H - 1 : y0 + 1));
            for (int c = 0; c < 3; c++) {
                double s00 = rgb[3 * ((size_t)ya * W + xa) + c], s10 = rgb[3 * ((size_t)ya * W + xb) + c];
                double s01 = rgb[3 * ((size_t)yb * W + xa) + c], s11 = rgb[3 * ((size_t)yb * W + xb) + c];
                out[3 * ((size_t)j * N + i) + c] = (float)((1 - tx) * (1 - ty) * s00 + tx * (1 - ty) * s10 +
                                                           (1 - tx) * ty * s01 + tx * ty * s11);
            }
        }
}
/* An image environment's table: the picture's luminance (equirect) or its octahedral resample's */
static void env_build_image(or_scene* S) {
    int oct = S->env.chart == PT_CHART_OCTAHEDRAL;
    int W = S->img_w, H = S->img_h;
    const float* rgb = S->img;
    float* rs = NULL;
    if (oct) {
        W = H = S->env.table_h > 0 ? S->env.table_h : 256;
        rs = (float*)malloc(sizeof(float) * 3 * (size_t)W * H);
        or_resample_oct(S->img, S->img_w, S->img_h, W, rs);
        rgb = rs;
    }
    float* Y = (float*)malloc(sizeof(float) * (size_t)W * H);
    for (size_t k = 0; k < (size_t)W * H; k++)
        Y[k] = (float)(0.2126 * (double)rgb[3 * k] + 0.7152 * (double)rgb[3 * k + 1] + 0.0722 * (double)rgb[3 * k + 2]);
    env_chain(S, Y, W, H, oct, S->env.compensated);
    free(rs);
}
/* the resample on its own (test probe) */
void or_resample_equirect_to_octahedral(const float* rgb, int W, int H, int N, float* out) { or_resample_oct(rgb, W, H, N, out); }

/* registerEnvironmentTable (engine/Engine.ts:450-460 -> importance.ts:16-127): the table rebuilt
 * from a caller's W x H RGB picture on `chart`; the luminance is the f64 weighted sum of the f32
 * channels (importance.ts:38-45). Returns 0, or -1 for a bad size or a negative / non-finite
 * sample (the reference throws). */
int or_scene_env_table_rgb(or_scene* S, const float* rgb, int W, int H, int chart, int compensated, double* total) {
    int oct = chart == PT_CHART_OCTAHEDRAL;
    if (W < 1 || H < 1 || (oct && W != H) || (S->env.type != PT_ENV_SKY && S->env.type != PT_ENV_IMAGE)) return -1;
    for (size_t k = 0; k < (size_t)3 * W * H; k++)
        if (!isfinite(rgb[k]) || rgb[k] < 0.0f) return -1;
    float* Y = (float*)malloc(sizeof(float) * (size_t)W * H);
    for (size_t k = 0; k < (size_t)W * H; k++)
        Y[k] = (float)(0.2126 * (double)rgb[3 * k] + 0.7152 * (double)rgb[3 * k + 1] + 0.0722 * (double)rgb[3 * k + 2]);
    env_chain(S, Y, W, H, oct, compensated);
    S->env.chart = oct ? PT_CHART_OCTAHEDRAL : PT_CHART_EQUIRECT;
    S->env.table_w = W; S->env.table_h = H; S->env.compensated = compensated ? 1 : 0;
    if (total) *total = S->env_total;
    return 0;
}

/* ---------------- lights (lights/quad/quad.glsl:14-41, features/lighting.ts:546-673) ---- */
typedef struct { v3 wi; float distance; sp radiance; float pdf; int flags; int light_id; } lsample_t;

static lsample_t light_sample_one(const or_scene* S, const pt_light* l, v3 p, float x0, float x1, const wl_t* wl) {
    lsample_t ls;
    if (l->kind == PT_LIGHT_QUAD) {
        v3 q = add(add(lv(l->corner), muls(lv(l->edge1), x0)), muls(lv(l->edge2), x1));
        v3 d = sub(q, p);
        float d2 = dot(d, d);
        ls.wi = muls(d, 1.0f / sqrtf(d2));
        ls.distance = sqrtf(d2);
        ls.radiance = lift(S, l->radiance, &l->s_radiance, wl);
        ls.flags = 0; ls.light_id = -1;
        float cos_l = dot(lv(l->normal), neg(ls.wi));
        if (cos_l <= 0.0f) { ls.pdf = 0.0f; return ls; }
        ls.pdf = d2 / (l->area * cos_l);
    } else {   /* point light (lights/point/point.glsl:11-23) */
        v3 d = sub(lv(l->corner), p);
        float d2 = dot(d, d);
        ls.wi = muls(d, 1.0f / sqrtf(d2));
        ls.distance = sqrtf(d2);
        ls.radiance = sdivs(lift(S, l->radiance, &l->s_radiance, wl), d2);
        ls.pdf = 1.0f; ls.flags = 1; ls.light_id = -1;
    }
    return ls;
}
static float light_pdf_one(const pt_light* l, v3 p, v3 light_p, v3 wi) {
    if (l->kind != PT_LIGHT_QUAD) return 0.0f;
    float cos_l = dot(lv(l->normal), neg(wi));
    if (cos_l <= 0.0f) return 0.0f;
    v3 d = sub(light_p, p);
    return dot(d, d) / (l->area * cos_l);
}
static lsample_t lighting_sample_finite(const or_scene* S, v3 p, float x0, float x1, const wl_t* wl) {
    lsample_t ls;
    int n = S->n_lights;
    if (n == 1) { ls = light_sample_one(S, &S->lights[0], p, x0, x1, wl); ls.light_id = 0; return ls; }
    int id = n - 1; float sel = S->lights[n - 1].select_pdf;
    float lo = n >= 2 ? S->lights[n - 2].cdf : 0.0f;
    for (int i = 0; i < n - 1; i++) {
        if (x0 < S->lights[i].cdf) { id = i; sel = S->lights[i].select_pdf; lo = i == 0 ? 0.0f : S->lights[i - 1].cdf; break; }
    }
    float xr = PLANT(2) ? x0 : (x0 - lo) / sel;
    ls = light_sample_one(S, &S->lights[id], p, clampf(xr, 0.0f, 0.9999999f), x1, wl);
    ls.light_id = id;
    ls.pdf *= sel;
    return ls;
}
static int env_samplable(const or_scene* S) {
    return S->env.type == PT_ENV_SKY || S->env.type == PT_ENV_IMAGE || (S->env.type == PT_ENV_CONSTANT && S->env.samplable);
}
static int env_select_live(const or_scene* S) { return env_samplable(S) && S->n_lights > 0; }
static lsample_t environment_sample(const or_scene* S, float x0, float x1, const wl_t* wl) {
    lsample_t ls;
    if (S->env.type == PT_ENV_CONSTANT) {   /* the constant sampler (features/environment.ts:74-87) */
        float z = 1.0f - 2.0f * x0;
        float r = sqrtf(fmaxf_(0.0f, 1.0f - z * z));
        float phi = PTM_TWO_PI * x1;
        ls.wi = V(r * COS(phi), z, r * SIN(phi));
        ls.distance = MAX_DIST;
        ls.radiance = env_radiance(S, ls.wi, wl);
        ls.pdf = 1.0f / (4.0f * PTM_PI);
        ls.flags = 0; ls.light_id = -1;
        return ls;
    }
    float a = env_uniform_fraction(S);
    if (x0 < a) {
        float y = 1.0f - 2.0f * (x0 / a);
        float r = sqrtf(fmaxf_(0.0f, 1.0f - y * y));
        float phi = PTM_TWO_PI * x1;
        ls.wi = V(r * COS(phi), y, r * SIN(phi));
    } else {
        ls.wi = env_sample_table(S, (x0 - a) / (1.0f - a), x1);
    }
    ls.distance = MAX_DIST;
    ls.radiance = env_radiance(S, ls.wi, wl);
    ls.pdf = environment_pdf(S, ls.wi);
    ls.flags = 0; ls.light_id = -1;
    return ls;
}
/* LightQuery (glsl/core/structs.glsl:83-87; pt.ts:125-140): the receiver point, its shading
 * normal (0 at a medium event) and whether the receiver's BSDF support is the sphere and NEE
 * evaluates it (modelTwoSidedShading, materials/index.ts:56-69: the rough dielectric here). */
typedef struct { v3 p, n; int two_sided; } lquery_t;

/* light_tree_importance (accel/light_tree/light_tree.glsl:28-77) */
static float lt_importance(const float* n0, const float* n1, const lquery_t* q) {
    v3 bmin = V(n0[0], n0[1], n0[2]), bmax = V(n1[0], n1[1], n1[2]);
    v3 c = muls(add(bmin, bmax), 0.5f);
    v3 e = sub(bmax, bmin);
    v3 d = sub(c, q->p);
    float d2raw = dot(d, d);
    float diag2 = 0.25f * dot(e, e);
    float base = n0[3] / fmaxf_(d2raw, fmaxf_(diag2, 1e-8f));
    if (dot(q->n, q->n) < 0.5f) return base;                       /* no orientation */
    /* the box corner farthest along n: mix(bmin, bmax, step(0, n)) */
    v3 far_c = V(q->n.x < 0.0f ? bmin.x : bmax.x, q->n.y < 0.0f ? bmin.y : bmax.y, q->n.z < 0.0f ? bmin.z : bmax.z);
    float h = dot(q->n, sub(far_c, q->p));
    const int two = q->two_sided && !PLANT(11);   /* plant 11: the query's sphere support ignored */
    if (two) {
        v3 near_c = sub(add(bmin, bmax), far_c);
        h = fmaxf_(h, -dot(q->n, sub(near_c, q->p)));
    } else if (h <= 0.0f) {
        return 0.0f;
    }
    float cos_c = dot(q->n, d) * (1.0f / sqrtf(fmaxf_(d2raw, 1e-12f)));
    if (two) cos_c = fabsf(cos_c);
    v3 dmx = V(fmaxf_(q->p.x - bmin.x, bmax.x - q->p.x), fmaxf_(q->p.y - bmin.y, bmax.y - q->p.y),
               fmaxf_(q->p.z - bmin.z, bmax.z - q->p.z));
    float lo = h * (1.0f / sqrtf(fmaxf_(dot(dmx, dmx), 1e-12f)));
    return base * fminf_(1.0f, fmaxf_(cos_c, lo));
}
/* light_tree_pick (lighting.ts:786-800): the stochastic descent, importance-proportional child
 * choice, a fresh random number per level; the product pmf out; -1 on a dead descent. */
static int lt_pick(const or_scene* S, const lquery_t* q, float xi, rng_t* rng, float* pmf) {
    *pmf = 1.0f;
    int ni = 0;
    const float* n1 = S->lt_nodes + 4;
    for (int lvl = 0; lvl < LT_MAX_DEPTH; lvl++) {
        if (n1[3] < 0.0f) break;
        int L = ni + 1, R = (int)n1[3];
        const float *l0 = S->lt_nodes + 8 * (size_t)L, *r0 = S->lt_nodes + 8 * (size_t)R;
        float il = lt_importance(l0, l0 + 4, q), ir = lt_importance(r0, r0 + 4, q);
        float isum = il + ir;
        if (isum <= 0.0f) { *pmf = 0.0f; return -1; }
        float p_l = il / isum;
        if (xi < p_l) { *pmf *= p_l; ni = L; n1 = l0 + 4; }
        else { *pmf *= (1.0f - p_l); ni = R; n1 = r0 + 4; }
        xi = rnd(rng);
    }
    return (int)(-n1[3] - 1.0f);
}
/* light_tree_pmf (lighting.ts:802-822): the light's trail replayed with the stored query */
static float lt_pmf(const or_scene* S, const lquery_t* q, int light) {
    uint32_t tw = (uint32_t)S->lt_trails[4 * (size_t)light];
    float pmf = 1.0f;
    int ni = 0;
    const float* n1 = S->lt_nodes + 4;
    for (int lvl = 0; lvl < LT_MAX_DEPTH; lvl++) {
        if (n1[3] < 0.0f) break;
        if (lvl == 24) tw = (uint32_t)S->lt_trails[4 * (size_t)light + 1];
        int L = ni + 1, R = (int)n1[3];
        const float *l0 = S->lt_nodes + 8 * (size_t)L, *r0 = S->lt_nodes + 8 * (size_t)R;
        float il = lt_importance(l0, l0 + 4, q), ir = lt_importance(r0, r0 + 4, q);
        float isum = il + ir;
        if (isum <= 0.0f) return 0.0f;
        float p_l = il / isum;
        if (((tw & 1u) == 0u) != PLANT(12)) { pmf *= p_l; ni = L; n1 = l0 + 4; }   /* plant 12: the bits read inverted */
        else { pmf *= (1.0f - p_l); ni = R; n1 = r0 + 4; }
        tw >>= 1u;
    }
    return pmf;
}
/* the tree's finite stage (generateLightSamplingBvh, lighting.ts:834-870): pick, then the
 * light's sample at (random(), x1), its pdf times the selection pmf */
static lsample_t lighting_sample_tree(const or_scene* S, const lquery_t* q, float x0, float x1, rng_t* rng, const wl_t* wl) {
    lsample_t ls;
    float sel;
    int id = lt_pick(S, q, x0, rng, &sel);
    if (id < 0) { ls.pdf = 0.0f; return ls; }
    float lx = rnd(rng);
    ls = light_sample_one(S, &S->lights[id], q->p, lx, x1, wl);
    ls.light_id = id;
    ls.pdf *= sel;
    return ls;
}
/* lighting_sample: env-only, finite-only or two-stage; the finite stage by the selection CDF or,
 * with tree != 0, the light tree (then rng feeds its descent) */
static lsample_t lighting_sample(const or_scene* S, const lquery_t* q, float x0, float x1, const wl_t* wl, rng_t* rng, int tree) {
    lsample_t ls;
    if (S->n_lights == 0) {
        if (env_samplable(S)) return environment_sample(S, x0, x1, wl);
        ls.pdf = 0.0f; return ls;
    }
    if (!env_select_live(S))
        return tree ? lighting_sample_tree(S, q, x0, x1, rng, wl) : lighting_sample_finite(S, q->p, x0, x1, wl);
    float ps = S->env.select_prob;
    if (x0 < ps) {
        ls = environment_sample(S, clampf(x0 / ps, 0.0f, 0.9999999f), x1, wl);
        ls.pdf *= ps;
        return ls;
    }
    float fx = clampf((x0 - ps) / (1.0f - ps), 0.0f, 0.9999999f);
    ls = tree ? lighting_sample_tree(S, q, fx, x1, rng, wl) : lighting_sample_finite(S, q->p, fx, x1, wl);
    ls.pdf *= (1.0f - ps);
    return ls;
}
static int light_of(const or_scene* S, int region) {
    if (region < 0) return -1;
    for (int i = 0; i < S->n_lights; i++) if (S->lights[i].region == region) return i;
    return -1;
}
/* lighting_pdf (lighting.ts:636-673; the tree's, :909-933, replays the trail) */
static float lighting_pdf(const or_scene* S, const lquery_t* q, v3 wi, int lid, v3 light_p, int tree) {
    const pt_light* l = &S->lights[lid];
    float sel = tree ? lt_pmf(S, q, lid)
                     : (S->n_lights == 1 ? 1.0f : (PLANT(1) ? S->lights[0].select_pdf : l->select_pdf));
    if (env_select_live(S)) sel = sel * (1.0f - S->env.select_prob);
    return sel * light_pdf_one(l, q->p, light_p, wi);
}

/* ---------------- homogeneous media (transport/volume/analytic/analytic.glsl:9-51) ---- */
static inline sp sigma_s_of(const or_scene* S, const pt_material* m, const wl_t* wl) { return lift(S, m->sigma_s, &m->s_sigma_s, wl); }
static inline sp sigma_a_of(const or_scene* S, const pt_material* m, const wl_t* wl) { return lift(S, m->sigma_a, &m->s_sigma_a, wl); }
typedef struct { int scattered; float t; sp weight; sp radiance; } msample_t;
static inline sp medium_emission_of(const or_scene* S, const pt_material* m, const wl_t* wl) {
    return lift(S, m->medium_emission, &m->s_medium_emission, wl);
}
static inline int medium_emits(const pt_material* m) {
    return m->medium_emission[0] != 0.0f || m->medium_emission[1] != 0.0f || m->medium_emission[2] != 0.0f;
}
/* derivedMajorant (features/materials.ts:774-789) in the two colour models' bounds: RGB keeps the
 * ceiling (components/color/rgb/rgb.ts:5); spectral widens it to the float32 step above
 * 4 c (1 + 2^-20)^5 / (1 - 2^-20)^2 (components/color/spectral/majorant.ts). */
static void derive_majorants(pt_material* m) {
    double c = 1e-6;
    for (int k = 0; k < 3; k++) { double v = (double)m->sigma_a[k] + (double)m->sigma_s[k]; if (v > c) c = v; }
    m->majorant[0] = (float)c;
    const double ceiling = (double)(float)c;
    const double up = 1.0 + ldexp(1.0, -20), down = 1.0 - ldexp(1.0, -20);
    const double bound = 4.0 * ceiling * up * up * up * up * up / (down * down);
    float r = (float)bound;
    uint32_t u; memcpy(&u, &r, 4); u++; memcpy(&r, &u, 4);
    m->majorant[1] = r;
}
/* Frozen sampling proposal (test infrastructure: off unless a test calls or_set_proposal; the
 * reference has no such mode). For one material, the analytic medium arm draws its channel and
 * distance, and the HG phase its direction, from fixed proposal coefficients (RGB), and the
 * path is weighted by the live ones: importance sampling with a proposal that does not depend
 * on the live coefficients, unbiased for every value of them. With Russian roulette off, a
 * path's decisions then do not move when the live coefficients do, so the same-seed central
 * difference of a render is the derivative of one smooth estimator; with the proposal equal to
 * the live coefficients every value is bit-identical to the plain walk. */
static int g_prop_mat = -1;
static float g_prop_ss[3], g_prop_sa[3], g_prop_g;
void or_set_proposal(int material, const float* sigma_s, const float* sigma_a, float g) {
    g_prop_mat = material;
    for (int c = 0; c < 3 && material >= 0; c++) { g_prop_ss[c] = sigma_s[c]; g_prop_sa[c] = sigma_a[c]; }
    g_prop_g = g;
}
static inline int proposal_of(const or_scene* S, const pt_material* m) {
    return g_prop_mat >= 0 && g_prop_mat < S->n_materials && m == &S->materials[g_prop_mat];
}

static msample_t medium_sample_analytic(const or_scene* S, const pt_material* m, float t_max, float x0, float x1, const wl_t* wl) {
    msample_t ms;
    sp sa = sigma_a_of(S, m, wl), ss = sigma_s_of(S, m, wl);
    sp st = sadd(sa, ss);
    sp sq = st;                                   /* the sampling extinction */
    if (!wl->spec && proposal_of(S, m))
        sq = sadd(SP(g_prop_sa[0], g_prop_sa[1], g_prop_sa[2], 0.0f), SP(g_prop_ss[0], g_prop_ss[1], g_prop_ss[2], 0.0f));
    const int N = wl->spec ? 4 : 3;
    int c = (int)(x0 * (float)N); if (c > N - 1) c = N - 1;     /* spectrum_sample_channel */
    float sc = s_lane(sq, c);
    float t = sc == 0.0f ? t_max : -LOG(1.0f - x1) / sc;
    if (t < t_max) {
        sp tr = s_exp(smuls(sneg(st), t));
        sp trq = s_exp(smuls(sneg(sq), t));
        float pdf = s_avg(wl, smul(sq, trq));
        ms.scattered = 1; ms.t = t;
        ms.weight = sdivs(smul(ss, tr), fmaxf_(pdf, 1e-20f));
    } else {
        sp tr = s_exp(smuls(sneg(st), t_max));
        sp trq = s_exp(smuls(sneg(sq), t_max));
        float pdf = s_avg(wl, trq);
        ms.scattered = 0; ms.t = t_max;
        ms.weight = sdivs(tr, fmaxf_(pdf, 1e-20f));
    }
    ms.radiance = sgrey(0.0f);
    return ms;
}
/* Deterministic absorbing arm (features/materials.ts:708-718) with the emission source term
 * eps * integral_0^len exp(-sigma_a s) ds (transport/volume/analytic/emission.glsl:4-13). */
static inline float absorption_integral1(float sigma, float len) {
    float tau = sigma * len;
    return tau < 0.1f ? len * (1.0f + tau * (-0.5f + tau * (1.0f / 6.0f + tau * (-1.0f / 24.0f + tau / 120.0f))))
                      : (1.0f - EXP(-tau)) / sigma;
}
static msample_t medium_sample_absorb(const or_scene* S, const pt_material* m, float t_max, const wl_t* wl) {
    msample_t ms;
    sp sa = sigma_a_of(S, m, wl);
    ms.scattered = 0; ms.t = t_max;
    ms.weight = s_exp(smuls(sneg(sa), t_max));
    ms.radiance = sgrey(0.0f);
    if (medium_emits(m)) {
        sp e = medium_emission_of(S, m, wl);
        sp I = SP(absorption_integral1(sa.x, t_max), absorption_integral1(sa.y, t_max),
                  absorption_integral1(sa.z, t_max), absorption_integral1(sa.w, t_max));
        ms.radiance = smul(e, I);
    }
    return ms;
}
/* tracking_cap (delta_tracking.glsl:46-50): ceil(lambda + 6.5 sqrt(lambda) + 12). */
static inline int tracking_cap(float lambda) { return (int)ceilf(lambda + 6.5f * sqrtf(lambda) + 12.0f); }
/* Delta tracking (delta_tracking.glsl:56-120; Kutz et al. 2017 Alg. 4 with the history-aware
 * average probabilities) on a constant medium at the scalar majorant: emission collected per
 * tentative collision before the lottery, absorption a pure terminator. xi carries the first
 * distance and event draws; later ones come from the stream. */
static msample_t medium_sample_delta(const or_scene* S, const pt_material* m, float t_max, float x0, float x1,
                                     rng_t* rng, const wl_t* wl) {
    msample_t ms;
    ms.scattered = 0; ms.t = t_max; ms.weight = sgrey(1.0f); ms.radiance = sgrey(0.0f);
    const float sb = wl->spec ? m->majorant[1] : m->majorant[0];
    const sp sa = sigma_a_of(S, m, wl), ss = sigma_s_of(S, m, wl), eps = medium_emission_of(S, m, wl);
    const sp st = sadd(sa, ss);
    const sp sn = SP(fmaxf_(sb - st.x, 0.0f), fmaxf_(sb - st.y, 0.0f), fmaxf_(sb - st.z, 0.0f), fmaxf_(sb - st.w, 0.0f));
    sp w = sgrey(1.0f);
    float t = 0.0f, xd = x0, xe = x1;
    const int cap = tracking_cap(sb * t_max);
    for (int i = 0; i < cap; i++) {
        t += -LOG(1.0f - xd) / sb;
        if (t >= t_max) { ms.weight = w; return ms; }
        ms.radiance = sadd(ms.radiance, sdivs(smul(w, eps), sb));
        float pa = s_avg(wl, smul(sa, w)), ps = s_avg(wl, smul(ss, w)), pn = s_avg(wl, smul(sn, w));
        float c = pa + ps + pn;
        if (c <= 0.0f) { ms.weight = sgrey(0.0f); return ms; }
        pa /= c; ps /= c; pn /= c;
        if (xe < pa) { ms.weight = sgrey(0.0f); return ms; }
        if (xe < pa + ps) {
            ms.scattered = 1; ms.t = t;
            ms.weight = sdivs(smul(w, ss), sb * fmaxf_(ps, 1e-20f));
            return ms;
        }
        w = smul(w, sdivs(sn, sb * fmaxf_(pn, 1e-20f)));
        xd = rnd(rng); xe = rnd(rng);
    }
    ms.weight = w;
    return ms;
}
static inline sp medium_transmittance(const or_scene* S, const pt_material* m, float len, const wl_t* wl) {
    return s_exp(smuls(sneg(sadd(sigma_a_of(S, m, wl), sigma_s_of(S, m, wl))), len));
}
/* HG phase (volume_scattering/hg/hg.glsl:11-42), forward convention. */
static inline float hg_value(v3 wi, v3 wo, float g) {
    float c = dot(wi, neg(wo));
    float d = 1.0f + g * g - 2.0f * g * c;
    return (1.0f - g * g) / (4.0f * PTM_PI * d * sqrtf(fmaxf_(d, 1e-8f)));
}
/* Rayleigh phase (volume_scattering/rayleigh/rayleigh.glsl): (3 / 16 pi)(1 + c^2), c from the
 * propagation direction, the constant folded in fp32; parameter-free and grayscale. */
static inline float rayleigh_value(v3 wi, v3 wo) {
    float c = dot(wi, neg(wo));
    return (3.0f / (16.0f * PTM_PI)) * (1.0f + c * c);
}
/* The medium's phase (kernel_phase.glsl dispatch over the volume scattering models). */
static inline float phase_value(int model, v3 wi, v3 wo, float g) {
    return model == PT_PHASE_RAYLEIGH ? rayleigh_value(wi, wo) : hg_value(wi, wo, g);
}
static inline float phase_pdf(int model, v3 wi, v3 wo, float g, const wl_t* wl) { return s_avg(wl, sgrey(phase_value(model, wi, wo, g))); }
static inline void build_basis(v3 n, v3* t, v3* b) {   /* glsl/core/math.glsl:63-67 */
    v3 up = fabsf(n.y) < 0.999f ? V(0.0f, 1.0f, 0.0f) : V(1.0f, 0.0f, 0.0f);
    *t = normalize(cross(up, n));
    *b = cross(n, *t);
}
/* hg_sample (hg.glsl) and rayleigh_sample (rayleigh.glsl:15-33: the marginal (3/8)(1 + mu^2)
 * inverted by Cardano's formula for mu^3 + 3 mu + (4 - 8 xi) = 0, u = 4 xi - 2, D = sqrt(u^2 + 1),
 * mu = cbrt(u + D) - cbrt(D - u), clamped); both weight 1, the same frame around -wo. */
static v3 phase_sample(int model, v3 wo, float g, float x0, float x1, float* pdf, const wl_t* wl) {
    float cos_theta;
    if (model == PT_PHASE_RAYLEIGH) {
        float u = 4.0f * x0 - 2.0f;
        float D = sqrtf(u * u + 1.0f);
        cos_theta = POW(u + D, 1.0f / 3.0f) + (PLANT(6) ? 1.0f : -1.0f) * POW(D - u, 1.0f / 3.0f);
        cos_theta = cos_theta < -1.0f ? -1.0f : (cos_theta > 1.0f ? 1.0f : cos_theta);
    } else if (fabsf(g) < 1e-3f) cos_theta = 1.0f - 2.0f * x0;
    else {
        float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * x0);
        cos_theta = (1.0f + g * g - sq * sq) / (2.0f * g);
    }
    float sin_theta = sqrtf(fmaxf_(0.0f, 1.0f - cos_theta * cos_theta));
    float phi = PTM_TWO_PI * x1;
    v3 t, b;
    v3 nw = neg(wo);
    build_basis(nw, &t, &b);
    v3 wi = normalize(add(add(muls(t, sin_theta * COS(phi)), muls(b, sin_theta * SIN(phi))), muls(nw, cos_theta)));
    *pdf = phase_pdf(model, wi, wo, g, wl);
    return wi;
}
/* shadow_transmittance, media walker (transport/shadow/media/media.glsl:29-62). */
static sp shadow_transmittance_media(const or_scene* S, v3 o, v3 d, v3 light_p, int crossings_left, int* n_any, const wl_t* wl,
                                     int len_mat, double* len_in) {
    sp T = sgrey(1.0f);
    int medium = region_at(S, o);
    for (int crossed = 0; crossed <= crossings_left; crossed++) {
        float remaining = length3(sub(light_p, o)) - SHADOW_BACKOFF;
        if (remaining <= 0.0f) return T;
        hit_t h;
        (*n_any)++;
        int hit_boundary = scene_intersect(S, o, d, &h) && h.t < remaining;
        float seg_len = hit_boundary ? h.t : remaining;
        int med_mat = material_of(S, medium);
        if (material_has_medium(S, med_mat)) T = smul(T, medium_transmittance(S, &S->materials[med_mat], seg_len, wl));
        if (len_in && med_mat == len_mat && material_has_medium(S, med_mat)) *len_in += (double)seg_len;
        if (!hit_boundary) return T;
        if (!is_null_interface(S, material_of(S, h.region_owner))) return sgrey(0.0f);
        medium = h.region_to;
        float side = dot(d, h.ng) >= 0.0f ? 1.0f : -1.0f;
        o = add(h.p, muls(muls(h.ng, side), h.eps));
        d = normalize(sub(light_p, o));
    }
    return sgrey(0.0f);
}

/* ---------------- the walk (integrators/pt/pt.ts:299-464; combiner.ts; kernel.glsl) ---- */
typedef struct {
    int estimator, max_bounces, rr_start; float rr_max;
    int width, height, max_null;
    int spectral, wl_uniform;
    int equiangular;          /* estimator.mediumLightSampling == 'equiangular' */
    int light_tree;           /* estimator.lighting.selection == 'bvh' */
} or_program;

static inline float power_heuristic(float pf, float pg) {
    float f2 = pf * pf;
    if (f2 >= 1.0e38f) return 1.0f;
    return f2 / fmaxf_(1e-20f, f2 + pg * pg);
}
static inline v3 ray_spawn_origin(const hit_t* hit, v3 wi) {
    float side = dot(wi, hit->ng) >= 0.0f ? 1.0f : -1.0f;
    return add(hit->p, muls(muls(hit->ng, side), hit->eps));
}

/* One camera path; returns radiance. The media branch, null interfaces and the media
 * shadow walker are the program's (f.media, f.nulls) parts of pt.ts:299-464. */
/* Path-replay adjoint state (FP64): D = dlog f/dtheta of the path so far with sampling
 * decisions detached, G = the path's gradient sum_contributions dL/dI . c . D.
 * Parameter layout: [0..2] sigma_s, [3..5] sigma_a (one channel each), [6] g, [7+k] albedo k. */
#define ADJ_N (7 + PT_ADJ_MAX_ALBEDO)
typedef struct { const pt_adjoint_params* p; double delta[3]; double D[ADJ_N]; double G[ADJ_N]; } adj_t;
static inline int adj_albedo_index(const adj_t* A, int mat) {
    for (int k = 0; k < A->p->n_albedo; k++) if (A->p->albedo_material[k] == mat) return k;
    return -1;
}
/* d log hg / dg at c = dot(wi, -wo) (volume_scattering/hg/hg.glsl:11-19) */
static inline double dlog_hg(double g, double c) {
    double d = 1.0 + g * g - 2.0 * g * c;
    return -2.0 * g / (1.0 - g * g) - (PLANT(3) ? -3.0 : 3.0) * (g - c) / d;
}
static inline void adj_scatter(adj_t* A, sp c, const double* D) {
    double e[3] = {A->delta[0] * (double)c.x, A->delta[1] * (double)c.y, A->delta[2] * (double)c.z};
    for (int ch = 0; ch < 3; ch++) { A->G[ch] += e[ch] * D[ch]; A->G[3 + ch] += e[ch] * D[3 + ch]; }
    double es = e[0] + e[1] + e[2];
    for (int k = 6; k < ADJ_N; k++) A->G[k] += es * D[k];
}

/* lighting_query_delta (features/lighting.ts:306-336): the CDF chain over the delta lights. */
static float lighting_query_delta(const or_scene* S, float uc, int* id) {
    const int n = S->n_lights;
    *id = n - 1;
    for (int i = 0; i < n - 1; i++) if (uc < S->lights[i].cdf) { *id = i; break; }
    return S->lights[*id].select_pdf;
}
/* equiangular_sample_direct (techniques/equiangular/equiangular.glsl:18-57): one direct-light
 * vertex per medium segment, t distributed uniformly in the angle subtended at the light. */
static sp equiangular_sample_direct(const or_scene* S, const or_program* P, v3 o, v3 d, sp T, const pt_material* Mm,
                                    float t_max, int crossings_left, rng_t* rng, int* n_any, const wl_t* wl) {
    int id;
    float select_pdf = lighting_query_delta(S, rnd(rng), &id);
    const pt_light* l = &S->lights[id];
    v3 pos = lv(l->corner);
    sp intensity = lift(S, l->radiance, &l->s_radiance, wl);
    v3 D = sub(pos, o);
    float t_c = dot(D, d);
    float h2 = fmaxf_(dot(D, D) - t_c * t_c, 1e-8f);
    float h = sqrtf(h2);
    float theta_a = ATAN2(0.0f - t_c, h);
    float theta_b = ATAN2(t_max - t_c, h);
    float u = rnd(rng);
    float theta = theta_a * (1.0f - u) + theta_b * u;
    float t = t_c + h * TAN(theta);
    float pdf_t = h / ((theta_b - theta_a) * (h2 + (t - t_c) * (t - t_c)));
    v3 p_evt = add(o, muls(d, t));
    sp T_seg = s_exp(smuls(sneg(sadd(sigma_a_of(S, Mm, wl), sigma_s_of(S, Mm, wl))), t));
    v3 to_light = sub(pos, p_evt);
    float d2 = fmaxf_(dot(to_light, to_light), 1e-8f);
    float dist = sqrtf(d2);
    v3 wi = V(to_light.x / dist, to_light.y / dist, to_light.z / dist);
    sp vis = shadow_transmittance_media(S, p_evt, wi, pos, crossings_left, n_any, wl, -2, NULL);
    if (s_black(wl, vis)) return sgrey(0.0f);
    float f = phase_value(Mm->phase, wi, neg(d), Mm->hg_g);
    sp c = smul(smul(smul(smul(T, T_seg), sigma_s_of(S, Mm, wl)), sgrey(f)), sdivs(intensity, d2));
    c = smul(c, vis);
    return sdivs(c, pdf_t * select_pdf);
}

static sp transport_trace(const or_scene* S, const or_program* P, v3 ro, v3 rd, rng_t* rng, int* n_closest, int* n_any,
                          const wl_t* wl, adj_t* A, float* aov) {
    sp T = sgrey(1.0f), L = sgrey(0.0f);
    int prev_was_delta = 1; float prev_bsdf_pdf = 0.0f;
    lquery_t prev_q = {ro, V(0.0f, 0.0f, 0.0f), 0};   /* kernel_init's prev_query (kernel.ts:35) */
    const int tree = P->light_tree;
    float eta_scale = 1.0f;
    int nee = P->estimator != PT_EST_PT, mis = P->estimator == PT_EST_MIS;
    int emitters = S->n_lights > 0;
    const int media = S->has_media;
    int current_medium = media ? region_at(S, ro) : -1;
    int null_crossings = 0;
    int collapsed = 0;
    for (int bounce = 0; bounce <= P->max_bounces; bounce++) {
        hit_t hit;
        (*n_closest)++;
        int boundary = scene_intersect(S, ro, rd, &hit);
        if (media) {
            int med_mat = material_of(S, current_medium);
            if (material_has_medium(S, med_mat)) {
                const pt_material* Mm = &S->materials[med_mat];
                if (P->equiangular && bounce < P->max_bounces)   /* per-segment direct light */
                    L = sadd(L, equiangular_sample_direct(S, P, ro, rd, T, Mm, boundary ? hit.t : MAX_DIST,
                                                          S->has_null ? P->max_null - null_crossings : 0, rng, n_any, wl));
                float x0 = rnd(rng), x1 = rnd(rng);
                const float seg = boundary ? hit.t : MAX_DIST;
                msample_t ms = Mm->medium_arm == PT_MED_DELTA ? medium_sample_delta(S, Mm, seg, x0, x1, rng, wl)
                             : Mm->medium_arm == PT_MED_ABSORB ? medium_sample_absorb(S, Mm, seg, wl)
                             : medium_sample_analytic(S, Mm, seg, x0, x1, wl);
                /* the source term is weighted by segment-start throughput (pt.ts:333-339) */
                if (medium_emits(Mm)) L = sadd(L, smul(T, ms.radiance));
                T = smul(T, ms.weight);
                if (A && med_mat == A->p->medium_material) {   /* f = sigma_s e^{-sigma_t t} | e^{-sigma_t t_max} | e^{-sigma_a t_max} */
                    const float* ssv = Mm->sigma_s;
                    for (int ch = 0; ch < 3; ch++) {
                        /* a channel with sigma_s = 0 scatters nothing (f = 0): no finite dlog f */
                        if (Mm->medium_arm != PT_MED_ABSORB)
                            A->D[ch] += (ms.scattered && ssv[ch] > 0.0f ? (PLANT(4) ? -1.0 : 1.0) / (double)ssv[ch] : 0.0) - (double)ms.t;
                        if (!PLANT(5)) A->D[3 + ch] -= (double)ms.t;
                    }
                }
                if (ms.scattered) {
                    if (bounce == P->max_bounces) break;
                    v3 p_evt = add(ro, muls(rd, ms.t));
                    v3 wo_med = neg(rd);
                    float g = Mm->hg_g;
                    const int phase = Mm->phase;
                    /* only a Henyey-Greenstein medium has a phase parameter (g) */
                    const int g_param = A && med_mat == A->p->medium_material && phase == PT_PHASE_HG;
                    const int frozen = !wl->spec && proposal_of(S, Mm);
                    const float gq = frozen ? g_prop_g : g;     /* the phase sampling's g */
                    if (nee && !P->equiangular) {   /* light_sample_direct_medium (techniques/light/light_medium.glsl:7-18) */
                        float y0 = rnd(rng), y1 = rnd(rng);
                        lquery_t qm = {p_evt, V(0.0f, 0.0f, 0.0f), 0};   /* light_query_medium */
                        lsample_t ls = lighting_sample(S, &qm, y0, y1, wl, rng, tree);
                        if (ls.pdf > 0.0f) {
                            v3 light_p = add(p_evt, muls(ls.wi, ls.distance));
                            double sh_len = 0.0;
                            sp vis = shadow_transmittance_media(S, p_evt, ls.wi, light_p,
                                                                S->has_null ? P->max_null - null_crossings : 0, n_any, wl,
                                                                A ? A->p->medium_material : -2, A ? &sh_len : NULL);
                            if (!s_black(wl, vis)) {
                                float f = phase_value(phase, ls.wi, wo_med, g);
                                float w = 1.0f;
                                if (mis && !(ls.flags & 1)) w = power_heuristic(ls.pdf, phase_pdf(phase, ls.wi, wo_med, gq, wl));
                                sp c = smul(smul(T, ls.radiance), sgrey(f));
                                c = smul(c, vis);
                                c = smuls(c, w);
                                L = sadd(L, sdivs(c, ls.pdf));
                                if (A) {
                                    double Dl[ADJ_N];
                                    memcpy(Dl, A->D, sizeof Dl);
                                    if (g_param) Dl[6] += dlog_hg(g, dot(ls.wi, neg(wo_med)));
                                    for (int ch = 0; ch < 6; ch++) Dl[ch] -= sh_len;
                                    adj_scatter(A, sdivs(c, ls.pdf), Dl);
                                }
                            }
                        }
                    }
                    /* kernel_sample_phase (techniques/kernel/kernel_phase.glsl:7-15) */
                    float z0 = rnd(rng), z1 = rnd(rng);
                    float ppdf;
                    v3 wi = phase_sample(phase, wo_med, gq, z0, z1, &ppdf, wl);
                    if (frozen) T = smuls(T, phase_value(phase, wi, wo_med, g) / phase_value(phase, wi, wo_med, gq));
                    if (g_param) A->D[6] += dlog_hg(g, dot(wi, neg(wo_med)));
                    prev_was_delta = 0; prev_bsdf_pdf = ppdf; prev_q.p = p_evt; prev_q.n = V(0.0f, 0.0f, 0.0f); prev_q.two_sided = 0;
                    ro = p_evt; rd = wi;
                    /* roulette_interior (pt.ts:230-283), present when some medium weights absorption;
                     * medium_survival = sigma_s / sigma_t there, 1 on the tracking arm (materials.ts:731-772) */
                    if (P->rr_start >= 0 && bounce >= P->rr_start && S->weighted_absorption) {
                        sp surv = sgrey(1.0f);
                        if (Mm->medium_arm == PT_MED_ANALYTIC) {
                            sp ss = sigma_s_of(S, Mm, wl), st = sadd(ss, sigma_a_of(S, Mm, wl));
                            surv = SP(ss.x / fmaxf_(st.x, 1e-9f), ss.y / fmaxf_(st.y, 1e-9f), ss.z / fmaxf_(st.z, 1e-9f),
                                      ss.w / fmaxf_(st.w, 1e-9f));
                        }
                        float p = s_max(wl, surv);
                        if (p <= 0.0f) break;
                        if (rnd(rng) > p) break;
                        T = sdivs(T, p);
                    }
                    continue;
                }
            }
        }
        if (!boundary) {
            /* kernel_score_miss + combiner_w_env */
            float w = 1.0f;
            if (env_samplable(S) && nee) {
                if (prev_was_delta) w = 1.0f;
                else if (mis) w = power_heuristic(prev_bsdf_pdf, env_select_live(S) ? S->env.select_prob * environment_pdf(S, rd) : environment_pdf(S, rd));
                else w = 0.0f;
            }
            sp Le = env_radiance(S, rd, wl);
            L = sadd(L, smul(smuls(T, w), Le));
            if (A) adj_scatter(A, smul(smuls(T, w), Le), A->D);
            break;
        }
        if (media && hit.region_from != current_medium) current_medium = hit.region_from;   /* self-heal */
        int mat = material_of(S, hit.region_owner);
        if (S->has_null && is_null_interface(S, mat)) {
            current_medium = hit.region_to;
            float side = dot(rd, hit.ng) >= 0.0f ? 1.0f : -1.0f;
            ro = add(hit.p, muls(muls(hit.ng, side), hit.eps));
            null_crossings++;
            if (null_crossings > P->max_null) break;
            bounce--;
            continue;
        }
        const pt_material* M = &S->materials[mat];
        v3 wo = neg(rd);
        /* the first-hit guides (pt.ts:422-433): first non-null surface, no event before it; aov =
         * albedo fact clamped to [0, 1] x weight (aov/albedo/albedo.glsl), faced shading normal x
         * the weight's channel mean (aov/normal/normal.glsl) */
        if (aov && bounce == 0 && !wl->spec) {
            const float* row = M->model == PT_MAT_LAMBERT ? M->albedo
                             : (M->model == PT_MAT_GGX || M->model == PT_MAT_MIRROR) ? M->f0 : M->transmittance;
            float cell[3];
            if (M->model == PT_MAT_CHECKER) {   /* checker_albedo at the hit */
                sp c = lambert_albedo(S, M, &hit, wl);
                cell[0] = c.x; cell[1] = c.y; cell[2] = c.z;
                row = cell;
            }
            aov[0] = clampf(row[0], 0.0f, 1.0f) * T.x;
            aov[1] = clampf(row[1], 0.0f, 1.0f) * T.y;
            aov[2] = clampf(row[2], 0.0f, 1.0f) * T.z;
            v3 n = dot(hit.frame.n, wo) < 0.0f ? neg(hit.frame.n) : hit.frame.n;
            float avg = ((T.x + T.y) + T.z) * (1.0f / 3.0f);
            aov[3] = n.x * avg; aov[4] = n.y * avg; aov[5] = n.z * avg;
        }
        /* kernel_score_emitter_hit */
        int mat_emit = material_of(S, hit.region_to);
        if (mat_emit >= 0 && S->materials[mat_emit].emissive) {
            float w = 1.0f;
            if (nee && emitters) {
                int lid = light_of(S, hit.region_to);
                if (lid < 0 || prev_was_delta) w = 1.0f;
                else if (mis) w = power_heuristic(prev_bsdf_pdf, lighting_pdf(S, &prev_q, rd, lid, hit.p, tree));
                else w = 0.0f;
            }
            const pt_material* E = &S->materials[mat_emit];
            L = sadd(L, smul(smuls(T, w), lift(S, E->emission, &E->s_emission, wl)));
            if (A) adj_scatter(A, smul(smuls(T, w), lift(S, E->emission, &E->s_emission, wl)), A->D);
        }
        if (bounce == P->max_bounces) break;
        /* collapse_wavelengths (pt.ts:440,466-475): a dielectric bounding a dispersive region sends
         * the path along the hero wavelength; the hero lane carries all four (once per path) */
        if (wl->spec && !collapsed && transmissive_model(M->model) &&
            (dispersive_region(S, hit.region_from) || dispersive_region(S, hit.region_to))) {
            T = smul(T, SP(4.0f, 0.0f, 0.0f, 0.0f));
            collapsed = 1;
        }
        /* light_sample_direct (techniques/light/light.glsl:8-28) */
        if (nee && has_nondelta(M)) {
            float x0 = rnd(rng), x1 = rnd(rng);
            lquery_t qs = {hit.p, hit.frame.n, M->model == PT_MAT_ROUGH_DIELECTRIC};   /* light_query_surface */
            lsample_t ls = lighting_sample(S, &qs, x0, x1, wl, rng, tree);
            if (ls.pdf > 0.0f) {
                v3 light_p = add(hit.p, muls(ls.wi, ls.distance));
                v3 so = ray_spawn_origin(&hit, ls.wi);
                v3 sd = normalize(sub(light_p, so));
                sp vis;
                double sh_len = 0.0;
                if (media) vis = shadow_transmittance_media(S, so, sd, light_p, S->has_null ? P->max_null - null_crossings : 0, n_any, wl,
                                                            A ? A->p->medium_material : -2, A ? &sh_len : NULL);
                else {
                    float maxDist = length3(sub(light_p, so)) - SHADOW_BACKOFF;
                    (*n_any)++;
                    vis = scene_any(S, so, sd, maxDist) ? sgrey(0.0f) : sgrey(1.0f);
                }
                if (!s_black(wl, vis)) {
                    sp f = surface_eval(S, M, ls.wi, wo, &hit, wl);
                    float cos_i = fabsf(dot(ls.wi, hit.frame.n));
                    float w = 1.0f;
                    if (mis && !(ls.flags & 1)) w = power_heuristic(ls.pdf, surface_pdf(S, M, ls.wi, wo, &hit, wl));
                    sp c = smul(smul(T, ls.radiance), f);
                    c = smuls(c, cos_i);
                    c = smul(c, vis);
                    c = smuls(c, w);
                    c = sdivs(c, ls.pdf);
                    L = sadd(L, c);
                    if (A) {
                        double Dl[ADJ_N];
                        memcpy(Dl, A->D, sizeof Dl);
                        int ak = M->model == PT_MAT_LAMBERT ? adj_albedo_index(A, mat) : -1;
                        if (ak >= 0) Dl[7 + ak] += 1.0 / (double)M->albedo[0];
                        for (int ch = 0; ch < 6; ch++) Dl[ch] -= sh_len;
                        adj_scatter(A, c, Dl);
                    }
                }
            }
        }
        /* kernel_sample_continuation */
        float uc = rnd(rng); float ux = rnd(rng); float uy = rnd(rng);
        isample_t bs = surface_sample(S, M, wo, &hit, uc, ux, uy, wl);
        if (s_black(wl, bs.weight)) break;
        T = smul(T, bs.weight);
        if (A && M->model == PT_MAT_LAMBERT) {
            int ak = adj_albedo_index(A, mat);
            if (ak >= 0) A->D[7 + ak] += 1.0 / (double)M->albedo[0];
        }
        prev_was_delta = (bs.flags & LOBE_DELTA) != 0; prev_bsdf_pdf = bs.pdf;
        prev_q.p = hit.p; prev_q.n = hit.frame.n; prev_q.two_sided = M->model == PT_MAT_ROUGH_DIELECTRIC;
        if (media && (bs.flags & LOBE_TRANSMISSION)) current_medium = hit.region_to;
        if (S->has_transmission && (bs.flags & LOBE_TRANSMISSION)) {
            float r = ior_of(S, hit.region_to, wl) / ior_of(S, hit.region_from, wl);
            eta_scale *= r * r;
        }
        /* roulette (pt.ts:204-222) */
        if (P->rr_start >= 0 && bounce >= P->rr_start) {
            float m = S->has_transmission ? s_max(wl, T) * eta_scale : s_max(wl, T);
            float p = fminf_(P->rr_max, m);
            if (p <= 0.0f) break;
            if (rnd(rng) > p) break;
            T = sdivs(T, p);
        }
        ro = ray_spawn_origin(&hit, bs.wi);
        rd = bs.wi;
    }
    return L;
}

/* ---------------- ctypes API ---------------- */
or_scene* or_scene_create(const pt_scene_desc* d) {
    or_scene* S = (or_scene*)calloc(1, sizeof(or_scene));
    S->n_tris = d->n_tris;
    S->n_proto = d->proto_vertices ? d->n_proto_tris : 0;
    const size_t ntot = (size_t)d->n_tris + (size_t)S->n_proto;
    S->tri = (float*)malloc(sizeof(float) * 9 * (ntot > 0 ? ntot : 1));
    memcpy(S->tri, d->tri_vertices, sizeof(float) * 9 * (size_t)d->n_tris);
    if (S->n_proto) memcpy(S->tri + 9 * (size_t)d->n_tris, d->proto_vertices, sizeof(float) * 9 * (size_t)S->n_proto);
    if (d->tri_normals || d->proto_normals) {
        S->nrm = (float*)calloc(9 * (ntot > 0 ? ntot : 1), sizeof(float));
        if (d->tri_normals) memcpy(S->nrm, d->tri_normals, sizeof(float) * 9 * (size_t)d->n_tris);
        if (d->proto_normals && S->n_proto)
            memcpy(S->nrm + 9 * (size_t)d->n_tris, d->proto_normals, sizeof(float) * 9 * (size_t)S->n_proto);
    }
    S->tri_object = (int32_t*)malloc(sizeof(int32_t) * (size_t)(d->n_tris > 0 ? d->n_tris : 1));
    memcpy(S->tri_object, d->tri_object, sizeof(int32_t) * (size_t)d->n_tris);
    S->n_objects = d->n_objects;
    S->objects = (pt_object*)malloc(sizeof(pt_object) * (size_t)(d->n_objects > 0 ? d->n_objects : 1));
    memcpy(S->objects, d->objects, sizeof(pt_object) * (size_t)d->n_objects);
    S->n_materials = d->n_materials;
    S->materials = (pt_material*)malloc(sizeof(pt_material) * (size_t)(d->n_materials > 0 ? d->n_materials : 1));
    memcpy(S->materials, d->materials, sizeof(pt_material) * (size_t)d->n_materials);
    S->n_lights = d->n_lights;
    S->lights = (pt_light*)malloc(sizeof(pt_light) * (size_t)(d->n_lights > 0 ? d->n_lights : 1));
    memcpy(S->lights, d->lights, sizeof(pt_light) * (size_t)d->n_lights);
    S->env = d->env;
    S->ambient_material = d->ambient_material;
    if (d->spectral_curves) {
        memcpy(S->spec_curves, d->spectral_curves, sizeof S->spec_curves);
        memcpy(S->xyz_rgb, d->xyz_to_rgb, sizeof S->xyz_rgb);
        S->cie_y = d->cie_y_integral;
        S->has_spectral = 1;
    }
    if (d->env.type == PT_ENV_IMAGE && d->env_image) {
        S->img_w = d->env_image_w; S->img_h = d->env_image_h;
        S->img = (float*)malloc(sizeof(float) * 3 * (size_t)S->img_w * (size_t)S->img_h);
        memcpy(S->img, d->env_image, sizeof(float) * 3 * (size_t)S->img_w * (size_t)S->img_h);
    }
    if (d->light_boxes && d->light_powers && d->n_lights > 0) {
        S->lt_nodes = (float*)malloc(sizeof(float) * 8 * (size_t)(2 * d->n_lights - 1));
        S->lt_trails = (float*)malloc(sizeof(float) * 4 * (size_t)d->n_lights);
        or_light_tree_build(d->light_boxes, d->light_powers, d->n_lights, S->lt_nodes, S->lt_trails, NULL);
    }
    if (d->spectral_atlas && d->spectral_z_nodes) {
        S->atlas = (float*)malloc(sizeof(float) * 4 * 3 * 64 * 64 * 64);
        memcpy(S->atlas, d->spectral_atlas, sizeof(float) * 4 * 3 * 64 * 64 * 64);
        memcpy(S->atlas_z, d->spectral_z_nodes, sizeof S->atlas_z);
    }
    for (int m = 0; m < S->n_materials; m++) {
        if (S->materials[m].has_medium) S->has_media = 1;
        if (S->materials[m].model == PT_MAT_NULL) S->has_null = 1;
        if (S->materials[m].has_medium && S->materials[m].medium_arm == PT_MED_ANALYTIC) S->weighted_absorption = 1;
        derive_majorants(&S->materials[m]);
    }
    S->env_rot_cs[0] = (float)cos((double)d->env.rotation);
    S->env_rot_cs[1] = (float)sin((double)d->env.rotation);
    for (int m = 0; m < S->n_materials; m++) if (transmissive_model(S->materials[m].model)) S->has_transmission = 1;
    /* per-mesh BLAS; instance batches get theirs over the prototype range */
    S->blas = (bvh_t*)calloc((size_t)(S->n_objects > 0 ? S->n_objects : 1), sizeof(bvh_t));
    S->obj_first = (int*)calloc((size_t)(S->n_objects > 0 ? S->n_objects : 1), sizeof(int));
    for (int o = 0; o < S->n_objects; o++)
        S->obj_first[o] = S->objects[o].kind == PT_OBJ_INSTANCES ? d->n_tris + S->objects[o].first_tri : S->objects[o].first_tri;
    int nm = 0;
    for (int o = 0; o < S->n_objects; o++) {
        const pt_object* ob = &S->objects[o];
        if (ob->kind != PT_OBJ_MESH && ob->kind != PT_OBJ_INSTANCES) continue;
        if (ob->kind == PT_OBJ_MESH) nm++;
        int n = ob->tri_count;
        double* boxes = (double*)malloc(sizeof(double) * 6 * (size_t)(n > 0 ? n : 1));
        for (int i = 0; i < n; i++) {
            const float* p = S->tri + 9 * (size_t)(S->obj_first[o] + i);
            double ext = 0;
            for (int a = 0; a < 3; a++) {
                double mn = fmin(fmin(p[a], p[3 + a]), p[6 + a]), mx = fmax(fmax(p[a], p[3 + a]), p[6 + a]);
                boxes[6 * i + a] = mn; boxes[6 * i + 3 + a] = mx;
                if (mx - mn > ext) ext = mx - mn;
            }
            /* Conservative triangle box: Moller-Trumbore accepts points up to 1e-6 (barycentric)
             * outside the triangle, so an exact box can cull an accepted hit; widen by 4e-6 of
             * the extent (and 4e-7 relative to the coordinate) so the walk equals brute force. */
            for (int a = 0; a < 3; a++) {
                double m = 4e-6 * ext + 1e-30;
                boxes[6 * i + a] -= fmax(m, 4e-7 * fabs(boxes[6 * i + a]));
                boxes[6 * i + 3 + a] += fmax(m, 4e-7 * fabs(boxes[6 * i + 3 + a]));
            }
        }
        bvh_build(&S->blas[o], boxes, n, 2);
        free(boxes);
    }
    S->n_mesh_objects = nm;
    double* tb = (double*)malloc(sizeof(double) * 6 * (size_t)(nm > 0 ? nm : 1));
    S->tlas_obj = (int*)malloc(sizeof(int) * (size_t)(nm > 0 ? nm : 1));
    int k = 0;
    for (int o = 0; o < S->n_objects; o++) {
        if (S->objects[o].kind != PT_OBJ_MESH || S->blas[o].n_nodes == 0) continue;
        const float* n = S->blas[o].nodes;
        for (int a = 0; a < 3; a++) { tb[6 * k + a] = n[a]; tb[6 * k + 3 + a] = n[4 + a]; }
        S->tlas_obj[k++] = o;
    }
    bvh_build(&S->tlas, tb, k, 1);
    free(tb);
    /* placements: world boxes of the prototype's local box (f64, padded), instance TLAS */
    S->inst_prim0 = d->n_tris + d->n_objects;
    S->n_inst = d->instances ? d->n_instances : 0;
    if (S->n_inst > 0) {
        S->inst = (float*)malloc(sizeof(float) * 8 * (size_t)S->n_inst);
        S->inst_batch = (int*)malloc(sizeof(int) * (size_t)S->n_inst);
        S->inst_base = (int*)malloc(sizeof(int) * (size_t)S->n_inst);
        for (int i = 0; i < S->n_inst; i++) {
            memcpy(S->inst + 8 * (size_t)i, d->instances[i].q, 4 * sizeof(float));
            memcpy(S->inst + 8 * (size_t)i + 4, d->instances[i].ts, 4 * sizeof(float));
            S->inst_batch[i] = -1;
        }
        for (int o = 0; o < S->n_objects; o++) {
            const pt_object* ob = &S->objects[o];
            if (ob->kind != PT_OBJ_INSTANCES) continue;
            for (int i = ob->first_instance; i < ob->first_instance + ob->instance_count; i++) S->inst_batch[i] = o;
        }
        int acc = 0;
        for (int i = 0; i < S->n_inst; i++) { S->inst_base[i] = acc; acc += S->objects[S->inst_batch[i]].tri_count; }
        double* ib = (double*)malloc(sizeof(double) * 6 * (size_t)S->n_inst);
        for (int i = 0; i < S->n_inst; i++) {
            const float* I = S->inst + 8 * (size_t)i;
            const pt_object* ob = &S->objects[S->inst_batch[i]];
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            double qx = -I[0], qy = -I[1], qz = -I[2], qw = I[3], sc = I[7];
            for (int c = 0; c < 8; c++) {
                double v[3];
                for (int a = 0; a < 3; a++) v[a] = sc * ((c >> a) & 1 ? ob->box_max[a] : ob->box_min[a]) - I[4 + a];
                /* forward rotation conj(q): v + 2 q x (q x v + w v) */
                double c1[3] = {qy * v[2] - qz * v[1] + qw * v[0], qz * v[0] - qx * v[2] + qw * v[1], qx * v[1] - qy * v[0] + qw * v[2]};
                double w[3] = {v[0] + 2.0 * (qy * c1[2] - qz * c1[1]), v[1] + 2.0 * (qz * c1[0] - qx * c1[2]), v[2] + 2.0 * (qx * c1[1] - qy * c1[0])};
                for (int a = 0; a < 3; a++) { if (w[a] < lo[a]) lo[a] = w[a]; if (w[a] > hi[a]) hi[a] = w[a]; }
            }
            double ext = fmax(hi[0] - lo[0], fmax(hi[1] - lo[1], hi[2] - lo[2]));
            for (int a = 0; a < 3; a++) {
                double m = 1e-5 * ext + 1e-30;
                ib[6 * i + a] = lo[a] - fmax(m, 1e-6 * fabs(lo[a]));
                ib[6 * i + 3 + a] = hi[a] + fmax(m, 1e-6 * fabs(hi[a]));
            }
        }
        bvh_build(&S->itlas, ib, S->n_inst, 1);
        free(ib);
    }
    if (S->env.type == PT_ENV_SKY) env_build(S);
    if (S->env.type == PT_ENV_IMAGE) env_build_image(S);
    return S;
}
void or_scene_destroy(or_scene* S) {
    if (!S) return;
    for (int o = 0; o < S->n_objects; o++) bvh_free(&S->blas[o]);
    free(S->blas); bvh_free(&S->tlas); free(S->tlas_obj);
    free(S->obj_first); bvh_free(&S->itlas); free(S->inst); free(S->inst_batch); free(S->inst_base);
    free(S->tri); free(S->nrm); free(S->tri_object); free(S->objects); free(S->materials); free(S->lights);
    free(S->env_cond); free(S->env_marg); free(S->atlas); free(S->lt_nodes); free(S->lt_trails); free(S->img);
    free(S);
}
void or_env_table(const or_scene* S, float* cond, float* marg) {
    if (S->env_cond) { memcpy(cond, S->env_cond, sizeof(float) * (size_t)S->env_w * S->env_h); memcpy(marg, S->env_marg, sizeof(float) * (size_t)S->env_h); }
}
double or_env_total(const or_scene* S) { return S->env_total; }

void or_trace_closest(const or_scene* S, const float* rays, int n, float* t_out, int32_t* prim_out) {
    /* rays: 7 arrays of n (ox, oy, oz, dx, dy, dz, tmax) */
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; i++) {
        v3 ro = V(rays[i], rays[n + i], rays[2 * n + i]);
        v3 rd = V(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        float tmax = rays[6 * n + i];
        int p = scene_closest(S, ro, rd, &tmax);
        t_out[i] = p >= 0 ? tmax : rays[6 * n + i];
        prim_out[i] = p;
    }
}
void or_trace_any(const or_scene* S, const float* rays, int n, uint8_t* hit_out) {
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; i++) {
        v3 ro = V(rays[i], rays[n + i], rays[2 * n + i]);
        v3 rd = V(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        hit_out[i] = (uint8_t)scene_any(S, ro, rd, rays[6 * n + i]);
    }
}
/* Brute-force closest hit over every triangle (pins the BVH walks). */
void or_trace_closest_brute(const or_scene* S, const float* rays, int n, float* t_out, int32_t* prim_out) {
    #pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; i++) {
        v3 ro = V(rays[i], rays[n + i], rays[2 * n + i]);
        v3 rd = V(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        float tmax = rays[6 * n + i]; int best = -1;
        for (int tri = 0; tri < S->n_tris; tri++) {
            v3 a, b, c, g; float t;
            tri_load(S, tri, &a, &b, &c);
            if (tri_test(ro, rd, a, b, c, &t, &g) && t > 0.0f && (t < tmax || (t == tmax && tri < best))) { tmax = t; best = tri; }
        }
        for (int o = 0; o < S->n_objects; o++) {
            const pt_object* ob = &S->objects[o]; float t; int ok = 0;
            if (ob->kind == PT_OBJ_QUAD) ok = quad_intersect(ro, rd, ob->p, &t);
            else if (ob->kind == PT_OBJ_SPHERE) ok = sphere_intersect(ro, rd, ld3(ob->p), ob->p[3], &t);
            if (ok && t < tmax) { tmax = t; best = S->n_tris + o; }
        }
        for (int k = 0; k < S->n_inst; k++) {   /* every placement, every prototype triangle */
            v3 lo, ld; inst_ray(S->inst + 8 * (size_t)k, ro, rd, &lo, &ld);
            const pt_object* ob = &S->objects[S->inst_batch[k]];
            for (int j = 0; j < ob->tri_count; j++) {
                v3 a, b, c, g; float t;
                tri_load(S, S->obj_first[S->inst_batch[k]] + j, &a, &b, &c);
                int prim = S->inst_prim0 + S->inst_base[k] + j;
                if (tri_test(lo, ld, a, b, c, &t, &g) && t > 0.0f && (t < tmax || (t == tmax && prim < best))) { tmax = t; best = prim; }
            }
        }
        t_out[i] = tmax; prim_out[i] = best;
    }
}
int or_region_at(const or_scene* S, const float* p) { return region_at(S, V(p[0], p[1], p[2])); }
/* concentric_disk over n square samples u[2n] -> out[2n] (pins the lens map). */
static void concentric_disk(float u0, float u1, float* dx, float* dy);
void or_concentric_disk(const float* u, int n, float* out) {
    for (int i = 0; i < n; i++) concentric_disk(u[2 * i], u[2 * i + 1], &out[2 * i], &out[2 * i + 1]);
}

/* camera_generateRay: pinhole (camera/pinhole/pinhole.glsl:4-18) or orthographic
 * (camera/orthographic/orthographic.glsl:6-21: the origin slides over the film plane of
 * half-height u_orthoScale, every ray looks along forward). ndcx already carries the aspect. */
/* concentric_disk (glsl/core/math.glsl:71-78): Shirley-Chiu square-to-disk. */
static void concentric_disk(float u0, float u1, float* dx, float* dy) {
    float ox = 2.0f * u0 - 1.0f, oy = 2.0f * u1 - 1.0f;
    if (ox == 0.0f && oy == 0.0f) { *dx = 0.0f; *dy = 0.0f; return; }
    float r, th;
    if (fabsf(ox) > fabsf(oy)) { r = ox; th = (PTM_PI / 4.0f) * (oy / ox); }
    else { r = oy; th = PTM_PI / 2.0f - (PTM_PI / 4.0f) * (ox / oy); }
    *dx = r * ptm_cos(th); *dy = r * ptm_sin(th);
}
/* fisheye.glsl FISHEYE_THETA */
static float fisheye_theta(int projection, float rho, float K) {
    if (projection == PT_FISHEYE_EQUISOLID) return 2.0f * ptm_asin(clampf(rho * K, 0.0f, 1.0f));
    if (projection == PT_FISHEYE_STEREOGRAPHIC) return 2.0f * ptm_atan(rho * K);
    if (projection == PT_FISHEYE_ORTHOGRAPHIC) return ptm_asin(clampf(rho * K, 0.0f, 1.0f));
    return rho * K;
}
static void camera_ray(const pt_params* par, v3 O, v3 F, v3 R, v3 U, float ndcx, float ndcy, float l0, float l1,
                       float filmx, float filmy, int W, int H, v3* org, v3* dir) {
    if (par->camera == PT_CAM_ORTHOGRAPHIC) {
        *org = add(add(O, muls(R, ndcx * par->ortho_scale)), muls(U, ndcy * par->ortho_scale));
        *dir = F;
    } else if (par->camera == PT_CAM_FISHEYE) {     /* camera/fisheye/fisheye.glsl:23-38 */
        *org = O;
        float rho = sqrtf(ndcx * ndcx + ndcy * ndcy);
        if (rho < 1e-6f) { *dir = F; return; }
        float th = fisheye_theta(par->projection, rho, par->fisheye_k);
        float tx = ndcx / rho, ty = ndcy / rho;
        *dir = normalize(add(muls(F, ptm_cos(th)), muls(add(muls(R, tx), muls(U, ty)), ptm_sin(th))));
    } else if (par->camera == PT_CAM_CYLINDRICAL) {   /* camera/cylindrical/cylindrical.glsl:9-21 */
        *org = O;
        float cx = filmx - 0.5f * (float)W, cy = filmy - 0.5f * (float)H;
        float phi = cx / par->cyl_focal, hh = cy / par->cyl_focal;
        *dir = normalize(add(add(muls(F, ptm_cos(phi)), muls(R, ptm_sin(phi))), muls(U, hh)));
    } else if (par->camera == PT_CAM_EQUIRECT) {      /* camera/equirect/equirect.glsl:9-22 */
        *org = O;
        float u = filmx / (float)W, v = filmy / (float)H;
        float phi = (u - 0.5f) * PTM_TWO_PI, th = (v - 0.5f) * PTM_PI;
        float ct = ptm_cos(th);
        *dir = normalize(add(muls(add(muls(F, ptm_cos(phi)), muls(R, ptm_sin(phi))), ct), muls(U, ptm_sin(th))));
    } else {
        *org = O;
        *dir = normalize(add(add(F, muls(R, ndcx * par->tan_fov)), muls(U, ndcy * par->tan_fov)));
        if (par->camera == PT_CAM_THINLENS) {
            /* thinlens/thinlens.glsl:6-33: the pinhole ray's point on the focus plane, seen from
             * a point of the lens disk */
            v3 focus = add(O, muls(*dir, par->focus_distance / dot(*dir, F)));
            float lx, ly; concentric_disk(l0, l1, &lx, &ly);
            *org = add(add(O, muls(R, lx * par->aperture)), muls(U, ly * par->aperture));
            *dir = normalize(sub(focus, *org));
        }
    }
}

/* Render nframes frames starting at sample_count into film (RGBA32F, row 0 = bottom),
 * following features/accumulation.ts:14-51 and accumulator/average/average.glsl:6-10.
 * Optional per-sample radiance output `samples` [nframes][H][W][3] (may be NULL).
 * Rows [y0, y1) only (a bounded sample); counts of closest/any queries are returned. */
/* average_update (components/accumulator/average): the running mean of the samples in order */
static void avg_update(float* px, uint32_t n, float x, float y, float z) {
    if (n == 0) { px[0] = x; px[1] = y; px[2] = z; }
    else {
        float w = 1.0f / ((float)n + 1.0f);
        px[0] = px[0] * (1.0f - w) + x * w;
        px[1] = px[1] * (1.0f - w) + y * w;
        px[2] = px[2] * (1.0f - w) + z * w;
    }
    px[3] = 1.0f;
}
static void render_impl(const or_scene* S, const pt_program* prog, const pt_params* par, uint32_t sample_count,
                        int nframes, int y0, int y1, float* film, float* samples, uint64_t* counts, float* aov_a,
                        float* aov_n);
void or_render(const or_scene* S, const pt_program* prog, const pt_params* par, uint32_t sample_count,
               int nframes, int y0, int y1, float* film, float* samples, uint64_t* counts) {
    render_impl(S, prog, par, sample_count, nframes, y0, y1, film, samples, counts, NULL, NULL);
}
/* The same render with the first-hit guide films (either may be NULL). */
void or_render_aovs(const or_scene* S, const pt_program* prog, const pt_params* par, uint32_t sample_count,
                    int nframes, float* film, float* aov_albedo, float* aov_normal, uint64_t* counts) {
    render_impl(S, prog, par, sample_count, nframes, 0, prog->height, film, NULL, counts, aov_albedo, aov_normal);
}
static void render_impl(const or_scene* S, const pt_program* prog, const pt_params* par, uint32_t sample_count,
                        int nframes, int y0, int y1, float* film, float* samples, uint64_t* counts, float* aov_a,
                        float* aov_n) {
    or_program P;
    P.estimator = prog->estimator; P.max_bounces = prog->max_bounces; P.rr_start = prog->rr_start;
    P.rr_max = prog->rr_max_survival; P.width = prog->width; P.height = prog->height;
    P.max_null = prog->max_null_crossings;
    P.spectral = prog->color == PT_COLOR_SPECTRAL; P.wl_uniform = prog->wavelengths == PT_WL_UNIFORM;
    P.equiangular = prog->medium_light_sampling == PT_MLS_EQUIANGULAR;
    P.light_tree = prog->light_selection == PT_SEL_BVH && S->lt_nodes != NULL;
    if (P.spectral && !S->has_spectral) { if (counts) { counts[0] = counts[1] = 0; } return; }
    int W = P.width, H = P.height;
    uint64_t nc = 0, na = 0;
    v3 F = lv(par->cam_forward), R = lv(par->cam_right), U = lv(par->cam_up), O = lv(par->cam_position);
    #pragma omp parallel for schedule(dynamic, 1) reduction(+:nc, na)
    for (int y = y0; y < y1; y++) {
        for (int x = 0; x < W; x++) {
            int c1 = 0, c2 = 0;
            float* px = film + 4 * ((size_t)y * W + x);
            for (int f = 0; f < nframes; f++) {
                uint32_t n = sample_count + (uint32_t)f;
                float fx = (float)x + 0.5f + (float)par->pixel_offset[0];
                float fy = (float)y + 0.5f + (float)par->pixel_offset[1];
                rng_t rng; rng_init(&rng, (uint32_t)fx, (uint32_t)fy, n, par->reset_salt);
                float j0 = rnd(&rng), j1 = rnd(&rng);
                float filmx = fx + (j0 - 0.5f), filmy = fy + (j1 - 0.5f);
                float l0 = rnd(&rng), l1 = rnd(&rng);   /* lens sample, consumed by every camera */
                wl_t wl; wl.spec = P.spectral; wl.uniform = P.wl_uniform;
                wl.lam[0] = wl.lam[1] = wl.lam[2] = wl.lam[3] = 0.0f;
                if (P.spectral) {                   /* wavelengths_sample (color/spectral/spectral.glsl:1-10) */
                    float u = rnd(&rng);
                    for (int k = 0; k < 4; k++) wl.lam[k] = wavelength_from_cdf(P.wl_uniform, ptm_fract(u + 0.25f * (float)k));
                }
                float ndcx = (2.0f * filmx / (float)W) - 1.0f;
                float ndcy = (2.0f * filmy / (float)H) - 1.0f;
                ndcx *= par->aspect;
                v3 org, dir;
                camera_ray(par, O, F, R, U, ndcx, ndcy, l0, l1, filmx, filmy, W, H, &org, &dir);
                float aov[6] = {0, 0, 0, 0, 0, 0};
                sp Ls = transport_trace(S, &P, org, dir, &rng, &c1, &c2, &wl, NULL, (aov_a || aov_n) ? aov : NULL);
                v3 Lr = V(Ls.x, Ls.y, Ls.z);
                if (P.spectral) {                   /* color_to_pixel (color/spectral/spectral.glsl:27-33) */
                    v3 pix = V(0.0f, 0.0f, 0.0f);
                    for (int k = 0; k < 4; k++)
                        pix = add(pix, muls(sensor_response(S, wl.lam[k]), s_lane(Ls, k) / wavelength_pdf(P.wl_uniform, wl.lam[k])));
                    Lr = muls(pix, 0.25f);
                }
                if (samples) {
                    float* sp = samples + 3 * (((size_t)f * H + y) * W + x);
                    sp[0] = Lr.x; sp[1] = Lr.y; sp[2] = Lr.z;
                }
                /* average_update, or oneshot_update (accumulator/oneshot/oneshot.glsl): the latest sample */
                const uint32_t na_ = prog->accumulation == PT_ACC_ONESHOT ? 0u : n;
                avg_update(px, na_, Lr.x, Lr.y, Lr.z);
                if (aov_a) avg_update(aov_a + 4 * ((size_t)y * W + x), na_, aov[0], aov[1], aov[2]);
                if (aov_n) avg_update(aov_n + 4 * ((size_t)y * W + x), na_, aov[3], aov[4], aov[5]);
            }
            nc += (uint64_t)c1; na += (uint64_t)c2;
        }
    }
    if (counts) { counts[0] = nc; counts[1] = na; }
}

/* Declared-arithmetic exports, so tests can pin them against correctly rounded math. */
/* The device's wide-BVH slab test (pathtracerglsl_amd/csrc/traversal.h trace_step) restated in
 * fp32 for the bound check (test infrastructure; the oracle's own walks use the exact binary
 * boxes): for ray k and a quantized child box (origin p, exponents e, planes qlo / qhi per
 * axis), out[2k] = the entry distance max(near planes, 0) and out[2k+1] = the exit distance
 * min(far planes), each plane q * dq + (oq -+ err) with dq = 2^e / d, oq = (p - o) / d and
 * err = 6u (|oq| + 255 |dq|) + 1e-30. rays: o.xyz, d.xyz per ray; boxes: p.xyz, e.xyz (as floats),
 * qlo.xyz, qhi.xyz per ray. */
void or_slab_probe(const float* rays, const float* boxes, int n, float* out) {
    const float REL = 6.0f * 5.9604644775390625e-8f, ABS = 1e-30f;
    for (int k = 0; k < n; k++) {
        const float* r = rays + 6 * (size_t)k;
        const float* b = boxes + 12 * (size_t)k;
        float te = 0.0f, tf = 3.0e38f;
        for (int a = 0; a < 3; a++) {
            float d = r[3 + a];
            d = fabsf(d) < 1e-20f ? copysignf(1e-20f, d) : d;
            const float inv = 1.0f / d;
            const float s = ldexpf(1.0f, (int)b[3 + a]);
            const float dq = s * inv, oq = (b[a] - r[a]) * inv;
            const float e = fmaf(255.0f, fabsf(dq), fabsf(oq)) * REL + ABS;
            const float qn = inv < 0.0f ? b[9 + a] : b[6 + a], qf = inv < 0.0f ? b[6 + a] : b[9 + a];
            const float tn = fmaf(qn, dq, oq - e), tfa = fmaf(qf, dq, oq + e);
            te = fmaxf(te, tn);
            tf = fminf(tf, tfa);
        }
        out[2 * (size_t)k] = te;
        out[2 * (size_t)k + 1] = tf;
    }
}

void or_math_eval(int fn, const float* x, const float* y, float* out, int n) {
    for (int i = 0; i < n; i++) {
        switch (fn) {
            case 0: out[i] = ptm_sin(x[i]); break;
            case 1: out[i] = ptm_cos(x[i]); break;
            case 2: out[i] = ptm_exp(x[i]); break;
            case 3: out[i] = ptm_log(x[i]); break;
            case 4: out[i] = ptm_atan2(y[i], x[i]); break;
            case 5: out[i] = ptm_tan(x[i]); break;
            case 6: out[i] = ptm_pow(x[i], y[i]); break;
            case 7: out[i] = ptm_atan(x[i]); break;
            case 8: out[i] = ptm_acos(x[i]); break;
            case 9: out[i] = ptm_tanh(x[i]); break;
            case 10: out[i] = ptm_atanh(x[i]); break;
            case 11: out[i] = ptm_asin(x[i]); break;
            default: out[i] = NAN;
        }
    }
}

/* Basic-op companions of or_math_eval for the device comparison. */
void or_basic_eval(int fn, const float* x, const float* y, float* out, int n) {
    for (int i = 0; i < n; i++) {
        float a = x[i], b = y[i];
        switch (fn) {
            case 20: out[i] = a / b; break;
            case 21: case 23: out[i] = sqrtf(a); break;
            case 22: out[i] = 1.0f / sqrtf(a); break;
            case 24: out[i] = 1.0f / a; break;
            case 25: out[i] = a * b + a; break;
            case 26: out[i] = a > b ? a : b; break;
            default: out[i] = NAN;
        }
    }
}

/* BSDF probe: sample the material at a hit with shading normal n (frame = ambient_frame(n)),
 * then evaluate f and pdf at the sampled direction. out[12]: wi.xyz, weight.xyz, pdf, flags,
 * f.xyz (eval at wi), pdf(wi). Region ior: from = -1 (air), to = the material itself. */
void or_bsdf_probe(const pt_material* m, const float* n, const float* wo, float uc, float ux, float uy, float* out) {
    or_scene S; memset(&S, 0, sizeof S);
    pt_object ob; memset(&ob, 0, sizeof ob); ob.kind = PT_OBJ_SPHERE; ob.material = 0; ob.closed = 1;
    S.objects = &ob; S.n_objects = 1; S.materials = (pt_material*)m; S.n_materials = 1;
    hit_t h; memset(&h, 0, sizeof h);
    h.frame = ambient_frame(V(0, 0, 0), V(n[0], n[1], n[2]));
    h.ng = h.frame.n; h.region_from = -1; h.region_to = 0;
    v3 w = V(wo[0], wo[1], wo[2]);
    wl_t wl; memset(&wl, 0, sizeof wl);
    isample_t s = surface_sample(&S, m, w, &h, uc, ux, uy, &wl);
    sp f = surface_eval(&S, m, s.wi, w, &h, &wl);
    float p = surface_pdf(&S, m, s.wi, w, &h, &wl);
    float r[12] = {s.wi.x, s.wi.y, s.wi.z, s.weight.x, s.weight.y, s.weight.z, s.pdf, (float)s.flags, f.x, f.y, f.z, p};
    memcpy(out, r, sizeof r);
}
/* The octahedral chart's two maps (ptm_oct_dir / ptm_oct_uv) over n points: inverse = 0 maps
 * uv [n][2] to directions [n][3]; inverse = 1 maps directions [n][3] to uv [n][2]. (c, s) is the
 * rotation's cosine and sine. */
void or_oct_map(const float* in, int n, int inverse, float c, float s, float* out) {
    for (int i = 0; i < n; i++) {
        if (inverse) ptm_oct_uv(in[3 * i], in[3 * i + 1], in[3 * i + 2], c, s, &out[2 * i], &out[2 * i + 1]);
        else ptm_oct_dir(in[2 * i], in[2 * i + 1], c, s, &out[3 * i], &out[3 * i + 1], &out[3 * i + 2]);
    }
}
/* Phase-function probes (RGB wavelengths): values at direction pairs (wi, wo: 6 floats each),
 * and samples about wo at xi (2 floats each) with their pdfs. model is a PT_PHASE_* value. */
void or_phase_value(int model, float g, const float* dirs, int n, float* out) {
    for (int i = 0; i < n; i++) {
        const float* d = dirs + 6 * (size_t)i;
        out[i] = phase_value(model, V(d[0], d[1], d[2]), V(d[3], d[4], d[5]), g);
    }
}
void or_phase_sample(int model, float g, const float* wo, const float* xi, int n, float* wi, float* pdf) {
    wl_t wl0; memset(&wl0, 0, sizeof wl0);
    for (int i = 0; i < n; i++) {
        v3 w = phase_sample(model, V(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), g, xi[2 * i], xi[2 * i + 1], &pdf[i], &wl0);
        wi[3 * i] = w.x; wi[3 * i + 1] = w.y; wi[3 * i + 2] = w.z;
    }
}
/* Spectral records at n wavelengths (one lane each), through the kernels' evaluation. */
void or_spectrum_eval(const or_scene* S, const pt_spectrum* r, const float* lam, int n, float* out) {
    for (int i = 0; i < n; i++) {
        wl_t wl; memset(&wl, 0, sizeof wl);
        wl.spec = 1;
        for (int k = 0; k < 4; k++) wl.lam[k] = lam[i];
        sp v = spectrum_eval(S, r, &wl);
        out[i] = v.x;
    }
}

/* The environment's radiance at directions: RGB (lam null; out [n][3]) or the spectral lift at
 * one wavelength per direction (out [n]). */
void or_env_radiance(const or_scene* S, const float* dirs, const float* lam, int n, float* out) {
    for (int i = 0; i < n; i++) {
        wl_t wl; memset(&wl, 0, sizeof wl);
        wl.spec = lam != NULL;
        for (int k = 0; k < 4; k++) wl.lam[k] = lam ? lam[i] : 0.0f;
        sp v = env_radiance(S, V(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]), &wl);
        if (lam) out[i] = v.x;
        else { out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z; }
    }
}
/* Light-tree probes over a built tree (nodes, trails as or_light_tree_build writes them). */
static or_scene* lt_probe_scene(const float* nodes, const float* trails) {
    static or_scene T;
    memset(&T, 0, sizeof T);
    T.lt_nodes = (float*)nodes; T.lt_trails = (float*)trails;
    return &T;
}
float or_light_tree_importance(const float* n0, const float* n1, const float* p, const float* n, int two_sided) {
    lquery_t q = {V(p[0], p[1], p[2]), V(n[0], n[1], n[2]), two_sided};
    return lt_importance(n0, n1, &q);
}
float or_light_tree_pmf(const float* nodes, const float* trails, int light, const float* p, const float* n, int two_sided) {
    lquery_t q = {V(p[0], p[1], p[2]), V(n[0], n[1], n[2]), two_sided};
    return lt_pmf(lt_probe_scene(nodes, trails), &q, light);
}
/* the descent from xi with the pixel stream (px, py, sample, salt) feeding the later levels;
 * returns the light (-1 dead) and its product pmf in *pmf */
int or_light_tree_pick(const float* nodes, float xi, const float* p, const float* n, int two_sided, uint32_t px, uint32_t py,
                       uint32_t sample, uint32_t salt, float* pmf) {
    lquery_t q = {V(p[0], p[1], p[2]), V(n[0], n[1], n[2]), two_sided};
    rng_t r;
    rng_init(&r, px, py, sample, salt);
    return lt_pick(lt_probe_scene(nodes, NULL), &q, xi, &r, pmf);
}
/* Environment probes: pdf of directions, and samples with their pdfs. */
void or_env_pdf(const or_scene* S, const float* dirs, int n, float* out) {
    for (int i = 0; i < n; i++) out[i] = environment_pdf(S, V(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
}
void or_env_sample(const or_scene* S, const float* xi, int n, float* dirs, float* pdf) {
    for (int i = 0; i < n; i++) {
        wl_t wl0; memset(&wl0, 0, sizeof wl0);
        lsample_t ls = environment_sample(S, xi[2 * i], xi[2 * i + 1], &wl0);
        dirs[3 * i] = ls.wi.x; dirs[3 * i + 1] = ls.wi.y; dirs[3 * i + 2] = ls.wi.z; pdf[i] = ls.pdf;
    }
}

/* FP64 replay of the path-replay adjoint: the forward paths of or_render (identical fp32 walk,
 * sampleCount 0..nframes-1), with D and G accumulated in double. dL [H][W][4]; grad
 * [PT_ADJ_N_PARAMS]. Rows are summed in order, so the result is deterministic. */
void or_adjoint(const or_scene* S, const pt_program* prog, const pt_params* par, int nframes,
                const pt_adjoint_params* adj, const float* dL, double* grad) {
    or_program P;
    P.estimator = prog->estimator; P.max_bounces = prog->max_bounces; P.rr_start = prog->rr_start;
    P.rr_max = prog->rr_max_survival; P.width = prog->width; P.height = prog->height;
    P.max_null = prog->max_null_crossings; P.spectral = 0; P.wl_uniform = 0; P.equiangular = 0; P.light_tree = 0;
    const int W = P.width, H = P.height;
    double* rows = (double*)calloc((size_t)H * ADJ_N, sizeof(double));
    v3 F = lv(par->cam_forward), R = lv(par->cam_right), U = lv(par->cam_up), O = lv(par->cam_position);
    #pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < H; y++) {
        double* gr = rows + (size_t)y * ADJ_N;
        for (int x = 0; x < W; x++) {
            adj_t A; memset(&A, 0, sizeof A); A.p = adj;
            const float* d = dL + 4 * ((size_t)y * W + x);
            for (int c = 0; c < 3; c++) A.delta[c] = (double)d[c] / (double)nframes;
            for (int f = 0; f < nframes; f++) {
                int c1 = 0, c2 = 0;
                float fx = (float)x + 0.5f + (float)par->pixel_offset[0];
                float fy = (float)y + 0.5f + (float)par->pixel_offset[1];
                rng_t rng; rng_init(&rng, (uint32_t)fx, (uint32_t)fy, (uint32_t)f, par->reset_salt);
                float j0 = rnd(&rng), j1 = rnd(&rng);
                float filmx = fx + (j0 - 0.5f), filmy = fy + (j1 - 0.5f);
                float l0 = rnd(&rng), l1 = rnd(&rng);
                float ndcx = (2.0f * filmx / (float)W) - 1.0f;
                float ndcy = (2.0f * filmy / (float)H) - 1.0f;
                ndcx *= par->aspect;
                v3 org, dir;
                camera_ray(par, O, F, R, U, ndcx, ndcy, l0, l1, filmx, filmy, W, H, &org, &dir);
                wl_t wl; memset(&wl, 0, sizeof wl);
                memset(A.D, 0, sizeof A.D);
                (void)transport_trace(S, &P, org, dir, &rng, &c1, &c2, &wl, &A, NULL);
            }
            for (int k = 0; k < ADJ_N; k++) gr[k] += A.G[k];
        }
    }
    for (int k = 0; k < ADJ_N; k++) grad[k] = 0.0;
    for (int y = 0; y < H; y++) for (int k = 0; k < ADJ_N; k++) grad[k] += rows[(size_t)y * ADJ_N + k];
    free(rows);
}
