// ptmi.hip — libptmi: the C ABI and the wavefront path-tracing pipeline for gfx950.
//
// One frame wave traces frames_in_flight samples for every pixel as independent paths in
// fixed slots of an SoA path state in HBM. Per bounce, four kernels run over compacted
// queues of path ids:
//   extend  — closest hit over the wide BVH (+ analytic objects) for the ray queue
//   bookkeep— queue counters (no host round trip: grids are fixed, sizes read on device)
//   shade   — hit fill, emitter/miss scoring, NEE sample (shadow queue), BSDF continuation,
//             Russian roulette, spawn (next ray queue)
//   shadow  — any-hit for the shadow queue; unoccluded contributions added to the path
// and a final accumulate kernel folds each slot's radiance into the film in sample order
// (average_update, components/accumulator/average/average.glsl:6-10). Radiance is added to
// a path in exactly the order the reference walk adds it (pt.ts:299-464): emission at
// vertex k (shade k), NEE of vertex k (shadow k), emission at vertex k+1 (shade k+1).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "builder.h"
#include "description.h"
#include "formats.h"
#include "dev_common.h"
#include "shading.h"
#include "traversal.h"

using namespace ptmi;

// ============================================================================ kernels

namespace {

constexpr int TBLOCK = 128;     // traversal blocks (2 waves)
constexpr int SBLOCK = 256;     // shading blocks
constexpr int MAX_GRID = 8192;

// Shading categories: the bin kernel sorts every path of a bounce by what its vertex needs,
// so each shade launch runs one code path (no material divergence inside a wave).
enum { CAT_MISS = 0, CAT_LAMBERT = 1, CAT_GGX = 2, CAT_DIELECTRIC = 3, CAT_MEDIUM = 4, CAT_NULL = 5, CAT_MIRROR = 6,
       CAT_ROUGH = 7, N_CAT = 8 };
// Categories whose shading classifies every hit's regions (the transmissive models read both
// sides' indices; media programs track the current medium).
constexpr bool cat_probes(int cat, bool media) { return cat == CAT_DIELECTRIC || cat == CAT_ROUGH || media; }
// counts[] slots
enum { C_RAYQ0 = 0, C_RAYQ1 = 1, C_SHADOW = 2, C_CAT0 = 3, C_EXT_FETCH = C_CAT0 + N_CAT, C_SHD_FETCH, C_CAM_MEDIUM,
       N_COUNTS = 16 };

struct Paths {
    float4* ray_o;   // origin.xyz
    float4* ray_d;   // direction.xyz
    float4* thr;     // throughput.xyz, eta_scale (RGB) | throughput, 4 lanes (spectral)
    float4* rad;     // radiance.xyz (RGB) | 4 lanes (spectral)
    uint2* rng;      // rng_base, rng_dim
    float4* prev;    // prev_query.p.xyz, prev_bsdf_pdf
    int* flags;      // bit 0 prev_was_delta, bits 8..15 null crossings, bits 16..31 bounce
    int* medium;     // current_medium region (media scenes)
    float2* hit;     // t (medium event t after binning), primitive id (bits)
    int* q[2];       // ray queues (current / next iteration)
    int* catq[N_CAT];// shading queues
    int* counts;
    int* s_path;
    float4* s_o;     // shadow origin.xyz, maxDist (opaque) | crossings left (media, int bits)
    float4* s_d;     // shadow direction.xyz, light pdf
    float4* s_c;     // partial contribution.xyz (before visibility), MIS weight
    float4* s_l;     // light point.xyz (media walker), MIS weight (spectral)
    float4* wl;      // hero wavelengths (spectral)
    float4* aov[2];  // first-hit guide samples (albedo, normal) of the paths in flight, or null
    float* eta;      // eta_scale (spectral; RGB keeps it in thr.w)
    unsigned long long* hist;   // [0] closest queries, [1] any-hit queries
    unsigned long long* trav;   // [0] nodes visited, [1] triangles tested (COUNT builds)
    // path-replay adjoint (ADJ kernels): SoA per parameter, [k * cap + path]
    float* adj_D;               // dlog f / dtheta of the path so far (sampling detached)
    float* adj_G;               // the path's gradient
    float* adj_sD;              // [k * cap + shadow entry]: D at the NEE vertex incl. local terms
    const float4* adj_delta;    // dL/dI per pixel (times 1/nframes)
    int cap, npix;
    int adj_medium;             // material whose medium sigma_s / sigma_a / g are parameters
    int adj_n_albedo;
    int adj_albedo[PT_ADJ_MAX_ALBEDO];
    // batched wavefront (BATCH kernels): one DevScene per scene, the scene of each path slot
    const DevScene* scenes;
    int* path_scene;
};
// The scene a path belongs to: the kernel argument, or its entry in the batch.
template <bool BATCH> PT_DEV const DevScene& scene_of(const DevScene& Sk, const Paths& P, int path) {
    return BATCH ? P.scenes[P.path_scene[path]] : Sk;
}
constexpr int ADJ_N = PT_ADJ_N_PARAMS;

// Adjoint helpers. Parameter layout: [0..2] sigma_s (channel c), [3..5] sigma_a (channel c),
// [6] g, [7 + k] albedo k; a contribution c scatters e = dL/dI * c into G with the channel
// structure of D (sigma terms act on one channel, g and albedos on all three).
// D is loaded once per vertex and written back only when a parameter factor changed it; G is
// read-modified-written only where a contribution lands (most vertices have none).
struct AdjState { float D[ADJ_N]; bool dirty; };
PT_DEV void adj_load(const Paths& P, int i, AdjState& A) {
#pragma unroll
    for (int k = 0; k < ADJ_N; k++) A.D[k] = P.adj_D[(size_t)k * P.cap + i];
    A.dirty = false;
}
PT_DEV void adj_store(const Paths& P, int i, const AdjState& A) {
    if (!A.dirty) return;
#pragma unroll
    for (int k = 0; k < ADJ_N; k++) P.adj_D[(size_t)k * P.cap + i] = A.D[k];
}
PT_DEV void adj_scatter(float* G, const float* D, f3 e);
PT_DEV void adj_add_G(const Paths& P, int i, const float* D, f3 e) {
    float G[ADJ_N];
#pragma unroll
    for (int k = 0; k < ADJ_N; k++) G[k] = P.adj_G[(size_t)k * P.cap + i];
    adj_scatter(G, D, e);
#pragma unroll
    for (int k = 0; k < ADJ_N; k++) P.adj_G[(size_t)k * P.cap + i] = G[k];
}
PT_DEV void adj_scatter(float* G, const float* D, f3 e) {
    G[0] += e.x * D[0]; G[1] += e.y * D[1]; G[2] += e.z * D[2];
    G[3] += e.x * D[3]; G[4] += e.y * D[4]; G[5] += e.z * D[5];
    const float es = (e.x + e.y) + e.z;
#pragma unroll
    for (int k = 6; k < ADJ_N; k++) G[k] += es * D[k];
}
PT_DEV f3 adj_e(const Paths& P, int path, sp4 c) {
    float4 d = P.adj_delta[path % P.npix];
    return mk3(d.x * c.x, d.y * c.y, d.z * c.z);
}
PT_DEV int adj_albedo_index(const Paths& P, int mat) {
    int r = -1;
    for (int k = 0; k < P.adj_n_albedo; k++) if (P.adj_albedo[k] == mat) r = k;
    return r;
}
// add v to D[7 + k] without a run-time register index
PT_DEV void adj_add_albedo(float* D, int k, float v) {
#pragma unroll
    for (int kk = 0; kk < PT_ADJ_MAX_ALBEDO; kk++) if (kk == k) D[7 + kk] += v;
}
// d log hg / dg at c = dot(wi, -wo) (volume_scattering/hg/hg.glsl:11-19)
PT_DEV float dlog_hg(float g, float c) {
    float d = 1.0f + g * g - 2.0f * g * c;
    return -2.0f * g / (1.0f - g * g) - 3.0f * (g - c) / d;
}

PT_DEV int fl_bounce(int f) { return (f >> 16) & 0xFFFF; }
PT_DEV int fl_nulls(int f) { return (f >> 8) & 0xFF; }
// bit 1: the path's wavelengths collapsed to the hero (pt.ts:466-475, once per path)
PT_DEV int mk_flags(bool delta, int nulls, int bounce, bool collapsed = false) {
    return (delta ? 1 : 0) | (collapsed ? 2 : 0) | (nulls << 8) | (bounce << 16);
}

template <bool SPEC> PT_DEV void load_thr(const Paths& P, int i, sp4& T, float& eta) {
    float4 t = P.thr[i];
    if (SPEC) { T = ld4(t); eta = P.eta[i]; } else { T = mk4(t.x, t.y, t.z, 0.0f); eta = t.w; }
}
template <bool SPEC> PT_DEV void store_thr(const Paths& P, int i, sp4 T, float eta) {
    if (SPEC) { P.thr[i] = st4(T); P.eta[i] = eta; } else P.thr[i] = make_float4(T.x, T.y, T.z, eta);
}
template <bool SPEC> PT_DEV sp4 load_rad(const Paths& P, int i) {
    float4 v = P.rad[i];
    return SPEC ? ld4(v) : mk4(v.x, v.y, v.z, 0.0f);
}
template <bool SPEC> PT_DEV void store_rad(const Paths& P, int i, sp4 L) {
    P.rad[i] = SPEC ? st4(L) : make_float4(L.x, L.y, L.z, 0.0f);
}
template <bool SPEC> PT_DEV Wl load_wl(const Paths& P, int i) {
    Wl w;
    w.lam = SPEC ? P.wl[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return w;
}

struct Camera {
    float ox, oy, oz, fx, fy, fz, rx, ry, rz, ux, uy, uz, tan_fov, aspect;
    uint32_t salt;
    int off_x, off_y;
    int ortho;          // PT_CAM_ORTHOGRAPHIC
    float ortho_scale;
    int thin;           // PT_CAM_THINLENS
    float aperture, focus;
    int model;          // the PT_CAM_* value (fisheye / cylindrical / equirect branches)
    int projection;     // fisheye radial map
    float fk, cyl_f;
};

// fisheye.glsl FISHEYE_THETA: the radial map of each projection.
PT_DEV float fisheye_theta(int projection, float rho, float K) {
    if (projection == PT_FISHEYE_EQUISOLID) return 2.0f * ptm_asin(gclamp(rho * K, 0.0f, 1.0f));
    if (projection == PT_FISHEYE_STEREOGRAPHIC) return 2.0f * ptm_atan(rho * K);
    if (projection == PT_FISHEYE_ORTHOGRAPHIC) return ptm_asin(gclamp(rho * K, 0.0f, 1.0f));
    return rho * K;
}

// path_state_init's current_medium = scene_region_at(camera origin), once per wave.
template <int STACK>
__global__ void __launch_bounds__(64) k_camera_medium(DevScene S, Paths P, Camera C) {
    __shared__ uint2 stk[STACK * 64];
    if (threadIdx.x == 0) P.counts[C_CAM_MEDIUM] = region_at<64>(S, mk3(C.ox, C.oy, C.oz), stk);
}
// The same per path when camera rays do not share an origin (orthographic), after raygen.
template <int STACK>
__global__ void __launch_bounds__(SBLOCK) k_ray_medium(DevScene S, Paths P, int n) {
    __shared__ uint2 stk[STACK * SBLOCK];
    for (int base = blockIdx.x * SBLOCK; base < n; base += gridDim.x * SBLOCK) {
        const int i = base + threadIdx.x;
        if (i < n) P.medium[i] = region_at<SBLOCK>(S, xyz(P.ray_o[i]), stk);
    }
}

// One camera sample of pixel (x, y) at sampleCount `sample` into path slot i.
template <bool SPEC, bool ADJ>
PT_DEV void raygen_path(const Paths& P, const Camera& C, int W, int H, int i, int x, int y, uint32_t sample, int media, int wl_uniform) {
    // main() (features/accumulation.ts:14-26): pixel = gl_FragCoord.xy + u_pixelOffset
    float fx = (float)x + 0.5f + (float)C.off_x;
    float fy = (float)y + 0.5f + (float)C.off_y;
    Rng r;
    r.base = pcg4d_x((uint32_t)fx, (uint32_t)fy, sample, C.salt);
    r.dim = 0;
    float j0 = r.next(), j1 = r.next();
    float filmx = fx + (j0 - 0.5f), filmy = fy + (j1 - 0.5f);   // pixel/box/box.glsl:5-7
    float l0 = r.next(), l1 = r.next();                           // lens sample (read by the thin lens)
    if (SPEC) {                                                   // wavelengths_sample (spectral.glsl:1-10)
        float u = r.next();
        float l[4];
        for (int k = 0; k < 4; k++) l[k] = wavelength_from_cdf(wl_uniform != 0, ptm_fract(u + 0.25f * (float)k));
        P.wl[i] = make_float4(l[0], l[1], l[2], l[3]);
        P.eta[i] = 1.0f;
    }
    // camera/pinhole/pinhole.glsl:4-18; orthographic/orthographic.glsl:6-21 (the origin slides
    // over the film plane, every ray looks along forward)
    float ndcx = (2.0f * filmx / (float)W) - 1.0f;
    float ndcy = (2.0f * filmy / (float)H) - 1.0f;
    ndcx *= C.aspect;
    f3 F = mk3(C.fx, C.fy, C.fz), R = mk3(C.rx, C.ry, C.rz), U = mk3(C.ux, C.uy, C.uz);
    f3 org = mk3(C.ox, C.oy, C.oz), d;
    if (C.ortho) {
        org = (org + R * (ndcx * C.ortho_scale)) + U * (ndcy * C.ortho_scale);
        d = F;
    } else if (C.model == PT_CAM_FISHEYE) {                    // camera/fisheye/fisheye.glsl:23-38
        const float rho = sqrtf(ndcx * ndcx + ndcy * ndcy);
        if (rho < 1e-6f) d = F;                                // centre pixel: on axis
        else {
            const float th = fisheye_theta(C.projection, rho, C.fk);
            const float tx = ndcx / rho, ty = ndcy / rho;
            d = normalize(F * ptm_cos(th) + (R * tx + U * ty) * ptm_sin(th));
        }
    } else if (C.model == PT_CAM_CYLINDRICAL) {                // camera/cylindrical/cylindrical.glsl:9-21
        const float cx = filmx - 0.5f * (float)W, cy = filmy - 0.5f * (float)H;
        const float phi = cx / C.cyl_f, hh = cy / C.cyl_f;
        d = normalize((F * ptm_cos(phi) + R * ptm_sin(phi)) + U * hh);
    } else if (C.model == PT_CAM_EQUIRECT) {                   // camera/equirect/equirect.glsl:9-22
        const float u = filmx / (float)W, v = filmy / (float)H;
        const float phi = (u - 0.5f) * PTM_TWO_PI, th = (v - 0.5f) * PTM_PI;
        const float ct = ptm_cos(th);
        d = normalize((F * ptm_cos(phi) + R * ptm_sin(phi)) * ct + U * ptm_sin(th));
    } else {
        d = normalize((F + R * (ndcx * C.tan_fov)) + U * (ndcy * C.tan_fov));
        if (C.thin) {                                              // camera/thinlens/thinlens.glsl:6-33
            f3 focus = org + d * (C.focus / dot(d, F));
            float2 lens = concentric_disk(l0, l1);
            org = (org + R * (lens.x * C.aperture)) + U * (lens.y * C.aperture);
            d = normalize(focus - org);
        }
    }
    P.ray_o[i] = f4(org, 0.0f);
    P.ray_d[i] = f4(d, 0.0f);
    P.thr[i] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    P.rad[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.rng[i] = make_uint2(r.base, r.dim);
    for (int a = 0; a < 2; a++)
        if (P.aov[a]) P.aov[a][i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);   // 0 unless a surface is reached
    P.prev[i] = f4(org, 0.0f);
    P.flags[i] = mk_flags(true, 0, 0);
    if (media) P.medium[i] = P.counts[C_CAM_MEDIUM];
    P.q[0][i] = i;
    if (ADJ) {
#pragma unroll
        for (int k = 0; k < ADJ_N; k++) { P.adj_D[(size_t)k * P.cap + i] = 0.0f; P.adj_G[(size_t)k * P.cap + i] = 0.0f; }
    }
}

template <bool SPEC, bool ADJ>
__global__ void __launch_bounds__(SBLOCK) k_raygen(Paths P, Camera C, int W, int H, int n_paths, uint32_t sample0, int media, int wl_uniform) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) P.counts[C_RAYQ0] = n_paths;
    if (i >= n_paths) return;
    int pix = i % (W * H), frame = i / (W * H);
    raygen_path<SPEC, ADJ>(P, C, W, H, i, pix % W, pix / W, sample0 + (uint32_t)frame, media, wl_uniform);
}

// Batched wavefront: a wave is a list of (scene, first frame, frames) chunks laid out back to
// back in the path slots; every scene keeps its own camera, film and sample order.
struct BatchEntry {
    int scene, W, H, npix;
    int f0, nf;
    long long base;       // first path slot of the chunk
    long long film;       // first pixel of the scene's film
};
__global__ void __launch_bounds__(SBLOCK) k_raygen_batch(Paths P, const BatchEntry* E, int ne, const Camera* cams, int n_paths) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) P.counts[C_RAYQ0] = n_paths;
    if (i >= n_paths) return;
    int lo = 0, hi = ne - 1;                      // last entry with base <= i
    while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (E[mid].base <= (long long)i) lo = mid; else hi = mid - 1; }
    const BatchEntry& e = E[lo];
    const int local = (int)((long long)i - e.base);
    const int pix = local % e.npix, frame = e.f0 + local / e.npix;
    P.path_scene[i] = e.scene;
    raygen_path<false, false>(P, cams[e.scene], e.W, e.H, i, pix % e.W, pix / e.W, (uint32_t)frame, 0, 0);
}
__global__ void __launch_bounds__(SBLOCK) k_accumulate_batch(const Paths P, const BatchEntry* E, float4* films) {
    const BatchEntry& e = E[blockIdx.y];
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= e.npix) return;
    float4* film = films + e.film;
    float4 acc = film[pix];
    for (int f = 0; f < e.nf; f++) {
        const uint32_t n = (uint32_t)(e.f0 + f);
        float4 x = P.rad[e.base + (long long)f * e.npix + pix];
        if (n == 0) acc = make_float4(x.x, x.y, x.z, 1.0f);
        else {
            float w = 1.0f / ((float)n + 1.0f);
            acc = make_float4(acc.x * (1.0f - w) + x.x * w, acc.y * (1.0f - w) + x.y * w, acc.z * (1.0f - w) + x.z * w, 1.0f);
        }
    }
    film[pix] = acc;
}
__global__ void __launch_bounds__(SBLOCK) k_fill_f4(float4* p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Persistent-lane work distribution: each wave owns a chunk of queue entries and hands them
// to lanes as they go idle (no per-ray atomics; one atomic per CHUNK entries per wave).
struct LaneFeeder {
    uint32_t next, end;
    bool exhausted;
    PT_DEV void init() { next = end = 0; exhausted = false; }
    // Idle lanes get a queue index (or -1). Convergent: every lane of the wave calls it.
    template <int CHUNK>
    PT_DEV int feed(int* counter, uint32_t n, bool idle, int min_idle) {
        uint64_t im = __ballot(idle);
        int nidle = (int)__popcll(im);
        if (nidle == 0 || exhausted) return -1;
        uint64_t am = __ballot(!idle);
        if (am != 0 && nidle < min_idle) return -1;
        int lane = (int)__lane_id();
        if (end == next) {
            uint32_t base = 0;
            if (lane == 0) base = (uint32_t)atomicAdd(counter, CHUNK);
            base = __shfl(base, 0);
            if (base >= n) { exhausted = true; return -1; }
            next = base;
            end = min(base + (uint32_t)CHUNK, n);
        }
        uint32_t avail = end - next;
        uint32_t pre = (uint32_t)__popcll(im & ((1ull << lane) - 1ull));
        int idx = (idle && pre < avail) ? (int)(next + pre) : -1;
        next += min((uint32_t)nidle, avail);
        return idx;
    }
};

#ifndef PTMI_FEED_CHUNK
#define PTMI_FEED_CHUNK 64
#endif
#ifndef PTMI_FEED_MIN_IDLE
#define PTMI_FEED_MIN_IDLE 8
#endif
#ifndef PTMI_POSTPONE_EXT
#define PTMI_POSTPONE_EXT 0
#endif
#ifndef PTMI_POSTPONE_SHD
#define PTMI_POSTPONE_SHD 0
#endif
constexpr int POSTPONE_EXT = PTMI_POSTPONE_EXT;    // triangle postponing below this many lanes (0: off)
constexpr int POSTPONE_SHD = PTMI_POSTPONE_SHD;    // the same for opaque shadow rays
constexpr int FEED_CHUNK = PTMI_FEED_CHUNK;        // queue entries a wave claims per atomic
constexpr int FEED_MIN_IDLE = PTMI_FEED_MIN_IDLE;  // idle lanes before a wave refills

// extend: closest hit for every ray of the current queue, then bin the path by shading category.
// INST: the scene has instance batches. The TLAS walk keeps a private stack, so only scenes
// with placements compile it in (the plain kernels then need no scratch memory).
template <int STACK, bool COUNT, bool BATCH = false, int TB = TBLOCK, bool INST = false>
__global__ void __launch_bounds__(TB) k_extend(DevScene Sk, Paths P, int qsel) {
    __shared__ uint2 stk[STACK * TB];
    const uint32_t n = (uint32_t)P.counts[qsel];
    const int* q = P.q[qsel];
    TraceCounters cnt = {0, 0};
    LaneFeeder feed; feed.init();
    TravState st;
    bool active = false;
    int path = -1;
    const DevScene* Sp = &Sk;
    for (;;) {
        int k = feed.feed<FEED_CHUNK>(&P.counts[C_EXT_FETCH], n, !active, FEED_MIN_IDLE);
        if (k >= 0) {
            path = q[k];
            Sp = &scene_of<BATCH>(Sk, P, path);
            float4 o4 = P.ray_o[path], d4 = P.ray_d[path];
            trace_init(st, xyz(o4), xyz(d4), MAX_DIST);
            active = true;
        }
        if (__ballot(active) == 0) {
            if (feed.exhausted) break;
            continue;
        }
        if (active && (Sp->n_cw_nodes == 0 ||
                       trace_step<false, TB, COUNT, POSTPONE_EXT, STACK>(Sp->cw_nodes, Sp->cw_tris, st, stk, cnt, Sp->cw_depth))) {
            float tmax = st.tmax; int prim = st.prim;
            trace_analytic<false>(*Sp, st.o, st.d, tmax, prim);
            if (INST) trace_instances<false, TB, COUNT>(*Sp, st.o, st.d, tmax, prim, stk, cnt);
            P.hit[path] = make_float2(tmax, __int_as_float(prim));
            active = false;
        }
    }
    if (COUNT) {
        atomicAdd(&P.trav[0], (unsigned long long)cnt.nodes);
        atomicAdd(&P.trav[1], (unsigned long long)cnt.tris);
    }
}

// shadow (opaque scenes): any hit for every NEE ray; unoccluded contributions are added to
// their path as ((a * vis) * w) / pdf, the reference's operation order with vis = 1.
template <int STACK, bool SPEC, bool ADJ, bool BATCH = false, int TB = TBLOCK, bool INST = false>
__global__ void __launch_bounds__(TB) k_shadow(DevScene Sk, Paths P) {
    __shared__ uint2 stk[STACK * TB];
    const int n = P.counts[C_SHADOW];
    TraceCounters cnt = {0, 0};
    for (int base = blockIdx.x * TB; base < n; base += gridDim.x * TB) {
        int k = base + threadIdx.x;
        if (k >= n) break;
        const DevScene& S = scene_of<BATCH>(Sk, P, P.s_path[k]);
        float4 o4 = P.s_o[k], d4 = P.s_d[k];
        f3 o = xyz(o4), d = xyz(d4);
        float maxd = o4.w; int prim = -1;
        bool blocked = trace_analytic<true>(S, o, d, maxd, prim);
        if (!blocked) blocked = trace_wide<true, STACK, TB, false, POSTPONE_SHD>(S, o, d, maxd, prim, stk, cnt);
        if (INST && !blocked) blocked = trace_instances<true, TB, false>(S, o, d, maxd, prim, stk, cnt);
        if (!blocked) {
            int i = P.s_path[k];
            float4 c = P.s_c[k];
            const float w = SPEC ? P.s_l[k].w : c.w;
            sp4 a = SPEC ? ld4(c) : mk4(c.x, c.y, c.z, 0.0f);
            const sp4 v = (a * w) / d4.w;
            store_rad<SPEC>(P, i, load_rad<SPEC>(P, i) + v);
            if (ADJ) {
                float D[ADJ_N];
#pragma unroll
                for (int kk = 0; kk < ADJ_N; kk++) D[kk] = P.adj_sD[(size_t)kk * P.cap + k];
                adj_add_G(P, i, D, adj_e(P, i, v));
            }
        }
    }
}

// shadow (media scenes): the segment walker of transport/shadow/media/media.glsl:29-62 —
// closest hits up to the light point, Beer-Lambert per segment in the current medium, pass
// through null interfaces, anything else blocks (dielectrics are opaque to shadow rays).
template <int STACK, bool SPEC, bool ADJ, bool INST = false>
__global__ void __launch_bounds__(TBLOCK) k_shadow_media(DevScene S, Paths P) {
    __shared__ uint2 stk[STACK * TBLOCK];
    const int n = P.counts[C_SHADOW];
    TraceCounters cnt = {0, 0};
    for (int base = blockIdx.x * TBLOCK; base < n; base += gridDim.x * TBLOCK) {
        int k = base + threadIdx.x;
        if (k >= n) break;
        float4 o4 = P.s_o[k], d4 = P.s_d[k], l4 = P.s_l[k];
        f3 o = xyz(o4), d = xyz(d4), light_p = xyz(l4);
        const int crossings_left = __float_as_int(o4.w);
        const Wl wl = load_wl<SPEC>(P, P.s_path[k]);
        sp4 T = sgrey(1.0f);
        float len_param = 0.0f;     // length travelled inside the parameter medium
        bool blocked = true;
        int medium = region_at<TBLOCK>(S, o, stk);
        for (int crossed = 0; crossed <= crossings_left; crossed++) {
            float remaining = length(light_p - o) - SHADOW_BACKOFF;
            if (remaining <= 0.0f) { blocked = false; break; }
            float tmax = MAX_DIST; int prim = -1;
            if (S.n_cw_nodes > 0) trace_wide<false, STACK, TBLOCK, false>(S, o, d, tmax, prim, stk, cnt);
            trace_analytic<false>(S, o, d, tmax, prim);
            if (INST) trace_instances<false, TBLOCK, false>(S, o, d, tmax, prim, stk, cnt);
            bool hit_boundary = prim >= 0 && tmax < remaining;
            float seg_len = hit_boundary ? tmax : remaining;
            int med_mat = material_of(S, medium);
            if (material_has_medium(S, med_mat)) T = T * medium_transmittance<SPEC>(S, S.materials[med_mat], seg_len, wl);
            if (ADJ && med_mat == P.adj_medium && material_has_medium(S, med_mat)) len_param += seg_len;
            if (!hit_boundary) { blocked = false; break; }
            Hit h; bool back;
            fill_hit(S, o, d, tmax, prim, h, back);
            if (!is_null_interface(S, material_of(S, h.owner))) break;
            classify_hit<TBLOCK>(S, h, back, true, stk);
            medium = h.region_to;
            float side = dot(d, h.ng) >= 0.0f ? 1.0f : -1.0f;
            o = h.p + (h.ng * side) * h.eps;
            d = normalize(light_p - o);
        }
        if (!blocked && !is_black<SPEC>(T)) {
            int i = P.s_path[k];
            float4 c = P.s_c[k];
            const float w = SPEC ? l4.w : c.w;
            sp4 a = SPEC ? ld4(c) : mk4(c.x, c.y, c.z, 0.0f);
            const sp4 v = ((a * T) * w) / d4.w;
            store_rad<SPEC>(P, i, load_rad<SPEC>(P, i) + v);
            if (ADJ) {
                float D[ADJ_N];
#pragma unroll
                for (int kk = 0; kk < ADJ_N; kk++) D[kk] = P.adj_sD[(size_t)kk * P.cap + k];
#pragma unroll
                for (int kk = 0; kk < 6; kk++) D[kk] -= len_param;     // e^{-sigma_t len} along the shadow ray
                adj_add_G(P, i, D, adj_e(P, i, v));
            }
        }
    }
}

// Block-aggregated queue appends: LDS atomics give each lane its slot inside the block's
// share, then one global atomic per queue per block. Every thread of the block must call it.
template <int K>
PT_DEV int block_append(int* const* counters, int which, int* s_cnt, int* s_base) {
    if (threadIdx.x < K) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    int local = which >= 0 ? atomicAdd(&s_cnt[which], 1) : -1;
    __syncthreads();
    if (threadIdx.x < K) s_base[threadIdx.x] = s_cnt[threadIdx.x] ? atomicAdd(counters[threadIdx.x], s_cnt[threadIdx.x]) : 0;
    __syncthreads();
    return which >= 0 ? s_base[which] + local : -1;
}

// bin: per path of the iteration, the medium segment first (media scenes: the analytic
// channel-MIS distance sample of pt.ts's media branch, which may end the segment in a
// scattering event), then the shading category of the vertex.
template <bool MEDIA, bool SPEC, bool ADJ, bool BATCH = false>
__global__ void __launch_bounds__(SBLOCK) k_bin(DevScene Sk, Paths P, int qsel, int equiangular, int max_bounces, int max_null) {
    __shared__ int s_cnt[N_CAT], s_base[N_CAT];
    int* ctr[N_CAT];
    for (int c = 0; c < N_CAT; c++) ctr[c] = &P.counts[C_CAT0 + c];
    int* sctr[1] = {&P.counts[C_SHADOW]};
    const int n = P.counts[qsel];
    const int* q = P.q[qsel];
    for (int base = blockIdx.x * SBLOCK; base < n; base += gridDim.x * SBLOCK) {
        int k = base + threadIdx.x;
        int cat = -1, path = 0;
        bool eq_shadow = false;
        f3 eq_o = mk3(0, 0, 0), eq_wi = mk3(0, 0, 0), eq_pos = mk3(0, 0, 0);
        sp4 eq_a = sgrey(0.0f);
        float eq_pdf = 1.0f;
        int eq_cross = 0;
        if (k < n) {
            path = q[k];
            const DevScene& S = scene_of<BATCH>(Sk, P, path);
            float2 hv = P.hit[path];
            int prim = __float_as_int(hv.y);
            if (MEDIA) {
                int med_mat = material_of(S, P.medium[path]);
                if (material_has_medium(S, med_mat)) {
                    uint2 rs = P.rng[path];
                    Rng rng; rng.base = rs.x; rng.dim = rs.y;
                    const float t_max = prim >= 0 ? hv.x : MAX_DIST;
                    const int fl = P.flags[path];
                    if (!ADJ && !BATCH && equiangular && fl_bounce(fl) < max_bounces) {
                        // equiangular_sample_direct (techniques/equiangular/equiangular.glsl:18-57): one
                        // direct-light vertex per medium segment, from segment-start throughput; the
                        // shadow walker runs before this iteration's shading, in the reference's order.
                        const pt_material& Mm = S.materials[med_mat];
                        const Wl wl = load_wl<SPEC>(P, path);
                        const float uc = rng.next();
                        int id = S.n_lights - 1;
                        for (int li = 0; li < S.n_lights - 1; li++) if (uc < S.lights[li].cdf) { id = li; break; }
                        const pt_light& l = S.lights[id];
                        const float select_pdf = l.select_pdf;
                        const f3 pos = ld3(l.corner);
                        const sp4 intensity = lift<SPEC>(S, l.radiance, l.s_radiance, wl);
                        const f3 o = xyz(P.ray_o[path]), d = xyz(P.ray_d[path]);
                        const f3 D = pos - o;
                        const float t_c = dot(D, d);
                        const float h2 = gmax(dot(D, D) - t_c * t_c, 1e-8f);
                        const float h = sqrtf(h2);
                        const float theta_a = ptm_atan2(0.0f - t_c, h);
                        const float theta_b = ptm_atan2(t_max - t_c, h);
                        const float u = rng.next();
                        const float theta = glsl_mix(theta_a, theta_b, u);
                        const float t = t_c + h * ptm_tan(theta);
                        const float pdf_t = h / ((theta_b - theta_a) * (h2 + (t - t_c) * (t - t_c)));
                        const f3 p_evt = o + d * t;
                        sp4 Tp; float etap;
                        load_thr<SPEC>(P, path, Tp, etap);
                        const sp4 T_seg = spectrum_exp((-(sigma_a_of<SPEC>(S, Mm, wl) + sigma_s_of<SPEC>(S, Mm, wl))) * t);
                        const f3 to_light = pos - p_evt;
                        const float d2 = gmax(dot(to_light, to_light), 1e-8f);
                        const float dist = sqrtf(d2);
                        eq_wi = mk3(to_light.x / dist, to_light.y / dist, to_light.z / dist);
                        const float f = hg_value(eq_wi, -d, Mm.hg_g);
                        eq_a = (((Tp * T_seg) * sigma_s_of<SPEC>(S, Mm, wl)) * sgrey(f)) * (intensity / d2);
                        eq_pdf = pdf_t * select_pdf;
                        eq_o = p_evt; eq_pos = pos;
                        eq_cross = S.has_null ? max_null - fl_nulls(fl) : 0;
                        eq_shadow = true;
                    }
                    float x0 = rng.next(), x1 = rng.next();
                    const pt_material& Mm = S.materials[med_mat];
                    const Wl wl = load_wl<SPEC>(P, path);
                    // medium_sample dispatch by the material's arm (features/materials.ts:675-728)
                    MSample ms = Mm.medium_arm == PT_MED_DELTA ? medium_sample_delta<SPEC>(S, Mm, t_max, x0, x1, rng, wl)
                               : Mm.medium_arm == PT_MED_ABSORB ? medium_sample_absorb<SPEC>(S, Mm, t_max, wl)
                               : medium_sample_analytic<SPEC>(S, Mm, t_max, x0, x1, wl);
                    sp4 T; float eta;
                    load_thr<SPEC>(P, path, T, eta);
                    // the source term, weighted by segment-start throughput (pt.ts:333-339)
                    if (medium_emits(Mm)) store_rad<SPEC>(P, path, load_rad<SPEC>(P, path) + T * ms.radiance);
                    store_thr<SPEC>(P, path, T * ms.weight, eta);
                    if (ADJ && med_mat == P.adj_medium) {   // f = sigma_s e^{-sigma_t t} | e^{-sigma_t t_max}
                        const float* ssv = S.materials[med_mat].sigma_s;
#pragma unroll
                        for (int ch = 0; ch < 3; ch++) {
                            const size_t ks = (size_t)ch * P.cap + path, ka = (size_t)(3 + ch) * P.cap + path;
                            if (Mm.medium_arm != PT_MED_ABSORB)
                                P.adj_D[ks] += (ms.scattered && ssv[ch] > 0.0f ? 1.0f / ssv[ch] : 0.0f) - ms.t;
                            P.adj_D[ka] -= ms.t;
                        }
                    }
                    P.rng[path] = make_uint2(rng.base, rng.dim);
                    if (ms.scattered) { cat = CAT_MEDIUM; P.hit[path] = make_float2(ms.t, hv.y); }
                }
            }
            if (cat < 0) {
                if (prim < 0) cat = CAT_MISS;
                else {
                    int owner = prim_owner(S, prim);
                    int model = S.materials[S.objects[owner].material].model;
                    cat = model == PT_MAT_LAMBERT ? CAT_LAMBERT
                        : model == PT_MAT_GGX ? CAT_GGX
                        : model == PT_MAT_DIELECTRIC ? CAT_DIELECTRIC
                        : model == PT_MAT_MIRROR ? CAT_MIRROR
                        : model == PT_MAT_ROUGH_DIELECTRIC ? CAT_ROUGH : CAT_NULL;
                }
            }
        }
        int slot = block_append<N_CAT>(ctr, cat, s_cnt, s_base);
        if (cat >= 0) P.catq[cat][slot] = path;
        if (MEDIA && !ADJ && !BATCH && equiangular) {
            int ss = block_append<1>(sctr, eq_shadow ? 0 : -1, s_cnt, s_base);
            if (eq_shadow) {   // the shadow walker adds ((a * vis) * 1) / pdf
                P.s_path[ss] = path;
                P.s_o[ss] = f4(eq_o, __int_as_float(eq_cross));
                P.s_d[ss] = f4(eq_wi, eq_pdf);
                P.s_c[ss] = SPEC ? st4(eq_a) : make_float4(eq_a.x, eq_a.y, eq_a.z, 1.0f);
                P.s_l[ss] = f4(eq_pos, 1.0f);
            }
        }
    }
}

// After the equiangular shadow rays of an iteration: count them and empty the shadow queue
// for the surface NEE of the same iteration.
__global__ void k_shadow_reset(Paths P) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        P.hist[1] += (unsigned long long)P.counts[C_SHADOW];
        P.counts[C_SHADOW] = 0;
    }
}

// Queue bookkeeping before an iteration (single thread).
__global__ void k_bookkeep(Paths P, int cur, int next, int first) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (!first) P.hist[1] += (unsigned long long)P.counts[C_SHADOW];
        if (cur >= 0) P.hist[0] += (unsigned long long)P.counts[cur];
        if (next >= 0) P.counts[next] = 0;
        P.counts[C_SHADOW] = 0;
        for (int c = 0; c < N_CAT; c++) P.counts[C_CAT0 + c] = 0;
        P.counts[C_EXT_FETCH] = 0;
        P.counts[C_SHD_FETCH] = 0;
    }
}

// The per-vertex body of transport_trace (integrators/pt/pt.ts:299-464) with the combiner
// (transport/combiner.ts:23-102) and the kernel/light techniques (techniques/kernel/kernel.glsl,
// kernel_phase.glsl, techniques/light/light.glsl, light_medium.glsl), for the paths of one
// category. MEDIA selects the media program (region tracking, medium NEE shadow walker).
template <int EST, int CAT, int STACK, bool MEDIA, bool SPEC, bool ADJ, bool BATCH = false>
__global__ void __launch_bounds__(SBLOCK) k_shade(DevScene Sk, Paths P, int qsel, int max_bounces,
                                                  int rr_start, float rr_max, int max_null, int equiangular = 0) {
    constexpr bool PROBE = cat_probes(CAT, MEDIA);
    // The region probe's LDS stack: static in the probing categories; the other surface
    // categories probe only at a back-face hit in a scene with an emissive solid, and are then
    // launched with the same stack as dynamic LDS (shade_dynamic_lds).
    constexpr bool STATIC_STK = PROBE && CAT != CAT_MISS && CAT != CAT_MEDIUM;
    __shared__ uint2 stk_static[STATIC_STK ? STACK * SBLOCK : 1];
    extern __shared__ uint2 stk_dynamic[];
    uint2* stk = STATIC_STK ? stk_static : stk_dynamic;
    __shared__ int s_cnt[2], s_base[2];
    int* ctr[2] = {&P.counts[qsel ^ 1], &P.counts[C_SHADOW]};
    const int n = P.counts[C_CAT0 + CAT];
    const int* q = P.catq[CAT];
    int* qn = P.q[qsel ^ 1];
    constexpr bool NEE = EST != PT_EST_PT;
    constexpr bool MIS = EST == PT_EST_MIS;
    constexpr int MODEL = CAT == CAT_LAMBERT ? PT_MAT_LAMBERT : CAT == CAT_GGX ? PT_MAT_GGX
                        : CAT == CAT_MIRROR ? PT_MAT_MIRROR : CAT == CAT_ROUGH ? PT_MAT_ROUGH_DIELECTRIC
                        : PT_MAT_DIELECTRIC;
    for (int base = blockIdx.x * SBLOCK; base < n; base += gridDim.x * SBLOCK) {
        const int k = base + threadIdx.x;
        const bool valid = k < n;
        const int i = valid ? q[k] : 0;
        bool cont = false, shadow = false;
        f3 s_o = mk3(0, 0, 0), s_d = mk3(0, 0, 0), s_l = mk3(0, 0, 0);
        sp4 s_a = sgrey(0.0f);
        float s_max = 0.0f, s_w = 0.0f, s_pdf = 1.0f;
        int s_cross = 0;
        AdjState A;
        float sD[ADJ_N];
        if (valid) {
            const DevScene& S = scene_of<BATCH>(Sk, P, i);
            if (ADJ && (CAT != CAT_MISS || S.env_type != PT_ENV_NONE)) adj_load(P, i, A);
            else A.dirty = false;
            f3 rd = xyz(P.ray_d[i]);
            sp4 T; float eta_scale;
            load_thr<SPEC>(P, i, T, eta_scale);
            sp4 L = load_rad<SPEC>(P, i);
            const Wl wl = load_wl<SPEC>(P, i);
            float4 pv = P.prev[i];
            f3 prev_qp = xyz(pv); float prev_pdf = pv.w;
            const int fl = P.flags[i];
            bool prev_delta = (fl & 1) != 0;
            bool collapsed = SPEC && (fl & 2) != 0;
            int nulls = fl_nulls(fl);
            const int bounce = fl_bounce(fl);
            int medium = MEDIA ? P.medium[i] : -1;
            const int crossings_left = S.has_null ? max_null - nulls : 0;
            if (CAT == CAT_MISS) {
                // kernel_score_miss + combiner_w_env
                float w = 1.0f;
                if (NEE && env_samplable(S)) {
                    if (prev_delta) w = 1.0f;
                    else if (MIS) w = power_heuristic(prev_pdf, env_select_live(S) ? S.env_select_prob * environment_pdf(S, rd) : environment_pdf(S, rd));
                    else w = 0.0f;
                }
                if (S.env_type != PT_ENV_NONE) {
                    const sp4 c = (T * w) * env_radiance<SPEC>(S, rd, wl);
                    L = L + c;
                    if (ADJ) adj_add_G(P, i, A.D, adj_e(P, i, c));
                }
            } else if (CAT == CAT_MEDIUM) {
                // ---- medium event (pt.ts media branch: scattered) ----
                if (bounce < max_bounces) {
                    const int mmat = material_of(S, medium);
                    const pt_material& Mm = S.materials[mmat];
                    const float g = Mm.hg_g;
                    const bool param_medium = ADJ && mmat == P.adj_medium;
                    f3 ro = xyz(P.ray_o[i]);
                    const float t = P.hit[i].x;
                    f3 p_evt = ro + rd * t;
                    f3 wo_med = -rd;
                    uint2 rs = P.rng[i];
                    Rng rng; rng.base = rs.x; rng.dim = rs.y;
                    if (NEE && !equiangular) {   // light_sample_direct_medium
                        float y0 = rng.next(), y1 = rng.next();
                        LSample ls = lighting_sample<SPEC>(S, p_evt, y0, y1, wl);
                        if (ls.pdf > 0.0f) {
                            s_l = p_evt + ls.wi * ls.distance;
                            s_o = p_evt; s_d = ls.wi; s_cross = crossings_left;
                            float f = hg_value(ls.wi, wo_med, g);
                            s_w = 1.0f;
                            if (MIS && !(ls.flags & 1)) s_w = power_heuristic(ls.pdf, hg_pdf<SPEC>(ls.wi, wo_med, g));
                            s_a = (T * ls.radiance) * sgrey(f);
                            s_pdf = ls.pdf;
                            shadow = true;
                            if (ADJ) {
#pragma unroll
                                for (int kk = 0; kk < ADJ_N; kk++) sD[kk] = A.D[kk];
                                if (param_medium) sD[6] += dlog_hg(g, dot(ls.wi, -wo_med));
                            }
                        }
                    }
                    // kernel_sample_phase
                    float z0 = rng.next(), z1 = rng.next();
                    float ppdf;
                    f3 wi = hg_sample<SPEC>(wo_med, g, z0, z1, ppdf);
                    if (param_medium) { A.D[6] += dlog_hg(g, dot(wi, -wo_med)); A.dirty = true; }
                    cont = true;
                    // roulette_interior, present when some medium weights absorption; medium_survival =
                    // sigma_s / max(sigma_t, 1e-9) there, 1 on the tracking arm (features/materials.ts:731-772)
                    if (rr_start >= 0 && bounce >= rr_start && S.weighted_absorption) {
                        sp4 surv = sgrey(1.0f);
                        if (Mm.medium_arm == PT_MED_ANALYTIC) {
                            sp4 ss = sigma_s_of<SPEC>(S, Mm, wl), st = ss + sigma_a_of<SPEC>(S, Mm, wl);
                            surv = mk4(ss.x / gmax(st.x, 1e-9f), ss.y / gmax(st.y, 1e-9f), ss.z / gmax(st.z, 1e-9f), ss.w / gmax(st.w, 1e-9f));
                        }
                        float p = spectrum_max<SPEC>(surv);
                        if (p <= 0.0f) cont = false;
                        else if (rng.next() > p) cont = false;
                        else T = T / p;
                    }
                    if (cont) {
                        P.ray_o[i] = f4(p_evt, 0.0f);
                        P.ray_d[i] = f4(wi, 0.0f);
                        store_thr<SPEC>(P, i, T, eta_scale);
                        P.prev[i] = f4(p_evt, ppdf);
                        P.flags[i] = mk_flags(false, nulls, bounce + 1, collapsed);
                    }
                    P.rng[i] = make_uint2(rng.base, rng.dim);
                }
            } else {
                float2 hv = P.hit[i];
                const int prim = __float_as_int(hv.y);
                f3 ro = xyz(P.ray_o[i]);
                const bool emitters = S.n_lights > 0;
                Hit h; bool back;
                fill_hit(S, ro, rd, hv.x, prim, h, back);
                const pt_material& M = S.materials[S.objects[h.owner].material];
                bool need = PROBE || (S.emissive_solid && back);
                classify_hit<SBLOCK>(S, h, back, need, stk);
                // first-hit guides (pt.ts:422-433): the first non-null surface with no event before
                if (!SPEC && CAT != CAT_NULL && bounce == 0 && (P.aov[0] || P.aov[1])) {
                    const sp4 Ta = T;                            // the path's weight on arrival
                    if (P.aov[0]) {                                  // aov/albedo/albedo.glsl
                        const float* row = MODEL == PT_MAT_LAMBERT ? M.albedo
                                         : (MODEL == PT_MAT_GGX || MODEL == PT_MAT_MIRROR) ? M.f0 : M.transmittance;
                        P.aov[0][i] = make_float4(gclamp(row[0], 0.0f, 1.0f) * Ta.x, gclamp(row[1], 0.0f, 1.0f) * Ta.y,
                                                  gclamp(row[2], 0.0f, 1.0f) * Ta.z, 0.0f);
                    }
                    if (P.aov[1]) {                                  // aov/normal/normal.glsl
                        const f3 n = dot(h.frame.n, -rd) < 0.0f ? -h.frame.n : h.frame.n;
                        const float avg = ((Ta.x + Ta.y) + Ta.z) * (1.0f / 3.0f);
                        P.aov[1][i] = make_float4(n.x * avg, n.y * avg, n.z * avg, 0.0f);
                    }
                }
                if (MEDIA && h.region_from != medium) medium = h.region_from;   // self-heal
                if (CAT == CAT_NULL) {
                    // null interface: not an optical event; no bounce consumed
                    medium = h.region_to;
                    float side = dot(rd, h.ng) >= 0.0f ? 1.0f : -1.0f;
                    nulls++;
                    cont = nulls <= max_null;
                    if (cont) {
                        P.ray_o[i] = f4(h.p + (h.ng * side) * h.eps, 0.0f);
                        P.flags[i] = mk_flags(prev_delta, nulls, bounce, collapsed);
                        if (MEDIA) P.medium[i] = medium;
                    }
                } else {
                    uint2 rs = P.rng[i];
                    Rng rng; rng.base = rs.x; rng.dim = rs.y;
                    f3 wo = -rd;
                    // kernel_score_emitter_hit
                    int mat_emit = h.region_to >= -1 ? material_of(S, h.region_to) : -1;
                    if (mat_emit >= 0) {
                        const pt_material& E = S.materials[mat_emit];
                        if (E.emissive) {
                            float w = 1.0f;
                            if (NEE && emitters) {
                                int lid = light_of(S, h.region_to);
                                if (lid < 0 || prev_delta) w = 1.0f;
                                else if (MIS) w = power_heuristic(prev_pdf, lighting_pdf(S, prev_qp, rd, lid, h.p));
                                else w = 0.0f;
                            }
                            const sp4 c = (T * w) * lift<SPEC>(S, E.emission, E.s_emission, wl);
                            L = L + c;
                            if (ADJ) adj_add_G(P, i, A.D, adj_e(P, i, c));
                        }
                    }
                    if (bounce < max_bounces) {
                        // collapse_wavelengths (pt.ts:440,466-475): a dielectric bounding a dispersive
                        // region sends the path along the hero wavelength, lane 0 carrying all four
                        if (SPEC && (CAT == CAT_DIELECTRIC || CAT == CAT_ROUGH) && !collapsed &&
                            (dispersive_region(S, h.region_from) || dispersive_region(S, h.region_to))) {
                            T = T * mk4(4.0f, 0.0f, 0.0f, 0.0f);
                            collapsed = true;
                        }
                        // light_sample_direct (skipped for delta-only materials)
                        if (NEE && CAT != CAT_DIELECTRIC && CAT != CAT_MIRROR) {   // delta lobes take no NEE
                            float x0 = rng.next(), x1 = rng.next();
                            LSample ls = lighting_sample<SPEC>(S, h.p, x0, x1, wl);
                            if (ls.pdf > 0.0f) {
                                f3 light_p = h.p + ls.wi * ls.distance;
                                float side = dot(ls.wi, h.ng) >= 0.0f ? 1.0f : -1.0f;
                                s_o = h.p + (h.ng * side) * h.eps;
                                s_d = normalize(light_p - s_o);
                                s_max = length(light_p - s_o) - SHADOW_BACKOFF;
                                s_l = light_p; s_cross = crossings_left;
                                sp4 f = surface_eval<MODEL, SPEC>(S, M, ls.wi, wo, h, wl);
                                float cos_i = fabsf(dot(ls.wi, h.frame.n));
                                s_w = 1.0f;
                                if (MIS && !(ls.flags & 1)) s_w = power_heuristic(ls.pdf, surface_pdf<MODEL, SPEC>(S, M, ls.wi, wo, h, wl));
                                s_a = ((T * ls.radiance) * f) * cos_i;
                                s_pdf = ls.pdf;
                                shadow = true;
                                if (ADJ) {
#pragma unroll
                                    for (int kk = 0; kk < ADJ_N; kk++) sD[kk] = A.D[kk];
                                    if (CAT == CAT_LAMBERT) {
                                        const int ak = adj_albedo_index(P, S.objects[h.owner].material);
                                        if (ak >= 0) adj_add_albedo(sD, ak, 1.0f / M.albedo[0]);
                                    }
                                }
                            }
                        }
                        // kernel_sample_continuation
                        float uc = rng.next(), ux = rng.next(), uy = rng.next();
                        ISample bs = surface_sample<MODEL, SPEC>(S, M, wo, h, uc, ux, uy, wl);
                        if (!is_black<SPEC>(bs.weight)) {
                            T = T * bs.weight;
                            if (ADJ && CAT == CAT_LAMBERT) {
                                const int ak = adj_albedo_index(P, S.objects[h.owner].material);
                                if (ak >= 0) { adj_add_albedo(A.D, ak, 1.0f / M.albedo[0]); A.dirty = true; }
                            }
                            prev_delta = (bs.flags & LOBE_DELTA) != 0;
                            prev_pdf = bs.pdf;
                            prev_qp = h.p;
                            if (MEDIA && (bs.flags & LOBE_TRANSMISSION)) medium = h.region_to;
                            if (S.has_transmission && (bs.flags & LOBE_TRANSMISSION)) {
                                float r = ior_of<SPEC>(S, h.region_to, wl) / ior_of<SPEC>(S, h.region_from, wl);
                                eta_scale *= r * r;
                            }
                            cont = true;
                            // roulette (pt.ts:204-222)
                            if (rr_start >= 0 && bounce >= rr_start) {
                                float m = S.has_transmission ? spectrum_max<SPEC>(T) * eta_scale : spectrum_max<SPEC>(T);
                                float p = gmin(rr_max, m);
                                if (p <= 0.0f) cont = false;
                                else if (rng.next() > p) cont = false;
                                else T = T / p;
                            }
                            if (cont) {
                                float side = dot(bs.wi, h.ng) >= 0.0f ? 1.0f : -1.0f;
                                P.ray_o[i] = f4(h.p + (h.ng * side) * h.eps, 0.0f);
                                P.ray_d[i] = f4(bs.wi, 0.0f);
                                store_thr<SPEC>(P, i, T, eta_scale);
                                P.prev[i] = f4(prev_qp, prev_pdf);
                                P.flags[i] = mk_flags(prev_delta, nulls, bounce + 1, collapsed);
                                P.rng[i] = make_uint2(rng.base, rng.dim);
                                if (MEDIA) P.medium[i] = medium;
                            }
                        }
                    }
                }
            }
            store_rad<SPEC>(P, i, L);
            if (ADJ) adj_store(P, i, A);
        }
        int slot = block_append<2>(ctr, cont ? 0 : -1, s_cnt, s_base);
        if (cont) qn[slot] = i;
        int sslot = block_append<2>(ctr, shadow ? 1 : -1, s_cnt, s_base);
        if (shadow) {
            P.s_path[sslot] = i;
            P.s_o[sslot] = f4(s_o, MEDIA ? __int_as_float(s_cross) : s_max);
            P.s_d[sslot] = f4(s_d, s_pdf);
            P.s_c[sslot] = SPEC ? st4(s_a) : make_float4(s_a.x, s_a.y, s_a.z, s_w);
            if (MEDIA || SPEC) P.s_l[sslot] = f4(s_l, s_w);
            if (ADJ) {
#pragma unroll
                for (int kk = 0; kk < ADJ_N; kk++) P.adj_sD[(size_t)kk * P.cap + sslot] = sD[kk];
            }
        }
    }
}

// Adjoint reductions. Deterministic: per pixel, the wave's paths are added in sample order
// into a per-pixel partial; after the last wave one block per parameter sums the partials
// in a fixed order. Atomic: block partials (fixed-order LDS tree) are added to the gradient
// with fp32 atomics.
__global__ void __launch_bounds__(SBLOCK) k_adj_pixel(const Paths P, float* pix_grad, int frames) {
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= P.npix) return;
    for (int k = 0; k < ADJ_N; k++) {
        float acc = pix_grad[(size_t)k * P.npix + pix];
        for (int f = 0; f < frames; f++) acc += P.adj_G[(size_t)k * P.cap + (size_t)f * P.npix + pix];
        pix_grad[(size_t)k * P.npix + pix] = acc;
    }
}
__global__ void __launch_bounds__(SBLOCK) k_adj_final(const float* pix_grad, int npix, float* grad) {
    __shared__ float red[SBLOCK];
    const int k = blockIdx.x;
    float acc = 0.0f;
    for (int p = threadIdx.x; p < npix; p += SBLOCK) acc += pix_grad[(size_t)k * npix + p];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = SBLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) grad[k] = red[0];
}
__global__ void __launch_bounds__(SBLOCK) k_adj_atomic(const Paths P, int n_paths, float* grad) {
    __shared__ float red[SBLOCK];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    for (int k = 0; k < ADJ_N; k++) {
        red[threadIdx.x] = i < n_paths ? P.adj_G[(size_t)k * P.cap + i] : 0.0f;
        __syncthreads();
        for (int w = SBLOCK / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) atomicAdd(&grad[k], red[0]);
        __syncthreads();
    }
}

// film: average_update (components/accumulator/average/average.glsl:6-10) of each frame's
// sample in sample order; spectral samples go through color_to_pixel
// (color/spectral/spectral.glsl:27-33) with the cie1931 sensor first.
template <bool SPEC>
__global__ void __launch_bounds__(SBLOCK) k_accumulate(const DevScene S, const Paths P, float4* film, int npix, int frames,
                                                       uint32_t sample0, int wl_uniform, float4* var, float4* aov0,
                                                       float4* aov1) {
    int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    // the guide pictures: average_update of their samples, in sample order
    float4* aovf[2] = {aov0, aov1};
    for (int a = 0; a < 2; a++) {
        if (!aovf[a] || !P.aov[a]) continue;
        float4 g = aovf[a][pix];
        for (int f = 0; f < frames; f++) {
            const uint32_t n = sample0 + (uint32_t)f;
            const float4 x = P.aov[a][(size_t)f * npix + pix];
            if (n == 0) g = make_float4(x.x, x.y, x.z, 1.0f);
            else {
                float w = 1.0f / ((float)n + 1.0f);
                g = make_float4(g.x * (1.0f - w) + x.x * w, g.y * (1.0f - w) + x.y * w, g.z * (1.0f - w) + x.z * w, 1.0f);
            }
        }
        aovf[a][pix] = g;
    }
    float4 acc = film[pix];
    float4 v = var ? var[pix] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int f = 0; f < frames; f++) {
        uint32_t n = sample0 + (uint32_t)f;
        const size_t i = (size_t)f * npix + pix;
        float4 x = P.rad[i];
        if (SPEC) {
            float4 lam = P.wl[i];
            f3 px = mk3(0.0f, 0.0f, 0.0f);
            for (int k = 0; k < 4; k++) {
                float l = lane(lam, k);
                px = px + sensor_response(S, l) * (lane(x, k) / wavelength_pdf(wl_uniform != 0, l));
            }
            px = px * 0.25f;
            x = make_float4(px.x, px.y, px.z, 0.0f);
        }
        const float4 prev = acc;
        if (n == 0) acc = make_float4(x.x, x.y, x.z, 1.0f);
        else {
            float w = 1.0f / ((float)n + 1.0f);
            acc = make_float4(acc.x * (1.0f - w) + x.x * w, acc.y * (1.0f - w) + x.y * w, acc.z * (1.0f - w) + x.z * w, 1.0f);
        }
        if (var) {   // variance_update (components/accumulator/variance/variance.glsl:10-16)
            if (n == 0) v = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
            else {
                float w = 1.0f / ((float)n + 1.0f);
                float dx = (x.x - prev.x) * (x.x - acc.x), dy = (x.y - prev.y) * (x.y - acc.y), dz = (x.z - prev.z) * (x.z - acc.z);
                v = make_float4(v.x * (1.0f - w) + dx * w, v.y * (1.0f - w) + dy * w, v.z * (1.0f - w) + dz * w, 1.0f);
            }
        }
    }
    film[pix] = acc;
    if (var) var[pix] = v;
}

template <int STACK, bool COUNT>
__global__ void __launch_bounds__(TBLOCK) k_trace_closest(DevScene S, const float* rays, int n, float* t_out, int* prim_out,
                                                          unsigned long long* trav) {
    __shared__ uint2 stk[STACK * TBLOCK];
    TraceCounters cnt = {0, 0};
    for (int base = blockIdx.x * TBLOCK; base < n; base += gridDim.x * TBLOCK) {
        int i = base + threadIdx.x;
        if (i >= n) break;
        f3 o = mk3(rays[i], rays[n + i], rays[2 * (size_t)n + i]);
        f3 d = mk3(rays[3 * (size_t)n + i], rays[4 * (size_t)n + i], rays[5 * (size_t)n + i]);
        float t0 = rays[6 * (size_t)n + i], tmax = t0; int prim = -1;
        trace_wide<false, STACK, TBLOCK, COUNT>(S, o, d, tmax, prim, stk, cnt);
        trace_analytic<false>(S, o, d, tmax, prim);
        trace_instances<false, TBLOCK, COUNT>(S, o, d, tmax, prim, stk, cnt);
        t_out[i] = tmax;
        prim_out[i] = prim;
    }
    if (COUNT) {
        atomicAdd(&trav[0], (unsigned long long)cnt.nodes);
        atomicAdd(&trav[1], (unsigned long long)cnt.tris);
    }
}

template <int STACK>
__global__ void __launch_bounds__(TBLOCK) k_trace_any(DevScene S, const float* rays, int n, uint8_t* hit_out) {
    __shared__ uint2 stk[STACK * TBLOCK];
    TraceCounters cnt = {0, 0};
    for (int base = blockIdx.x * TBLOCK; base < n; base += gridDim.x * TBLOCK) {
        int i = base + threadIdx.x;
        if (i >= n) break;
        f3 o = mk3(rays[i], rays[n + i], rays[2 * (size_t)n + i]);
        f3 d = mk3(rays[3 * (size_t)n + i], rays[4 * (size_t)n + i], rays[5 * (size_t)n + i]);
        float tmax = rays[6 * (size_t)n + i]; int prim = -1;
        bool h = trace_analytic<true>(S, o, d, tmax, prim);
        if (!h) h = trace_wide<true, STACK, TBLOCK, false>(S, o, d, tmax, prim, stk, cnt);
        if (!h) h = trace_instances<true, TBLOCK, false>(S, o, d, tmax, prim, stk, cnt);
        hit_out[i] = h ? 1 : 0;
    }
}

}  // namespace

// ============================================================================ host side

namespace {
thread_local std::string g_err;
pt_status fail(pt_status s, const std::string& msg) { g_err = msg; return s; }

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(e_ == hipErrorOutOfMemory ? PT_E_OOM : PT_E_DEVICE,                  \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                 \
    } while (0)

// ---------------------------------------------------------------- environment bake
// The procedural sky's importance table on the device. The reference renders the environment
// into a W x H equirect target (compiler/EnvironmentBake.ts:29-80) and builds the CDFs on the CPU
// (components/env/importance.ts:16-127). Here every step runs on the GPU, in the order and
// precision of the host builder (builder.cpp build_env_table), so the tables are bit-identical:
// texels and blur taps are independent (one thread each), and the sums the reference
// accumulates sequentially in f64 stay sequential (one thread per row for the conditional CDFs,
// one thread for the marginal CDF and for the compensation mean).
// Per-texel kernels are grid-stride (the launch grid is capped at MAX_GRID blocks).
__global__ void __launch_bounds__(256) k_env_luma(float* Y, int W, int H, float sx, float sy, float sz, int oct) {
    for (int k = blockIdx.x * 256 + threadIdx.x; k < W * H; k += gridDim.x * 256) {
        int i = k % W, j = k / W;
        float u = ((float)i + 0.5f) / (float)W, v = ((float)j + 0.5f) / (float)H;
        float dx, dy, dz;                                   // env_chart_dir, the bake's rotation 0
        if (oct) ptm_oct_dir(u, v, 1.0f, 0.0f, &dx, &dy, &dz);
        else {
            float phi = (u - 0.5f) * PTM_TWO_PI - 0.0f;
            float theta = v * PTM_PI;
            float st = ptm_sin(theta);
            dx = ptm_cos(phi) * st; dy = ptm_cos(theta); dz = ptm_sin(phi) * st;
        }
        float s = 0.5f * (dy + 1.0f);
        float kk = ptm_smoothstep(0.9994f, 0.9999f, dx * sx + dy * sy + dz * sz);
        float cx = 0.06f * (1.0f - s) + 0.35f * s + 18.0f * kk;
        float cy = 0.08f * (1.0f - s) + 0.45f * s + 16.0f * kk;
        float cz = 0.12f * (1.0f - s) + 0.65f * s + 13.0f * kk;
        Y[k] = (float)(0.2126 * (double)cx + 0.7152 * (double)cy + 0.0722 * (double)cz);
    }
}

// 3 x 3 tent, u wraps (the phi seam), v clamps (the poles).
__global__ void __launch_bounds__(256) k_env_blur(const float* __restrict__ Y, float* __restrict__ B, int W, int H, int oct) {
    for (int k = blockIdx.x * 256 + threadIdx.x; k < W * H; k += gridDim.x * 256) {
        int i = k % W, j = k / W;
        double sum = 0.0, wsum = 0.0;
        for (int dj = -1; dj <= 1; dj++) {
            int jj = min(H - 1, max(0, j + dj));
            for (int di = -1; di <= 1; di++) {
                int ii = oct ? min(W - 1, max(0, i + di)) : (i + di + W) % W;
                double w = (dj == 0 ? 2 : 1) * (di == 0 ? 2 : 1);
                sum += w * (double)Y[(size_t)jj * W + ii];
                wsum += w;
            }
        }
        B[k] = (float)(sum / wsum);
    }
}

// MIS compensation mean: sum_j sum_i B dOmega_j / sum dOmega, accumulated by thread 0 in texel
// order; the block stages whole rows through LDS, 8192 texels (32 KB) at a time. One block of
// 256 threads; W <= PT_ENV_MAX_W.
constexpr int PT_ENV_MAX_W = 4096;
constexpr int PT_ENV_STAGE = 8192;
__global__ void __launch_bounds__(256) k_env_mean(const float* __restrict__ B, const double* __restrict__ dO, int W, int H,
                                                  double* mean) {
    __shared__ __align__(16) float stage[PT_ENV_STAGE];
    const int rows = PT_ENV_STAGE / W;                   // >= 2 rows per stage
    double num = 0.0, den = 0.0;
    for (int j0 = 0; j0 < H; j0 += rows) {
        const int nr = min(rows, H - j0);
        for (int i = threadIdx.x; i < nr * W; i += 256) stage[i] = B[(size_t)j0 * W + i];
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int r = 0; r < nr; r++) {
                double d = dO[j0 + r];
                const float* row = stage + r * W;
                int i = 0;
                // 16 texels per step: four 16-byte LDS reads issued ahead of the dependent adds
                // (rows start 16-byte aligned when W is a multiple of 4)
                for (; (W & 3) == 0 && i + 16 <= W; i += 16) {
                    float4 q[4];
                    for (int m = 0; m < 4; m++) q[m] = *(const float4*)(row + i + 4 * m);
                    for (int m = 0; m < 4; m++) {
                        num += (double)q[m].x * d; den += d;
                        num += (double)q[m].y * d; den += d;
                        num += (double)q[m].z * d; den += d;
                        num += (double)q[m].w * d; den += d;
                    }
                }
                for (; i < W; i++) { num += (double)row[i] * d; den += d; }
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *mean = den > 0.0 ? num / den : 0.0;
}

__global__ void __launch_bounds__(256) k_env_compensate(float* B, int n, const double* mean) {
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
        double t = (double)B[k] - *mean;
        B[k] = (float)(t > 0 ? t : 0);
    }
}

// Conditional CDF of row j (thread j) and the row's mass.
__global__ void __launch_bounds__(64) k_env_rows(const float* __restrict__ B, const double* __restrict__ sinT, int W, int H,
                                                 float* __restrict__ cond, double* __restrict__ rows) {
    for (int j = blockIdx.x * 64 + threadIdx.x; j < H; j += gridDim.x * 64) {
        const float* r = B + (size_t)j * W;
        double st = sinT[j], sum = 0.0;
        for (int i = 0; i < W; i++) { double y = (double)r[i]; sum += (0.0 < y ? y : 0.0) * st; }
        rows[j] = sum;
        double c = 0.0;
        for (int i = 0; i < W; i++) {
            double y = (double)r[i];
            c += (0.0 < y ? y : 0.0) * st;
            cond[(size_t)j * W + i] = (float)(sum > 0 ? c / sum : 0);
        }
    }
}

// Marginal CDF over rows and the table's total mass (the env.totalWeight parameter).
__global__ void k_env_marginal(const double* __restrict__ rows, int H, float* __restrict__ marg, double* total_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double total = 0.0;
    for (int j = 0; j < H; j++) total += rows[j];
    double acc = 0.0;
    for (int j = 0; j < H; j++) { acc += rows[j]; marg[j] = (float)(total > 0 ? acc / total : 0); }
    *total_out = total;
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

int grid_for(long long n, int block) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return (int)std::min<long long>(g, MAX_GRID);
}
}  // namespace

struct TimingSlot { std::vector<std::pair<hipEvent_t, hipEvent_t>> ev; float ms = 0; uint32_t launches = 0; uint64_t items = 0; };

// One uploaded scene: its device buffers and the DevScene view of them.
struct SceneStore {
    std::vector<std::unique_ptr<DevBuf>> bufs;
    DevScene dev{};
    int stack = 16;
    uint32_t wide_depth = 0;
    bool has_ggx = false;
    bool has_mirror = false;
    bool has_rough = false;
    std::vector<pt_material> host_materials;
    hipStream_t stream = nullptr;
    template <class T> pt_status upload(const T* host, size_t n, const T** out) {
        auto b = std::make_unique<DevBuf>();
        b->bytes = std::max<size_t>(n, 1) * sizeof(T);
        HIP_TRY(hipMalloc(&b->p, b->bytes));
        if (n) HIP_TRY(hipMemcpyAsync(b->p, host, n * sizeof(T), hipMemcpyHostToDevice, stream));
        *out = (const T*)b->p;
        bufs.push_back(std::move(b));
        return PT_OK;
    }
    template <class T> pt_status alloc(size_t n, T** out) {
        auto b = std::make_unique<DevBuf>();
        b->bytes = std::max<size_t>(n, 1) * sizeof(T);
        HIP_TRY(hipMalloc(&b->p, b->bytes));
        *out = (T*)b->p;
        bufs.push_back(std::move(b));
        return PT_OK;
    }
};

struct pt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<DevBuf>> scene_bufs;
    DevScene dev{};
    bool has_scene = false;
    int stack = 16;
    uint32_t wide_depth = 0;
    int persist_blocks = 2048;   // persistent traversal grid (set from the device's CU count)
    int cu_count = 256;
    int shade_blocks = 2048;
    bool has_ggx = false;
    bool has_mirror = false;
    bool has_rough = false;
    std::vector<pt_material> host_materials;   // the uploaded records (adjoint checks, updates)
    bool timing = false;
    std::map<std::string, TimingSlot> timings;
    std::vector<std::string> timing_order;
    DevBuf trav;   // 2 x u64 counters for instrumented traces
    DevBuf rays_tmp, t_tmp, p_tmp;

    template <class T> pt_status upload(const T* host, size_t n, const T** out) {
        auto b = std::make_unique<DevBuf>();
        b->bytes = std::max<size_t>(n, 1) * sizeof(T);
        HIP_TRY(hipMalloc(&b->p, b->bytes));
        if (n) HIP_TRY(hipMemcpyAsync(b->p, host, n * sizeof(T), hipMemcpyHostToDevice, stream));
        *out = (const T*)b->p;
        scene_bufs.push_back(std::move(b));
        return PT_OK;
    }
    void ev_begin(const char* name, hipEvent_t* b) {
        if (!timing) return;
        (void)hipEventCreate(b);
        (void)hipEventRecord(*b, stream);
        (void)name;
    }
    void ev_end(const char* name, hipEvent_t b, uint64_t items) {
        if (!timing) return;
        hipEvent_t e;
        (void)hipEventCreate(&e);
        (void)hipEventRecord(e, stream);
        auto it = timings.find(name);
        if (it == timings.end()) { timing_order.push_back(name); it = timings.emplace(name, TimingSlot()).first; }
        it->second.ev.emplace_back(b, e);
        it->second.launches++;
        it->second.items += items;
    }
};

struct pt_renderer {
    pt_ctx* ctx = nullptr;
    pt_program prog{};
    DevBuf film, var_film, aov_film[2];
    uint32_t sample_count = 0;
    long long cap = 0;    // path slots allocated
    DevBuf b_ray_o, b_ray_d, b_thr, b_rad, b_rng, b_prev, b_flags, b_medium, b_hit, b_q0, b_q1, b_counts,
        b_spath, b_so, b_sd, b_sc, b_sl, b_wl, b_eta, b_hist, b_trav, b_cat[N_CAT], b_aov[2];
    // adjoint buffers (allocated by the first pt_backward)
    DevBuf b_adjD, b_adjG, b_adjsD, b_delta, b_pixgrad, b_grad;
    DevBuf b_pscene;      // scene of each path slot (batched wavefront)
    long long adj_cap = 0;
    int adj_mode = PT_REDUCE_DETERMINISTIC;
    Paths paths{};
    bool count_traversal = false;
};

#define TIMED(ctx, name, items, launch)             \
    do {                                            \
        hipEvent_t ev_b_ = nullptr;                 \
        (ctx)->ev_begin(name, &ev_b_);              \
        launch;                                     \
        (ctx)->ev_end(name, ev_b_, items);          \
    } while (0)

extern "C" {

const char* pt_last_error(void) { return g_err.c_str(); }

double pt_env_total_weight(const pt_env* env) {
    if (!env || env->type != PT_ENV_SKY) return 0.0;
    std::vector<float> cond, marg;
    double total = 0.0;
    build_env_table(*env, cond, marg, total);
    return total;
}

pt_status pt_program_from_description(const char* json, size_t len, int32_t width, int32_t height, int32_t frames_in_flight,
                                      pt_program* out) {
    if (!json || !out) return fail(PT_E_INVALID, "pt_program_from_description: null argument");
    std::string err;
    pt_status st = ptmi_lower_description(json, len, width, height, frames_in_flight, out, err);
    return st == PT_OK ? PT_OK : fail(st, err);
}

pt_status pt_renderer_load_description(pt_ctx* ctx, const char* json, size_t len, int32_t width, int32_t height,
                                       int32_t frames_in_flight, pt_renderer** out) {
    pt_program pg;
    pt_status st = pt_program_from_description(json, len, width, height, frames_in_flight, &pg);
    return st == PT_OK ? pt_renderer_load(ctx, &pg, out) : st;
}

pt_status pt_obj_parse(const char* text, size_t len, uint32_t flags, pt_mesh_data** out) {
    if (!out || (!text && len)) return fail(PT_E_INVALID, "pt_obj_parse: null argument");
    *out = nullptr;
    auto* m = new pt_mesh_data();
    std::string err;
    pt_status st = ptmi_parse_obj(text ? text : "", len, flags, *m, err);
    if (st != PT_OK) { delete m; return fail(st, err); }
    *out = m;
    return PT_OK;
}

pt_status pt_mesh_data_info(const pt_mesh_data* m, uint32_t* n_vertices, uint32_t* n_indices, int32_t* has_normals,
                            int32_t* has_uvs) {
    if (!m) return fail(PT_E_INVALID, "pt_mesh_data_info: null mesh");
    if (n_vertices) *n_vertices = (uint32_t)(m->positions.size() / 3);
    if (n_indices) *n_indices = (uint32_t)m->indices.size();
    if (has_normals) *has_normals = m->has_normals;
    if (has_uvs) *has_uvs = m->has_uvs;
    return PT_OK;
}

pt_status pt_mesh_data_copy(const pt_mesh_data* m, float* positions, uint32_t* indices, float* normals, float* uvs) {
    if (!m) return fail(PT_E_INVALID, "pt_mesh_data_copy: null mesh");
    if (positions) std::memcpy(positions, m->positions.data(), m->positions.size() * 4);
    if (indices) std::memcpy(indices, m->indices.data(), m->indices.size() * 4);
    if (normals && m->has_normals) std::memcpy(normals, m->normals.data(), m->normals.size() * 4);
    if (uvs && m->has_uvs) std::memcpy(uvs, m->uvs.data(), m->uvs.size() * 4);
    return PT_OK;
}

void pt_mesh_data_free(pt_mesh_data* m) { delete m; }

pt_status pt_inst_parse(const void* buffer, size_t len, pt_inst_table* out) {
    if (!out || (!buffer && len)) return fail(PT_E_INVALID, "pt_inst_parse: null argument");
    std::string err;
    static const uint8_t empty = 0;
    pt_status st = ptmi_parse_inst(buffer ? (const uint8_t*)buffer : &empty, len, *out, err);
    return st == PT_OK ? PT_OK : fail(st, err);
}

pt_status pt_finish(pt_ctx* ctx) {
    if (!ctx) return fail(PT_E_INVALID, "pt_finish: null ctx");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PT_OK;
}

pt_status pt_ctx_create(int device_index, pt_ctx** out) {
    if (!out) return fail(PT_E_INVALID, "pt_ctx_create: out is null");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device_index < 0 || device_index >= n) return fail(PT_E_INVALID, "pt_ctx_create: no such device");
    HIP_TRY(hipSetDevice(device_index));
    auto* c = new pt_ctx();
    c->device = device_index;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return fail(PT_E_DEVICE, "stream creation failed"); }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device_index) == hipSuccess && prop.multiProcessorCount > 0) {
        c->persist_blocks = prop.multiProcessorCount * 8;   // 16 waves per CU of 2-wave blocks
        c->cu_count = prop.multiProcessorCount;
        c->shade_blocks = prop.multiProcessorCount * 8;
    }
    c->trav.bytes = 2 * sizeof(unsigned long long);
    if (hipMalloc(&c->trav.p, c->trav.bytes) != hipSuccess) { delete c; return fail(PT_E_OOM, "counter allocation failed"); }
    *out = c;
    return PT_OK;
}

void pt_ctx_destroy(pt_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->timings)
        for (auto& pr : kv.second.ev) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
    ctx->scene_bufs.clear();
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

// The delta-tracking majorant of a medium record (features/materials.ts:774-789
// derivedMajorant: max(1e-6, max_c(sigma_a + sigma_s)) on the live values), as the RGB colour
// model keeps it (components/color/rgb/rgb.ts:5) and as the spectral model bounds it: the float32
// step above 4 c (1 + 2^-20)^5 / (1 - 2^-20)^2 (components/color/spectral/majorant.ts).
static void derive_majorants(pt_material& m) {
    double c = 1e-6;
    for (int k = 0; k < 3; k++) c = std::max(c, (double)m.sigma_a[k] + (double)m.sigma_s[k]);
    m.majorant[0] = (float)c;
    const double ceiling = (double)(float)c;
    const double up = 1.0 + std::ldexp(1.0, -20), down = 1.0 - std::ldexp(1.0, -20);
    const double bound = 4.0 * ceiling * up * up * up * up * up / (down * down);
    m.majorant[1] = std::nextafter((float)bound, INFINITY);
}

// Depth of a binary BVH (nodes of 8 floats: min.xyz, count or -1; max.xyz, offset or right
// child); the device walks keep a 64-entry private stack.
static int binary_depth(const std::vector<float>& nodes) {
    if (nodes.empty()) return 0;
    std::vector<std::pair<int, int>> todo{{0, 1}};
    int depth = 0;
    while (!todo.empty()) {
        auto [n, d] = todo.back(); todo.pop_back();
        depth = std::max(depth, d);
        if (nodes[8 * (size_t)n + 3] < 0.0f) { todo.push_back({n + 1, d + 1}); todo.push_back({(int)nodes[8 * (size_t)n + 7], d + 1}); }
    }
    return depth;
}

static pt_status validate_scene(const pt_scene_desc* sd) {
    if (!sd) return fail(PT_E_INVALID, "pt_scene_upload: null argument");
    if (sd->abi_version != PT_ABI_VERSION) return fail(PT_E_INVALID, "pt_scene_upload: ABI version mismatch");
    if (sd->n_tris < 0 || sd->n_objects < 0 || sd->n_materials <= 0 || sd->n_lights < 0)
        return fail(PT_E_INVALID, "pt_scene_upload: negative count or no materials");
    if (sd->n_tris > 0 && (!sd->tri_vertices || !sd->tri_object)) return fail(PT_E_MISSING_INPUT, "missing channel: triangles");
    if (!sd->materials) return fail(PT_E_MISSING_INPUT, "missing channel: materials");
    for (int m = 0; m < sd->n_materials; m++)
        if (sd->materials[m].has_medium && (sd->materials[m].medium_arm < PT_MED_ANALYTIC || sd->materials[m].medium_arm > PT_MED_DELTA))
            return fail(PT_E_INVALID, "material " + std::to_string(m) + ": unknown medium arm");
    if (sd->n_objects > 0 && !sd->objects) return fail(PT_E_MISSING_INPUT, "missing channel: objects");
    if (!sd->materials) return fail(PT_E_MISSING_INPUT, "missing channel: materials");
    if (sd->n_lights > 0 && !sd->lights) return fail(PT_E_MISSING_INPUT, "missing channel: lights");
    for (int t = 0; t < sd->n_tris; t++) {
        int o = sd->tri_object[t];
        if (o < 0 || o >= sd->n_objects || sd->objects[o].kind != PT_OBJ_MESH)
            return fail(PT_E_INVALID, "triangle " + std::to_string(t) + " has an invalid owning object");
    }
    const int n_proto = sd->proto_vertices ? sd->n_proto_tris : 0;
    const int n_inst = sd->instances ? sd->n_instances : 0;
    if (n_proto < 0 || n_inst < 0) return fail(PT_E_INVALID, "pt_scene_upload: negative prototype or instance count");
    std::vector<int> owner((size_t)std::max(n_inst, 1), -1);
    for (int o = 0; o < sd->n_objects; o++) {
        const pt_object& ob = sd->objects[o];
        if (ob.kind != PT_OBJ_INSTANCES) continue;
        if (ob.tri_count <= 0 || ob.first_tri < 0 || ob.first_tri + ob.tri_count > n_proto)
            return fail(PT_E_INVALID, "object " + std::to_string(o) + ": prototype triangle range out of bounds");
        if (ob.instance_count <= 0 || ob.first_instance < 0 || ob.first_instance + ob.instance_count > n_inst)
            return fail(PT_E_INVALID, "object " + std::to_string(o) + ": instance range out of bounds");
        for (int i = ob.first_instance; i < ob.first_instance + ob.instance_count; i++) {
            if (owner[i] >= 0) return fail(PT_E_INVALID, "instance " + std::to_string(i) + " belongs to two batches");
            owner[i] = o;
            if (!(sd->instances[i].ts[3] > 0.0f)) return fail(PT_E_INVALID, "instance " + std::to_string(i) + ": scale must be > 0");
        }
    }
    for (int i = 0; i < n_inst; i++)
        if (owner[i] < 0) return fail(PT_E_INVALID, "instance " + std::to_string(i) + " belongs to no batch");
    for (int o = 0; o < sd->n_objects; o++) {
        const pt_object& ob = sd->objects[o];
        if (ob.kind < PT_OBJ_MESH || ob.kind > PT_OBJ_INSTANCES) return fail(PT_E_INVALID, "object " + std::to_string(o) + ": unknown kind");
        if (ob.smooth && ((ob.kind == PT_OBJ_MESH && !sd->tri_normals) || (ob.kind == PT_OBJ_INSTANCES && !sd->proto_normals)))
            return fail(PT_E_MISSING_INPUT, "missing channel: normals (object " + std::to_string(o) + " is smooth)");
        if (ob.smooth && ob.kind != PT_OBJ_MESH && ob.kind != PT_OBJ_INSTANCES)
            return fail(PT_E_INVALID, "object " + std::to_string(o) + ": only meshes shade smooth");
        if (ob.material < 0 || ob.material >= sd->n_materials) return fail(PT_E_INVALID, "object " + std::to_string(o) + ": bad material");
        if (ob.kind == PT_OBJ_MESH && (ob.first_tri < 0 || ob.tri_count < 0 || ob.first_tri + ob.tri_count > sd->n_tris))
            return fail(PT_E_INVALID, "object " + std::to_string(o) + ": triangle range out of bounds");
    }
    return PT_OK;
}

// Build every device structure of one scene (wide BVH, containment, env table) into SS.
// The procedural sky's importance table, baked on the device into SS (k_env_* above). The f64
// row weights sin(theta_j) and texel solid angles are the host builder's expressions, evaluated
// on the host (H values each). ev0 / ev1, when given, bracket the kernels on SS.stream.
static pt_status bake_env_table(const pt_env& env, SceneStore& SS, DevScene& D, hipEvent_t ev0 = nullptr,
                                hipEvent_t ev1 = nullptr) {
    const bool oct = env.chart == PT_CHART_OCTAHEDRAL;
    const int H = env.table_h > 0 ? env.table_h : 256, W = oct ? H : (env.table_w > 0 ? env.table_w : 512);
    if (W > PT_ENV_MAX_W) return fail(PT_E_UNSUPPORTED, "environment table wider than 4096 texels");
    const long long n = (long long)W * H;
    std::vector<double> sin_t((size_t)H), d_omega((size_t)H);
    for (int j = 0; j < H; j++) {
        sin_t[j] = oct ? 1.0 : std::sin(M_PI * (j + 0.5) / H);
        d_omega[j] = oct ? (4 * M_PI) / (W * H) : (2 * M_PI / W) * (M_PI / H) * std::sin(M_PI * (j + 0.5) / H);
    }
    float sx = 0.4f, sy = 0.35f, sz = 0.2f;               // the sun direction, normalised in f32
    float d = sx * sx + sy * sy + sz * sz;
    float inv = 1.0f / sqrtf(d);
    sx *= inv; sy *= inv; sz *= inv;
    pt_status st;
    float *Y, *B, *cond, *marg;
    double *rows, *aux;
    const double *d_sin, *d_dom;
    if ((st = SS.alloc((size_t)n, &Y)) != PT_OK || (st = SS.alloc((size_t)n, &B)) != PT_OK ||
        (st = SS.alloc((size_t)n, &cond)) != PT_OK || (st = SS.alloc((size_t)H, &marg)) != PT_OK ||
        (st = SS.alloc((size_t)H, &rows)) != PT_OK || (st = SS.alloc(2, &aux)) != PT_OK ||
        (st = SS.upload(sin_t.data(), sin_t.size(), &d_sin)) != PT_OK ||
        (st = SS.upload(d_omega.data(), d_omega.size(), &d_dom)) != PT_OK)
        return st;
    if (ev0) HIP_TRY(hipEventRecord(ev0, SS.stream));
    hipLaunchKernelGGL(k_env_luma, dim3(grid_for(n, 256)), dim3(256), 0, SS.stream, Y, W, H, sx, sy, sz, oct ? 1 : 0);
    hipLaunchKernelGGL(k_env_blur, dim3(grid_for(n, 256)), dim3(256), 0, SS.stream, (const float*)Y, B, W, H, oct ? 1 : 0);
    if (env.compensated) {
        hipLaunchKernelGGL(k_env_mean, dim3(1), dim3(256), 0, SS.stream, (const float*)B, d_dom, W, H, aux);
        hipLaunchKernelGGL(k_env_compensate, dim3(grid_for(n, 256)), dim3(256), 0, SS.stream, B, (int)n,
                           (const double*)aux);
    }
    hipLaunchKernelGGL(k_env_rows, dim3(grid_for(H, 64)), dim3(64), 0, SS.stream, (const float*)B, d_sin, W, H, cond, rows);
    hipLaunchKernelGGL(k_env_marginal, dim3(1), dim3(64), 0, SS.stream, (const double*)rows, H, marg, aux + 1);
    if (ev1) HIP_TRY(hipEventRecord(ev1, SS.stream));
    HIP_TRY(hipGetLastError());
    D.env_cond = cond;
    D.env_marg = marg;
    D.env_total = aux + 1;
    return PT_OK;
}

static pt_status build_scene(const pt_scene_desc* sd, hipStream_t stream, SceneStore& SS) {
    SS.stream = stream;
    DevScene& D = SS.dev;
    D = DevScene{};
    // wide BVH over every mesh triangle
    WideBVH wb;
    try { build_wide_bvh(sd->tri_vertices, sd->n_tris, wb); }
    catch (const std::exception& e) { return fail(PT_E_INVALID, std::string("wide BVH build: ") + e.what()); }
    SS.wide_depth = wb.max_depth;
    if (wb.max_depth + 1 <= 16) SS.stack = 16;
    else if (wb.max_depth + 1 <= 32) SS.stack = 32;
    else return fail(PT_E_UNSUPPORTED, "wide BVH deeper than 31 levels");
    pt_status st;
    if ((st = SS.upload((const uint4*)wb.nodes.data(), wb.nodes.size() / 4, &D.cw_nodes)) != PT_OK) return st;
    if ((st = SS.upload((const float4*)wb.tris.data(), wb.tris.size() / 4, &D.cw_tris)) != PT_OK) return st;
    D.n_cw_nodes = (int)wb.n_nodes;
    D.cw_depth = (int)wb.max_depth;
    const int n_proto = sd->proto_vertices ? sd->n_proto_tris : 0;
    const size_t n_shade = (size_t)sd->n_tris + (size_t)n_proto;
    std::vector<float> shade(n_shade * 12);
    for (size_t t = 0; t < n_shade; t++) {
        const float* src = t < (size_t)sd->n_tris ? sd->tri_vertices + 9 * t : sd->proto_vertices + 9 * (t - sd->n_tris);
        for (int v = 0; v < 3; v++) {
            for (int a = 0; a < 3; a++) shade[12 * t + 4 * v + a] = src[3 * v + a];
            shade[12 * t + 4 * v + 3] = 0.0f;
        }
    }
    if ((st = SS.upload((const float4*)shade.data(), n_shade * 3, &D.tri_shade)) != PT_OK) return st;
    if (sd->tri_normals || sd->proto_normals) {
        std::vector<float> nrm(n_shade * 9, 0.0f);
        if (sd->tri_normals) std::copy(sd->tri_normals, sd->tri_normals + 9 * (size_t)sd->n_tris, nrm.begin());
        if (sd->proto_normals && n_proto)
            std::copy(sd->proto_normals, sd->proto_normals + 9 * (size_t)n_proto, nrm.begin() + 9 * (size_t)sd->n_tris);
        if ((st = SS.upload(nrm.data(), nrm.size(), &D.tri_nrm)) != PT_OK) return st;
    }
    if ((st = SS.upload(sd->tri_object, (size_t)sd->n_tris, &D.tri_object)) != PT_OK) return st;
    D.n_tris = sd->n_tris;
    if ((st = SS.upload(sd->objects, (size_t)sd->n_objects, &D.objects)) != PT_OK) return st;
    D.n_objects = sd->n_objects;
    std::vector<int> analytic;
    for (int o = 0; o < sd->n_objects; o++)
        if (sd->objects[o].kind == PT_OBJ_QUAD || sd->objects[o].kind == PT_OBJ_SPHERE) analytic.push_back(o);
    if ((st = SS.upload(analytic.data(), analytic.size(), &D.analytic)) != PT_OK) return st;
    D.n_analytic = (int)analytic.size();
    SS.host_materials.assign(sd->materials, sd->materials + sd->n_materials);
    for (pt_material& m : SS.host_materials) {
        derive_majorants(m);
        if (m.has_medium && m.medium_arm == PT_MED_ANALYTIC) D.weighted_absorption = 1;
    }
    if ((st = SS.upload(SS.host_materials.data(), SS.host_materials.size(), &D.materials)) != PT_OK) return st;
    if ((st = SS.upload(sd->lights, (size_t)sd->n_lights, &D.lights)) != PT_OK) return st;
    D.n_lights = sd->n_lights;
    SS.has_ggx = false;
    SS.has_mirror = false;
    SS.has_rough = false;
    for (int m = 0; m < sd->n_materials; m++) {
        if (transmissive_model(sd->materials[m].model)) D.has_transmission = 1;
        if (sd->materials[m].model == PT_MAT_ROUGH_DIELECTRIC) SS.has_rough = true;
        if (sd->materials[m].model == PT_MAT_GGX) SS.has_ggx = true;
        if (sd->materials[m].model == PT_MAT_MIRROR) SS.has_mirror = true;
        if (sd->materials[m].has_medium) D.has_media = 1;
        if (sd->materials[m].model == PT_MAT_NULL) D.has_null = 1;
    }
    if (sd->ambient_material >= sd->n_materials) return fail(PT_E_INVALID, "pt_scene_upload: ambient_material out of range");
    D.ambient_material = sd->ambient_material < 0 ? -1 : sd->ambient_material;
    if (sd->spectral_curves) {
        if ((st = SS.upload((const float4*)sd->spectral_curves, 95, &D.spec_curves)) != PT_OK) return st;
        for (int k = 0; k < 9; k++) D.xyz_rgb[k] = sd->xyz_to_rgb[k];
        D.cie_y = sd->cie_y_integral;
    }
    for (int o = 0; o < sd->n_objects; o++) {
        const pt_object& ob = sd->objects[o];
        bool solid = ob.kind == PT_OBJ_SPHERE || ((ob.kind == PT_OBJ_MESH || ob.kind == PT_OBJ_INSTANCES) && ob.closed);
        if (solid && sd->materials[ob.material].emissive) D.emissive_solid = 1;
    }
    // containment: per closed mesh (and closed instance batch) binary BVH, closest-triangle distance.
    // Triangle ids are shade indices: world triangles first, prototype triangles after n_tris.
    auto obj_tris = [&](const pt_object& ob) {
        return ob.kind == PT_OBJ_INSTANCES ? sd->proto_vertices + 9 * (size_t)ob.first_tri
                                           : sd->tri_vertices + 9 * (size_t)ob.first_tri;
    };
    auto obj_shade0 = [&](const pt_object& ob) { return ob.kind == PT_OBJ_INSTANCES ? sd->n_tris + ob.first_tri : ob.first_tri; };
    std::vector<float> onodes;
    std::vector<int> obase((size_t)std::max(sd->n_objects, 1), -1), oorder(std::max(n_shade, (size_t)1), 0);
    std::vector<int> titems, wide_items;
    for (int o = 0; o < sd->n_objects; o++) {
        const pt_object& ob = sd->objects[o];
        const bool batch = ob.kind == PT_OBJ_INSTANCES;
        if (!(ob.kind == PT_OBJ_MESH || batch) || ob.tri_count == 0) continue;
        if (batch || ob.closed) wide_items.push_back(o);
        if (!ob.closed) continue;
        if (!batch) titems.push_back(o);
        const float* tv = obj_tris(ob);
        std::vector<double> boxes(6 * (size_t)ob.tri_count);
        for (int i = 0; i < ob.tri_count; i++) {
            const float* p = tv + 9 * (size_t)i;
            for (int a = 0; a < 3; a++) {
                boxes[6 * i + a] = std::min(std::min(p[a], p[3 + a]), p[6 + a]);
                boxes[6 * i + 3 + a] = std::max(std::max(p[a], p[3 + a]), p[6 + a]);
            }
        }
        BinaryBVH bb;
        build_binary_bvh(boxes.data(), ob.tri_count, 2, bb);
        if (binary_depth(bb.nodes) >= 60) return fail(PT_E_UNSUPPORTED, "object " + std::to_string(o) + ": containment BVH too deep");
        obase[o] = (int)(onodes.size() / 8);
        onodes.insert(onodes.end(), bb.nodes.begin(), bb.nodes.end());
        const int s0 = obj_shade0(ob);
        for (int i = 0; i < ob.tri_count; i++) oorder[s0 + i] = s0 + (int)bb.order[i];
    }
    if ((st = SS.upload((const float4*)onodes.data(), onodes.size() / 4, &D.obj_nodes)) != PT_OK) return st;
    if ((st = SS.upload(obase.data(), obase.size(), &D.obj_node_base)) != PT_OK) return st;
    if ((st = SS.upload(oorder.data(), oorder.size(), &D.obj_order)) != PT_OK) return st;
    // per closed mesh / instance batch wide BVH (inside probe, prototype walk), concatenated
    std::vector<uint32_t> own; std::vector<float> owt; std::vector<int> oroot((size_t)std::max(sd->n_objects, 1), -1);
    uint32_t obj_depth = 0;
    for (int o : wide_items) {
        const pt_object& ob = sd->objects[o];
        WideBVH w;
        try { build_wide_bvh(obj_tris(ob), ob.tri_count, w); }
        catch (const std::exception& e) { return fail(PT_E_INVALID, std::string("object BVH build: ") + e.what()); }
        uint32_t noff = (uint32_t)(own.size() / 20), toff = (uint32_t)(owt.size() / 12);
        for (uint32_t k = 0; k < w.n_nodes; k++) { w.nodes[20 * k + 4] += noff; w.nodes[20 * k + 5] += toff; }
        const uint32_t s0 = (uint32_t)obj_shade0(ob);
        for (size_t k = 0; k < w.tris.size() / 12; k++)
            w.tris[12 * k + 3] = ptm_u2f(s0 + ptm_f2u(w.tris[12 * k + 3]));
        own.insert(own.end(), w.nodes.begin(), w.nodes.end());
        owt.insert(owt.end(), w.tris.begin(), w.tris.end());
        oroot[o] = (int)noff;
        obj_depth = std::max(obj_depth, w.max_depth);
    }
    if (obj_depth + 1 > (uint32_t)SS.stack) {
        if (obj_depth + 1 <= 32) SS.stack = 32;
        else return fail(PT_E_UNSUPPORTED, "object BVH deeper than 31 levels");
    }
    if ((st = SS.upload((const uint4*)own.data(), own.size() / 4, &D.ow_nodes)) != PT_OK) return st;
    if ((st = SS.upload((const float4*)owt.data(), owt.size() / 4, &D.ow_tris)) != PT_OK) return st;
    if ((st = SS.upload(oroot.data(), oroot.size(), &D.ow_root)) != PT_OK) return st;
    // placements (features/intersection.ts:981-1066): per instance the world->local rotation and
    // (translation, scale), owning batch, first primitive offset; a binary TLAS over the world
    // boxes of the prototype's local box (f64, padded so float ray transforms stay inside)
    const int n_inst = sd->instances ? sd->n_instances : 0;
    D.inst_prim0 = sd->n_tris + sd->n_objects;
    D.n_instances = n_inst;
    if (n_inst > 0) {
        std::vector<float4> iq(n_inst), its(n_inst);
        std::vector<int> ibatch(n_inst, -1), ibase(n_inst, 0);
        for (int o = 0; o < sd->n_objects; o++) {
            const pt_object& ob = sd->objects[o];
            if (ob.kind != PT_OBJ_INSTANCES) continue;
            for (int i = ob.first_instance; i < ob.first_instance + ob.instance_count; i++) ibatch[i] = o;
        }
        long long acc = 0;
        for (int i = 0; i < n_inst; i++) {
            const pt_instance& I = sd->instances[i];
            iq[i] = make_float4(I.q[0], I.q[1], I.q[2], I.q[3]);
            its[i] = make_float4(I.ts[0], I.ts[1], I.ts[2], I.ts[3]);
            ibase[i] = (int)acc;
            acc += sd->objects[ibatch[i]].tri_count;
        }
        if ((long long)D.inst_prim0 + acc >= 0x7fffffffLL) return fail(PT_E_UNSUPPORTED, "instanced primitive ids exceed 31 bits");
        std::vector<double> ib(6 * (size_t)n_inst);
        for (int i = 0; i < n_inst; i++) {
            const pt_instance& I = sd->instances[i];
            const pt_object& ob = sd->objects[ibatch[i]];
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            const double ux = -(double)I.q[0], uy = -(double)I.q[1], uz = -(double)I.q[2], w = I.q[3], sc = I.ts[3];
            for (int c = 0; c < 8; c++) {
                double v[3];
                for (int k = 0; k < 3; k++) v[k] = sc * (double)((c >> k) & 1 ? ob.box_max[k] : ob.box_min[k]) - (double)I.ts[k];
                // local -> world: rotate s * p - t by conj(q), v + 2 u x (u x v + w v)
                const double c1[3] = {uy * v[2] - uz * v[1] + w * v[0], uz * v[0] - ux * v[2] + w * v[1], ux * v[1] - uy * v[0] + w * v[2]};
                const double r[3] = {v[0] + 2.0 * (uy * c1[2] - uz * c1[1]), v[1] + 2.0 * (uz * c1[0] - ux * c1[2]),
                                     v[2] + 2.0 * (ux * c1[1] - uy * c1[0])};
                for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], r[k]); hi[k] = std::max(hi[k], r[k]); }
            }
            const double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
            for (int k = 0; k < 3; k++) {
                const double m = 1e-5 * ext + 1e-30;
                ib[6 * i + k] = lo[k] - std::max(m, 1e-6 * std::fabs(lo[k]));
                ib[6 * i + 3 + k] = hi[k] + std::max(m, 1e-6 * std::fabs(hi[k]));
            }
        }
        BinaryBVH tl;
        build_binary_bvh(ib.data(), n_inst, 1, tl);
        if (binary_depth(tl.nodes) >= 60) return fail(PT_E_UNSUPPORTED, "instance TLAS too deep");
        std::vector<int> iorder(tl.order.begin(), tl.order.end());
        if ((st = SS.upload((const float4*)tl.nodes.data(), tl.nodes.size() / 4, &D.inst_nodes)) != PT_OK) return st;
        D.n_inst_nodes = (int)(tl.nodes.size() / 8);
        if ((st = SS.upload(iorder.data(), iorder.size(), &D.inst_order)) != PT_OK) return st;
        if ((st = SS.upload(iq.data(), iq.size(), &D.inst_q)) != PT_OK) return st;
        if ((st = SS.upload(its.data(), its.size(), &D.inst_ts)) != PT_OK) return st;
        if ((st = SS.upload(ibatch.data(), ibatch.size(), &D.inst_batch)) != PT_OK) return st;
        if ((st = SS.upload(ibase.data(), ibase.size(), &D.inst_base)) != PT_OK) return st;
    }
    // uniform grid over the closed-mesh boxes (containment candidates)
    D.has_grid = titems.empty() ? 0 : 1;
    if (D.has_grid) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        std::vector<float> ext;
        for (int o : titems) {
            const pt_object& ob = sd->objects[o];
            float e = 0.0f;
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], ob.box_min[a]); hi[a] = std::max(hi[a], ob.box_max[a]);
                e = std::max(e, ob.box_max[a] - ob.box_min[a]);
            }
            ext.push_back(e);
        }
        std::nth_element(ext.begin(), ext.begin() + ext.size() / 2, ext.end());
        float cell = std::max(ext[ext.size() / 2], 1e-6f);
        for (int a = 0; a < 3; a++) {
            float e = hi[a] - lo[a];
            int n = (int)std::ceil(e / cell);
            n = std::min(std::max(n, 1), 128);
            D.grid_lo[a] = lo[a];
            D.grid_n[a] = n;
            D.grid_inv[a] = e > 0.0f ? (float)n / e : 0.0f;
        }
        auto cell_of = [&](int a, float x) {
            int c = (int)std::floor((x - D.grid_lo[a]) * D.grid_inv[a]);
            return std::min(std::max(c, 0), D.grid_n[a] - 1);
        };
        const size_t ncell = (size_t)D.grid_n[0] * D.grid_n[1] * D.grid_n[2];
        std::vector<int2> cells(ncell, make_int2(0, 0));
        std::vector<std::vector<int>> lists(ncell);
        for (int o : titems) {
            const pt_object& ob = sd->objects[o];
            int c0[3], c1[3];
            for (int a = 0; a < 3; a++) { c0[a] = cell_of(a, ob.box_min[a]); c1[a] = cell_of(a, ob.box_max[a]); }
            for (int z = c0[2]; z <= c1[2]; z++)
                for (int y = c0[1]; y <= c1[1]; y++)
                    for (int x = c0[0]; x <= c1[0]; x++) lists[((size_t)z * D.grid_n[1] + y) * D.grid_n[0] + x].push_back(o);
        }
        std::vector<int> items;
        for (size_t c = 0; c < ncell; c++) {
            cells[c] = make_int2((int)items.size(), (int)lists[c].size());
            items.insert(items.end(), lists[c].begin(), lists[c].end());
        }
        if ((st = SS.upload(cells.data(), cells.size(), &D.grid_cells)) != PT_OK) return st;
        if ((st = SS.upload(items.data(), items.size(), &D.grid_items)) != PT_OK) return st;
    }
    // environment
    D.env_type = sd->env.type;
    if (sd->env.type == PT_ENV_SKY) {
        if ((st = bake_env_table(sd->env, SS, D)) != PT_OK) return st;
        D.env_chart = sd->env.chart == PT_CHART_OCTAHEDRAL ? PT_CHART_OCTAHEDRAL : PT_CHART_EQUIRECT;
        D.env_h = sd->env.table_h > 0 ? sd->env.table_h : 256;
        D.env_w = D.env_chart == PT_CHART_OCTAHEDRAL ? D.env_h : (sd->env.table_w > 0 ? sd->env.table_w : 512);
        D.env_compensated = sd->env.compensated;
        D.env_intensity = sd->env.intensity;
        D.env_rotation = sd->env.rotation;
        D.env_rot_c = (float)std::cos((double)sd->env.rotation);
        D.env_rot_s = (float)std::sin((double)sd->env.rotation);
        D.env_select_prob = sd->env.select_prob;
    } else if (sd->env.type == PT_ENV_CONSTANT) {
        D.env_intensity = sd->env.intensity;
        D.env_color = make_float4(sd->env.color[0], sd->env.color[1], sd->env.color[2], 0.0f);
        D.env_s_color = sd->env.s_color;
        D.env_samplable = sd->env.samplable ? 1 : 0;
        D.env_select_prob = sd->env.select_prob;
    } else if (sd->env.type != PT_ENV_NONE) {
        return fail(PT_E_UNSUPPORTED, "environment type not provided by this build");
    }
    HIP_TRY(hipStreamSynchronize(SS.stream));
    return PT_OK;
}

pt_status pt_scene_upload(pt_ctx* ctx, const pt_scene_desc* sd) {
    if (!ctx) return fail(PT_E_INVALID, "pt_scene_upload: null argument");
    pt_status st = validate_scene(sd);
    if (st != PT_OK) return st;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->scene_bufs.clear();
    ctx->has_scene = false;
    ctx->dev = DevScene{};
    SceneStore SS;
    if ((st = build_scene(sd, ctx->stream, SS)) != PT_OK) return st;
    ctx->scene_bufs = std::move(SS.bufs);
    ctx->dev = SS.dev;
    ctx->stack = SS.stack;
    ctx->wide_depth = SS.wide_depth;
    ctx->has_ggx = SS.has_ggx;
    ctx->has_mirror = SS.has_mirror;
    ctx->has_rough = SS.has_rough;
    ctx->host_materials = std::move(SS.host_materials);
    ctx->has_scene = true;
    return PT_OK;
}

pt_status pt_renderer_load(pt_ctx* ctx, const pt_program* prog, pt_renderer** out) {
    if (!ctx || !prog || !out) return fail(PT_E_INVALID, "pt_renderer_load: null argument");
    if (prog->abi_version != PT_ABI_VERSION) return fail(PT_E_INVALID, "pt_renderer_load: ABI version mismatch");
    if (prog->estimator < PT_EST_PT || prog->estimator > PT_EST_MIS) return fail(PT_E_UNSUPPORTED, "unknown estimator");
    if (prog->color != PT_COLOR_RGB && prog->color != PT_COLOR_SPECTRAL) return fail(PT_E_UNSUPPORTED, "unknown color model");
    if (prog->wavelengths != PT_WL_VISIBLE && prog->wavelengths != PT_WL_UNIFORM) return fail(PT_E_UNSUPPORTED, "unknown wavelength sampler");
    if (prog->medium_light_sampling != PT_MLS_VERTEX && prog->medium_light_sampling != PT_MLS_EQUIANGULAR)
        return fail(PT_E_UNSUPPORTED, "unknown medium light sampling");
    if (prog->volume_sampling != PT_VS_ANALYTIC && prog->volume_sampling != PT_VS_DELTA_TRACKING)
        return fail(PT_E_UNSUPPORTED, "unknown volume sampling");
    if (prog->medium_light_sampling == PT_MLS_EQUIANGULAR && prog->estimator != PT_EST_NEE)
        return fail(PT_E_UNSUPPORTED, "equiangular medium NEE runs under the nee estimator");
    if (prog->color == PT_COLOR_SPECTRAL && ctx->has_scene && ctx->dev.env_type == PT_ENV_SKY)
        return fail(PT_E_UNSUPPORTED, "spectral programs with the procedural sky are not provided by this build");
    if (prog->width <= 0 || prog->height <= 0 || prog->max_bounces < 0) return fail(PT_E_INVALID, "bad film size or bounce budget");
    if (prog->aovs & ~(PT_AOV_ALBEDO | PT_AOV_NORMAL)) return fail(PT_E_INVALID, "unknown aov bits");
    if (prog->aovs && prog->color != PT_COLOR_RGB)
        return fail(PT_E_UNSUPPORTED, "first-hit guides (aovs) are provided for RGB programs only");
    if (prog->env_chart < 0 || prog->env_chart > PT_CHART_OCTAHEDRAL + 1) return fail(PT_E_INVALID, "unknown env chart");
    if (prog->env_chart && ctx->has_scene && ctx->dev.env_type == PT_ENV_SKY && prog->env_chart - 1 != ctx->dev.env_chart)
        return fail(PT_E_INVALID, std::string("the program's envSampler.chart is '") +
                                      (prog->env_chart - 1 == PT_CHART_OCTAHEDRAL ? "octahedral" : "equirect") +
                                      "' but the scene's sky table was built on the other chart (pt_env.chart)");
    HIP_TRY(hipSetDevice(ctx->device));
    auto* r = new pt_renderer();
    r->ctx = ctx;
    r->prog = *prog;
    if (r->prog.frames_in_flight < 1) r->prog.frames_in_flight = 1;
    r->film.bytes = (size_t)prog->width * prog->height * sizeof(float4);
    if (hipMalloc(&r->film.p, r->film.bytes) != hipSuccess) { delete r; return fail(PT_E_OOM, "film allocation failed"); }
    if (prog->variance) {
        r->var_film.bytes = r->film.bytes;
        if (hipMalloc(&r->var_film.p, r->var_film.bytes) != hipSuccess) { delete r; return fail(PT_E_OOM, "film allocation failed"); }
    }
    for (int a = 0; a < 2; a++) {
        if (!(prog->aovs & (1 << a))) continue;
        r->aov_film[a].bytes = r->film.bytes;
        if (hipMalloc(&r->aov_film[a].p, r->aov_film[a].bytes) != hipSuccess) { delete r; return fail(PT_E_OOM, "film allocation failed"); }
        (void)hipMemsetAsync(r->aov_film[a].p, 0, r->aov_film[a].bytes, ctx->stream);
    }
    (void)hipMemsetAsync(r->film.p, 0, r->film.bytes, ctx->stream);
    *out = r;
    return PT_OK;
}

void pt_renderer_unload(pt_renderer* r) {
    if (!r) return;
    (void)hipSetDevice(r->ctx->device);
    (void)hipStreamSynchronize(r->ctx->stream);
    delete r;
}

}  // extern "C"

static pt_status ensure_paths(pt_renderer* r, long long n) {
    if (r->cap >= n) return PT_OK;
    struct Entry { DevBuf* b; size_t elem; };
    std::vector<Entry> list = {
        {&r->b_ray_o, 16}, {&r->b_ray_d, 16}, {&r->b_thr, 16}, {&r->b_rad, 16}, {&r->b_rng, 8}, {&r->b_prev, 16},
        {&r->b_flags, 4}, {&r->b_hit, 8}, {&r->b_q0, 4}, {&r->b_q1, 4}, {&r->b_spath, 4}, {&r->b_so, 16},
        {&r->b_sd, 16}, {&r->b_sc, 16}, {&r->b_sl, 16}, {&r->b_medium, 4}, {&r->b_wl, 16}, {&r->b_eta, 4}, {&r->b_pscene, 4}};
    for (int c = 0; c < N_CAT; c++) list.push_back({&r->b_cat[c], 4});   // one queue per shading category
    for (int a = 0; a < 2; a++)
        if (r->prog.aovs & (1 << a)) list.push_back({&r->b_aov[a], 16});  // first-hit guide samples
    for (auto& e : list) {
        if (e.b->p) { HIP_TRY(hipFree(e.b->p)); e.b->p = nullptr; }
        e.b->bytes = e.elem * (size_t)n;
        HIP_TRY(hipMalloc(&e.b->p, e.b->bytes));
    }
    if (!r->b_counts.p) {
        r->b_counts.bytes = N_COUNTS * sizeof(int);
        HIP_TRY(hipMalloc(&r->b_counts.p, r->b_counts.bytes));
        HIP_TRY(hipMemset(r->b_counts.p, 0, r->b_counts.bytes));
    }
    if (!r->b_hist.p) { r->b_hist.bytes = 2 * sizeof(unsigned long long); HIP_TRY(hipMalloc(&r->b_hist.p, r->b_hist.bytes)); }
    if (!r->b_trav.p) { r->b_trav.bytes = 2 * sizeof(unsigned long long); HIP_TRY(hipMalloc(&r->b_trav.p, r->b_trav.bytes)); }
    Paths& P = r->paths;
    P.ray_o = (float4*)r->b_ray_o.p; P.ray_d = (float4*)r->b_ray_d.p; P.thr = (float4*)r->b_thr.p;
    P.rad = (float4*)r->b_rad.p; P.rng = (uint2*)r->b_rng.p; P.prev = (float4*)r->b_prev.p;
    P.flags = (int*)r->b_flags.p; P.hit = (float2*)r->b_hit.p; P.q[0] = (int*)r->b_q0.p; P.q[1] = (int*)r->b_q1.p;
    P.counts = (int*)r->b_counts.p; P.s_path = (int*)r->b_spath.p; P.s_o = (float4*)r->b_so.p;
    P.s_d = (float4*)r->b_sd.p; P.s_c = (float4*)r->b_sc.p; P.s_l = (float4*)r->b_sl.p;
    P.medium = (int*)r->b_medium.p; P.wl = (float4*)r->b_wl.p; P.eta = (float*)r->b_eta.p;
    P.path_scene = (int*)r->b_pscene.p; P.hist = (unsigned long long*)r->b_hist.p;
    P.trav = (unsigned long long*)r->b_trav.p;
    for (int a = 0; a < 2; a++) P.aov[a] = (r->prog.aovs & (1 << a)) ? (float4*)r->b_aov[a].p : nullptr;
    for (int c = 0; c < N_CAT; c++) {
        P.catq[c] = (int*)r->b_cat[c].p;
        if (!P.catq[c]) return fail(PT_E_OOM, "ensure_paths: a shading queue was not allocated");
    }
    P.cap = (int)n;
    P.npix = r->prog.width * r->prog.height;
    r->cap = n;
    return PT_OK;
}

// Dynamic LDS for a k_shade launch: the region probe's stack for the surface categories
// without a static one, in scenes whose emissive solids make them probe at back-face hits.
template <int CAT, bool MEDIA, int STACK>
static size_t shade_dynamic_lds(bool emissive_solid) {
    constexpr bool STATIC_STK = cat_probes(CAT, MEDIA) && CAT != CAT_MISS && CAT != CAT_MEDIUM;
    constexpr bool SURFACE = CAT != CAT_MISS && CAT != CAT_MEDIUM;
    return (!STATIC_STK && SURFACE && emissive_solid) ? (size_t)STACK * SBLOCK * sizeof(uint2) : 0;
}

template <int EST, int STACK, bool MEDIA, bool SPEC, bool ADJ>
static void launch_shade(pt_renderer* r, int cur) {
    pt_ctx* ctx = r->ctx;
    const pt_program& pg = r->prog;
    const DevScene& S = ctx->dev;
    hipStream_t st = ctx->stream;
    const int g = ctx->shade_blocks;
    const int mn = pg.max_null_crossings;
#define PT_SHADE(name, CATV) TIMED(ctx, name, 0, hipLaunchKernelGGL((k_shade<EST, CATV, STACK, MEDIA, SPEC, ADJ>), dim3(g), dim3(SBLOCK), (shade_dynamic_lds<CATV, MEDIA, STACK>(S.emissive_solid != 0)), st, S, r->paths, cur, pg.max_bounces, pg.rr_start, pg.rr_max_survival, mn, pg.medium_light_sampling))
    PT_SHADE("shade_miss", CAT_MISS);
    PT_SHADE("shade_lambert", CAT_LAMBERT);
    if (ctx->has_ggx) PT_SHADE("shade_ggx", CAT_GGX);
    if (ctx->has_mirror) PT_SHADE("shade_mirror", CAT_MIRROR);
    if (ctx->has_rough) PT_SHADE("shade_rough_dielectric", CAT_ROUGH);
    if (S.has_transmission) PT_SHADE("shade_dielectric", CAT_DIELECTRIC);
    if (MEDIA) PT_SHADE("shade_medium", CAT_MEDIUM);
    if (S.has_null) PT_SHADE("shade_null", CAT_NULL);
#undef PT_SHADE
}

template <int STACK, bool MEDIA, bool SPEC, bool ADJ>
static void launch_wave(pt_renderer* r, const Camera& cam, int frames, uint32_t sample0) {
    static_assert(!(SPEC && ADJ), "the adjoint runs RGB programs");
    pt_ctx* ctx = r->ctx;
    const DevScene& S = ctx->dev;
    Paths& P = r->paths;
    const pt_program& pg = r->prog;
    const int W = pg.width, H = pg.height;
    const long long n = (long long)W * H * frames;
    hipStream_t st = ctx->stream;
    if (MEDIA) hipLaunchKernelGGL((k_camera_medium<STACK>), dim3(1), dim3(64), 0, st, S, P, cam);
    TIMED(ctx, "raygen", n,  hipLaunchKernelGGL((k_raygen<SPEC, ADJ>), dim3((unsigned)((n + SBLOCK - 1) / SBLOCK)), dim3(SBLOCK), 0, st, P, cam, W, H, (int)n, sample0, MEDIA ? 1 : 0, pg.wavelengths));
    if (MEDIA && (cam.ortho || cam.thin)) hipLaunchKernelGGL((k_ray_medium<STACK>), dim3(ctx->shade_blocks), dim3(SBLOCK), 0, st, S, P, (int)n);
    int cur = 0;
    const int tgrid = ctx->persist_blocks;
    // a null crossing consumes an iteration but no bounce (pt.ts null-interface branch)
    const int iterations = pg.max_bounces + 1 + (S.has_null ? std::max(pg.max_null_crossings, 0) : 0);
    for (int it = 0; it < iterations; it++) {
        hipLaunchKernelGGL(k_bookkeep, dim3(1), dim3(64), 0, st, P, cur, cur ^ 1, it == 0 ? 1 : 0);
        const bool inst = S.n_inst_nodes > 0;
        if (r->count_traversal) {
            if (inst) TIMED(ctx, "extend_count", 0, hipLaunchKernelGGL((k_extend<STACK, true, false, TBLOCK, true>), dim3(tgrid), dim3(TBLOCK), 0, st, S, P, cur));
            else TIMED(ctx, "extend_count", 0, hipLaunchKernelGGL((k_extend<STACK, true>), dim3(tgrid), dim3(TBLOCK), 0, st, S, P, cur));
        } else {
            if (inst) TIMED(ctx, "extend", 0, hipLaunchKernelGGL((k_extend<STACK, false, false, TBLOCK, true>), dim3(tgrid), dim3(TBLOCK), 0, st, S, P, cur));
            else TIMED(ctx, "extend", 0, hipLaunchKernelGGL((k_extend<STACK, false>), dim3(tgrid), dim3(TBLOCK), 0, st, S, P, cur));
        }
        const int eq = MEDIA && !ADJ && pg.medium_light_sampling == PT_MLS_EQUIANGULAR ? 1 : 0;
        TIMED(ctx, "bin", 0, hipLaunchKernelGGL((k_bin<MEDIA, SPEC, ADJ>), dim3(ctx->shade_blocks), dim3(SBLOCK), 0, st, S, P, cur, eq, pg.max_bounces, pg.max_null_crossings));
        if (eq) {
            if (inst) TIMED(ctx, "shadow_media", 0, hipLaunchKernelGGL((k_shadow_media<STACK, SPEC, ADJ, true>), dim3(tgrid), dim3(TBLOCK), 0, st, S, P));
            else TIMED(ctx, "shadow_media", 0, hipLaunchKernelGGL((k_shadow_media<STACK, SPEC, ADJ>), dim3(tgrid), dim3(TBLOCK), 0, st, S, P));
            hipLaunchKernelGGL(k_shadow_reset, dim3(1), dim3(64), 0, st, P);
        }
        switch (pg.estimator) {
            case PT_EST_PT: launch_shade<PT_EST_PT, STACK, MEDIA, SPEC, ADJ>(r, cur); break;
            case PT_EST_NEE: launch_shade<PT_EST_NEE, STACK, MEDIA, SPEC, ADJ>(r, cur); break;
            default: launch_shade<PT_EST_MIS, STACK, MEDIA, SPEC, ADJ>(r, cur);
        }
        if (pg.estimator != PT_EST_PT) {
            if (MEDIA && inst) TIMED(ctx, "shadow_media", 0, hipLaunchKernelGGL((k_shadow_media<STACK, SPEC, ADJ, true>), dim3(tgrid), dim3(TBLOCK), 0, st, S, P));
            else if (MEDIA) TIMED(ctx, "shadow_media", 0, hipLaunchKernelGGL((k_shadow_media<STACK, SPEC, ADJ>), dim3(tgrid), dim3(TBLOCK), 0, st, S, P));
            else if (inst) TIMED(ctx, "shadow", 0, hipLaunchKernelGGL((k_shadow<STACK, SPEC, ADJ, false, TBLOCK, true>), dim3(tgrid), dim3(TBLOCK), 0, st, S, P));
            else TIMED(ctx, "shadow", 0, hipLaunchKernelGGL((k_shadow<STACK, SPEC, ADJ>), dim3(tgrid), dim3(TBLOCK), 0, st, S, P));
        }
        cur ^= 1;
    }
    hipLaunchKernelGGL(k_bookkeep, dim3(1), dim3(64), 0, st, P, -1, -1, 0);
    const int npix = W * H;
    if (ADJ) {
        if (r->adj_mode == PT_REDUCE_DETERMINISTIC)
            TIMED(ctx, "adj_reduce", npix, hipLaunchKernelGGL(k_adj_pixel, dim3((npix + SBLOCK - 1) / SBLOCK), dim3(SBLOCK), 0, st, P, (float*)r->b_pixgrad.p, frames));
        else
            TIMED(ctx, "adj_reduce", n, hipLaunchKernelGGL(k_adj_atomic, dim3((unsigned)((n + SBLOCK - 1) / SBLOCK)), dim3(SBLOCK), 0, st, P, (int)n, (float*)r->b_grad.p));
        return;
    }
    TIMED(ctx, "accumulate", npix, hipLaunchKernelGGL((k_accumulate<SPEC>), dim3((npix + SBLOCK - 1) / SBLOCK), dim3(SBLOCK), 0, st, S, P, (float4*)r->film.p, npix, frames, sample0, pg.wavelengths, (float4*)r->var_film.p, (float4*)r->aov_film[0].p, (float4*)r->aov_film[1].p));
}

extern "C" {

}  // extern "C"

// The camera models pt_params can select, with the thin lens's parameter domain
// (camera/thinlens/thinlens.ts:15-16: aperture nonnegative, focusDistance positive).
static pt_status check_camera(const pt_params* prm) {
    if (prm->camera == PT_CAM_PINHOLE || prm->camera == PT_CAM_ORTHOGRAPHIC || prm->camera == PT_CAM_EQUIRECT) return PT_OK;
    if (prm->camera == PT_CAM_FISHEYE) {
        if (prm->projection < PT_FISHEYE_EQUIDISTANT || prm->projection > PT_FISHEYE_ORTHOGRAPHIC)
            return fail(PT_E_INVALID, "fisheye: unknown projection");
        if (!(prm->fisheye_k > 0.0f) || !std::isfinite(prm->fisheye_k)) return fail(PT_E_INVALID, "fisheye: fisheye_k must be finite and > 0");
        return PT_OK;
    }
    if (prm->camera == PT_CAM_CYLINDRICAL) {
        if (!(prm->cyl_focal > 0.0f) || !std::isfinite(prm->cyl_focal)) return fail(PT_E_INVALID, "cylindrical: cyl_focal must be finite and > 0");
        return PT_OK;
    }
    if (prm->camera != PT_CAM_THINLENS) return fail(PT_E_UNSUPPORTED, "unknown camera");
    if (!(prm->aperture >= 0.0f) || !(prm->focus_distance > 0.0f) || !std::isfinite(prm->aperture) || !std::isfinite(prm->focus_distance))
        return fail(PT_E_INVALID, "thin lens: aperture must be >= 0 and focus_distance > 0");
    return PT_OK;
}

static Camera make_camera(const pt_params* prm) {
    Camera cam;
    cam.ox = prm->cam_position[0]; cam.oy = prm->cam_position[1]; cam.oz = prm->cam_position[2];
    cam.fx = prm->cam_forward[0]; cam.fy = prm->cam_forward[1]; cam.fz = prm->cam_forward[2];
    cam.rx = prm->cam_right[0]; cam.ry = prm->cam_right[1]; cam.rz = prm->cam_right[2];
    cam.ux = prm->cam_up[0]; cam.uy = prm->cam_up[1]; cam.uz = prm->cam_up[2];
    cam.tan_fov = prm->tan_fov; cam.aspect = prm->aspect; cam.salt = prm->reset_salt;
    cam.off_x = prm->pixel_offset[0]; cam.off_y = prm->pixel_offset[1];
    cam.ortho = prm->camera == PT_CAM_ORTHOGRAPHIC ? 1 : 0;
    cam.ortho_scale = prm->ortho_scale;
    cam.thin = prm->camera == PT_CAM_THINLENS ? 1 : 0;
    cam.aperture = prm->aperture;
    cam.focus = prm->focus_distance;
    cam.model = prm->camera;
    cam.projection = prm->projection;
    cam.fk = prm->fisheye_k;
    cam.cyl_f = prm->cyl_focal;
    return cam;
}

// Trace nframes samples (sampleCount sample0, sample0 + 1, ...) in waves of frames_in_flight.
// The forward program accumulates into the film; the adjoint program (ADJ) replays the same
// paths and reduces their gradients instead.
static pt_status run_waves(pt_renderer* r, const pt_params* prm, uint32_t nframes, uint32_t sample0, bool adj) {
    pt_ctx* ctx = r->ctx;
    const pt_program& pg = r->prog;
    const int F = std::max(1, std::min<int>(pg.frames_in_flight, (int)std::max<uint32_t>(nframes, 1)));
    const bool spectral = pg.color == PT_COLOR_SPECTRAL;
    if (spectral && !ctx->dev.spec_curves)
        return fail(PT_E_MISSING_INPUT, "missing input: spectral_curves (the spectral program needs the CIE/D65 curves in pt_scene_desc)");
    HIP_TRY(hipMemsetAsync(r->b_hist.p, 0, r->b_hist.bytes, ctx->stream));
    HIP_TRY(hipMemsetAsync(r->b_trav.p, 0, r->b_trav.bytes, ctx->stream));
    const Camera cam = make_camera(prm);
    uint32_t done = 0;
    while (done < nframes) {
        int f = (int)std::min<uint32_t>((uint32_t)F, nframes - done);
        const uint32_t s0 = sample0 + done;
        const int v = (ctx->stack == 16 ? 0 : 4) | (ctx->dev.has_media ? 2 : 0) | (spectral ? 1 : 0);
        if (adj) {
            switch (v) {
                // the media, spectral and adjoint programs run with the 32-entry stack, which
                // holds every accepted BVH (depth <= 31); only the plain RGB path keeps two depths
                case 0: case 4: launch_wave<32, false, false, true>(r, cam, f, s0); break;
                case 2: case 6: launch_wave<32, true, false, true>(r, cam, f, s0); break;
                default: return fail(PT_E_UNSUPPORTED, "the adjoint runs RGB programs only");
            }
        } else {
            switch (v) {
                case 0: launch_wave<16, false, false, false>(r, cam, f, s0); break;
                case 4: launch_wave<32, false, false, false>(r, cam, f, s0); break;
                case 1: case 5: launch_wave<32, false, true, false>(r, cam, f, s0); break;
                case 2: case 6: launch_wave<32, true, false, false>(r, cam, f, s0); break;
                default: launch_wave<32, true, true, false>(r, cam, f, s0);
            }
        }
        HIP_TRY(hipGetLastError());
        done += (uint32_t)f;
    }
    return PT_OK;
}

extern "C" {

pt_status pt_render_frames(pt_renderer* r, const pt_params* prm, uint32_t nframes) {
    if (!r || !prm) return fail(PT_E_INVALID, "pt_render_frames: null argument");
    pt_ctx* ctx = r->ctx;
    if (!ctx->has_scene) return fail(PT_E_MISSING_INPUT, "missing input: no scene uploaded (pt_scene_upload)");
    pt_status st0;
    if ((st0 = check_camera(prm)) != PT_OK) return st0;
    HIP_TRY(hipSetDevice(ctx->device));
    const pt_program& pg = r->prog;
    const int F = std::max(1, std::min<int>(pg.frames_in_flight, (int)std::max<uint32_t>(nframes, 1)));
    pt_status st = ensure_paths(r, (long long)pg.width * pg.height * F);
    if (st != PT_OK) return st;
    st = run_waves(r, prm, nframes, r->sample_count, false);
    if (st != PT_OK) return st;
    r->sample_count += nframes;
    return PT_OK;
}

pt_status pt_forward(pt_renderer* r, const pt_params* prm, uint32_t nframes) {
    if (!r || !prm) return fail(PT_E_INVALID, "pt_forward: null argument");
    pt_status st = pt_clear_accumulation(r);
    if (st != PT_OK) return st;
    return pt_render_frames(r, prm, nframes);
}

pt_status pt_backward(pt_renderer* r, const pt_params* prm, uint32_t nframes, const pt_adjoint_params* adj,
                      const float* dL_dimage, float* dparams) {
    if (!r || !prm || !adj || !dL_dimage || !dparams) return fail(PT_E_INVALID, "pt_backward: null argument");
    pt_ctx* ctx = r->ctx;
    if (!ctx->has_scene) return fail(PT_E_MISSING_INPUT, "missing input: no scene uploaded (pt_scene_upload)");
    const pt_program& pg = r->prog;
    if (pg.color != PT_COLOR_RGB) return fail(PT_E_UNSUPPORTED, "pt_backward: the adjoint runs RGB programs only");
    if (pt_status sc = check_camera(prm); sc != PT_OK) return sc;
    if (pg.medium_light_sampling != PT_MLS_VERTEX) return fail(PT_E_UNSUPPORTED, "pt_backward: the adjoint runs vertex medium NEE only");
    for (const pt_material& m : ctx->host_materials)
        if (m.has_medium && (m.medium_arm == PT_MED_DELTA || m.medium_emission[0] != 0.0f || m.medium_emission[1] != 0.0f ||
                             m.medium_emission[2] != 0.0f))
            return fail(PT_E_UNSUPPORTED, "pt_backward: emissive media (and the delta-tracking arm) are outside the adjoint");
    if (nframes == 0) return fail(PT_E_INVALID, "pt_backward: nframes must be >= 1");
    if (adj->reduce_mode != PT_REDUCE_ATOMIC && adj->reduce_mode != PT_REDUCE_DETERMINISTIC)
        return fail(PT_E_INVALID, "pt_backward: unknown reduce_mode");
    if (adj->n_albedo < 0 || adj->n_albedo > PT_ADJ_MAX_ALBEDO) return fail(PT_E_INVALID, "pt_backward: n_albedo out of range");
    const std::vector<pt_material>& mats = ctx->host_materials;
    const int nm = (int)mats.size();
    if (adj->medium_material >= nm) return fail(PT_E_INVALID, "pt_backward: medium_material out of range");
    if (adj->medium_material >= 0) {
        const pt_material& m = mats[adj->medium_material];
        if (!m.has_medium) return fail(PT_E_INVALID, "pt_backward: medium_material has no medium");
        // sigma_s = 0 is allowed: such a channel scatters nothing, and its d/dsigma_s holds
        // the extinction term only (the in-scattering term has no sampled path)
        for (int c = 0; c < 3; c++)
            if (!(m.sigma_s[c] >= 0.0f) || !(m.sigma_a[c] >= 0.0f))
                return fail(PT_E_INVALID, "pt_backward: medium parameters need sigma_s, sigma_a >= 0");
        if (!(m.hg_g > -1.0f && m.hg_g < 1.0f)) return fail(PT_E_INVALID, "pt_backward: HG g must be in (-1, 1)");
    }
    for (int k = 0; k < adj->n_albedo; k++) {
        const int a = adj->albedo_material[k];
        if (a < 0 || a >= nm) return fail(PT_E_INVALID, "pt_backward: albedo_material out of range");
        const pt_material& m = mats[a];
        if (m.model != PT_MAT_LAMBERT || !(m.albedo[0] > 0.0f) || m.albedo[0] != m.albedo[1] || m.albedo[1] != m.albedo[2])
            return fail(PT_E_INVALID, "pt_backward: albedo parameters must be Lambert materials with a grey albedo > 0");
    }
    HIP_TRY(hipSetDevice(ctx->device));
    const int F = std::max(1, std::min<int>(pg.frames_in_flight, (int)nframes));
    pt_status st = ensure_paths(r, (long long)pg.width * pg.height * F);
    if (st != PT_OK) return st;
    const size_t npix = (size_t)pg.width * pg.height;
    if (r->adj_cap != r->cap) {
        struct { DevBuf* b; size_t bytes; } list[] = {
            {&r->b_adjD, (size_t)ADJ_N * r->cap * 4}, {&r->b_adjG, (size_t)ADJ_N * r->cap * 4},
            {&r->b_adjsD, (size_t)ADJ_N * r->cap * 4}, {&r->b_delta, npix * 16}, {&r->b_pixgrad, (size_t)ADJ_N * npix * 4},
            {&r->b_grad, (size_t)ADJ_N * 4}};
        for (auto& e : list) {
            if (e.b->p) { HIP_TRY(hipFree(e.b->p)); e.b->p = nullptr; }
            e.b->bytes = e.bytes;
            HIP_TRY(hipMalloc(&e.b->p, e.b->bytes));
        }
        r->adj_cap = r->cap;
    }
    // dL/dI of the film average: each sample enters with weight 1/nframes
    std::vector<float> delta(npix * 4);
    const float inv_n = 1.0f / (float)nframes;
    for (size_t i = 0; i < npix * 4; i++) delta[i] = dL_dimage[i] * inv_n;
    HIP_TRY(hipMemcpyAsync(r->b_delta.p, delta.data(), npix * 16, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemsetAsync(r->b_pixgrad.p, 0, r->b_pixgrad.bytes, ctx->stream));
    HIP_TRY(hipMemsetAsync(r->b_grad.p, 0, r->b_grad.bytes, ctx->stream));
    Paths& P = r->paths;
    P.adj_D = (float*)r->b_adjD.p; P.adj_G = (float*)r->b_adjG.p; P.adj_sD = (float*)r->b_adjsD.p;
    P.adj_delta = (const float4*)r->b_delta.p;
    P.adj_medium = adj->medium_material;
    P.adj_n_albedo = adj->n_albedo;
    for (int k = 0; k < PT_ADJ_MAX_ALBEDO; k++) P.adj_albedo[k] = k < adj->n_albedo ? adj->albedo_material[k] : -1;
    r->adj_mode = adj->reduce_mode;
    st = run_waves(r, prm, nframes, 0, true);
    if (st != PT_OK) return st;
    if (adj->reduce_mode == PT_REDUCE_DETERMINISTIC)
        TIMED(ctx, "adj_final", npix, hipLaunchKernelGGL(k_adj_final, dim3(ADJ_N), dim3(SBLOCK), 0, ctx->stream,
                                                          (const float*)r->b_pixgrad.p, (int)npix, (float*)r->b_grad.p));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(dparams, r->b_grad.p, (size_t)ADJ_N * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- batched wavefront
struct pt_batch {
    pt_ctx* ctx = nullptr;
    std::vector<std::unique_ptr<SceneStore>> stores;
    std::vector<pt_program> progs;
    std::vector<Camera> cams;
    std::vector<long long> film_off;   // first pixel of each scene's film
    long long film_pixels = 0;
    int max_npix = 0, min_stack = 16;
    bool any_ggx = false, any_dielectric = false, any_mirror = false, any_rough = false, any_emissive_solid = false;
    DevBuf d_scenes, d_cams, d_entries, films;
    pt_renderer* r = nullptr;          // path slots
    pt_batch_config cfg{};
    std::vector<BatchEntry> entries;
    std::vector<std::pair<int, int>> waves;   // (first entry, entry count)
    std::vector<long long> wave_paths;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint32_t> graph_spp;
    pt_batch_config graph_cfg{};
    ~pt_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        delete r;
    }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr; graph = nullptr; graph_spp.clear();
    }
};

static bool same_cfg(const pt_batch_config& a, const pt_batch_config& b) {
    return a.wave_paths == b.wave_paths && a.trace_block == b.trace_block && a.trace_blocks_per_cu == b.trace_blocks_per_cu &&
           a.shade_blocks_per_cu == b.shade_blocks_per_cu && a.stack == b.stack && a.use_graph == b.use_graph;
}

// Waves of at most wave_paths slots; each scene contributes whole frames, in sample order.
static pt_status batch_schedule(pt_batch* b, const uint32_t* spp) {
    b->entries.clear(); b->waves.clear(); b->wave_paths.clear();
    const long long cap = b->cfg.wave_paths;
    long long used = 0;
    int first = 0;
    for (size_t s = 0; s < b->stores.size(); s++) {
        const int W = b->progs[s].width, H = b->progs[s].height, npix = W * H;
        uint32_t f = 0;
        while (f < spp[s]) {
            long long nf = std::min<long long>((long long)(spp[s] - f), (cap - used) / npix);
            if (nf <= 0) {
                b->waves.emplace_back(first, (int)b->entries.size() - first);
                b->wave_paths.push_back(used);
                first = (int)b->entries.size(); used = 0;
                continue;
            }
            BatchEntry e;
            e.scene = (int)s; e.W = W; e.H = H; e.npix = npix; e.f0 = (int)f; e.nf = (int)nf;
            e.base = used; e.film = b->film_off[s];
            b->entries.push_back(e);
            used += nf * npix; f += (uint32_t)nf;
        }
    }
    if (used > 0) { b->waves.emplace_back(first, (int)b->entries.size() - first); b->wave_paths.push_back(used); }
    return PT_OK;
}

template <int EST, int STACK>
static void batch_shade(pt_batch* b, int cur) {
    pt_ctx* ctx = b->ctx;
    const pt_program& pg = b->progs[0];
    const DevScene& S0 = b->stores[0]->dev;
    const int g = ctx->cu_count * b->cfg.shade_blocks_per_cu;
    Paths& P = b->r->paths;
    hipStream_t st = ctx->stream;
#define PT_BSHADE(CATV) hipLaunchKernelGGL((k_shade<EST, CATV, STACK, false, false, false, true>), dim3(g), dim3(SBLOCK), (shade_dynamic_lds<CATV, false, STACK>(b->any_emissive_solid)), st, S0, P, cur, pg.max_bounces, pg.rr_start, pg.rr_max_survival, pg.max_null_crossings, 0)
    PT_BSHADE(CAT_MISS);
    PT_BSHADE(CAT_LAMBERT);
    if (b->any_ggx) PT_BSHADE(CAT_GGX);
    if (b->any_mirror) PT_BSHADE(CAT_MIRROR);
    if (b->any_rough) PT_BSHADE(CAT_ROUGH);
    if (b->any_dielectric) PT_BSHADE(CAT_DIELECTRIC);
#undef PT_BSHADE
}

template <int STACK, int TB>
static void batch_waves(pt_batch* b) {
    pt_ctx* ctx = b->ctx;
    const pt_program& pg = b->progs[0];
    const DevScene& S0 = b->stores[0]->dev;
    Paths& P = b->r->paths;
    hipStream_t st = ctx->stream;
    const int tgrid = ctx->cu_count * b->cfg.trace_blocks_per_cu;
    const int sgrid = ctx->cu_count * b->cfg.shade_blocks_per_cu;
    const BatchEntry* dE = (const BatchEntry*)b->d_entries.p;
    for (size_t w = 0; w < b->waves.size(); w++) {
        const int e0 = b->waves[w].first, ne = b->waves[w].second;
        const long long n = b->wave_paths[w];
        hipLaunchKernelGGL(k_raygen_batch, dim3((unsigned)((n + SBLOCK - 1) / SBLOCK)), dim3(SBLOCK), 0, st, P, dE + e0, ne,
                           (const Camera*)b->d_cams.p, (int)n);
        int cur = 0;
        for (int it = 0; it <= pg.max_bounces; it++) {
            hipLaunchKernelGGL(k_bookkeep, dim3(1), dim3(64), 0, st, P, cur, cur ^ 1, it == 0 ? 1 : 0);
            hipLaunchKernelGGL((k_extend<STACK, false, true, TB>), dim3(tgrid), dim3(TB), 0, st, S0, P, cur);
            hipLaunchKernelGGL((k_bin<false, false, false, true>), dim3(sgrid), dim3(SBLOCK), 0, st, S0, P, cur, 0, pg.max_bounces, pg.max_null_crossings);
            switch (pg.estimator) {
                case PT_EST_PT: batch_shade<PT_EST_PT, STACK>(b, cur); break;
                case PT_EST_NEE: batch_shade<PT_EST_NEE, STACK>(b, cur); break;
                default: batch_shade<PT_EST_MIS, STACK>(b, cur);
            }
            if (pg.estimator != PT_EST_PT)
                hipLaunchKernelGGL((k_shadow<STACK, false, false, true, TB>), dim3(tgrid), dim3(TB), 0, st, S0, P);
            cur ^= 1;
        }
        hipLaunchKernelGGL(k_bookkeep, dim3(1), dim3(64), 0, st, P, -1, -1, 0);
        int maxpix = 0;
        for (int e = e0; e < e0 + ne; e++) maxpix = std::max(maxpix, b->entries[e].npix);
        hipLaunchKernelGGL(k_accumulate_batch, dim3((maxpix + SBLOCK - 1) / SBLOCK, ne), dim3(SBLOCK), 0, st, P, dE + e0,
                           (float4*)b->films.p);
    }
}

static void batch_launch(pt_batch* b) {
    hipLaunchKernelGGL(k_fill_f4, dim3(1024), dim3(SBLOCK), 0, b->ctx->stream, (float4*)b->films.p, b->film_pixels);
    const int v = (b->cfg.stack == 16 ? 0 : 3) + (b->cfg.trace_block == 64 ? 0 : (b->cfg.trace_block == 128 ? 1 : 2));
    switch (v) {
        case 0: batch_waves<16, 64>(b); break;
        case 1: batch_waves<16, 128>(b); break;
        case 2: batch_waves<16, 256>(b); break;
        case 3: batch_waves<32, 64>(b); break;
        case 4: batch_waves<32, 128>(b); break;
        default: batch_waves<32, 256>(b);
    }
}

static pt_status batch_check_config(const pt_batch* b, const pt_batch_config& c) {
    if (c.trace_block != 64 && c.trace_block != 128 && c.trace_block != 256) return fail(PT_E_INVALID, "pt_batch: trace_block must be 64, 128 or 256");
    if (c.stack != 16 && c.stack != 32) return fail(PT_E_INVALID, "pt_batch: stack must be 16 or 32");
    if (c.stack < b->min_stack) return fail(PT_E_INVALID, "pt_batch: stack shallower than a scene's BVH needs");
    if (c.wave_paths < b->max_npix || c.wave_paths > (1 << 26)) return fail(PT_E_INVALID, "pt_batch: wave_paths must hold the largest film and be <= 2^26");
    if (c.trace_blocks_per_cu < 1 || c.trace_blocks_per_cu > 32 || c.shade_blocks_per_cu < 1 || c.shade_blocks_per_cu > 32)
        return fail(PT_E_INVALID, "pt_batch: blocks per CU must be in [1, 32]");
    return PT_OK;
}

extern "C" {

pt_status pt_batch_create(pt_ctx* ctx, int32_t n, const pt_scene_desc* scenes, const pt_program* programs,
                          const pt_params* params, pt_batch** out) {
    if (!ctx || !scenes || !programs || !params || !out) return fail(PT_E_INVALID, "pt_batch_create: null argument");
    if (n <= 0) return fail(PT_E_INVALID, "pt_batch_create: no scenes");
    for (int s = 0; s < n; s++)
        if (pt_status sc = check_camera(&params[s]); sc != PT_OK) return sc;
    for (int s = 0; s < n; s++)
        if (programs[s].aovs) return fail(PT_E_UNSUPPORTED, "pt_batch_create: first-hit guides (aovs) are not batched");
    for (int s = 0; s < n; s++)
        if (programs[s].env_chart && scenes[s].env.type == PT_ENV_SKY &&
            programs[s].env_chart - 1 != (scenes[s].env.chart == PT_CHART_OCTAHEDRAL ? PT_CHART_OCTAHEDRAL : PT_CHART_EQUIRECT))
            return fail(PT_E_INVALID, "pt_batch_create: a program's envSampler.chart differs from its scene's sky table chart");
    const pt_program& p0 = programs[0];
    for (int s = 0; s < n; s++) {
        const pt_program& p = programs[s];
        if (p.abi_version != PT_ABI_VERSION || scenes[s].abi_version != PT_ABI_VERSION) return fail(PT_E_INVALID, "pt_batch_create: ABI version mismatch");
        if (p.estimator != p0.estimator || p.max_bounces != p0.max_bounces || p.rr_start != p0.rr_start || p.rr_max_survival != p0.rr_max_survival)
            return fail(PT_E_UNSUPPORTED, "pt_batch_create: every program must share estimator, maxBounces and Russian roulette");
        if (p.color != PT_COLOR_RGB) return fail(PT_E_UNSUPPORTED, "pt_batch_create: batches run RGB programs");
        if (p.width <= 0 || p.height <= 0) return fail(PT_E_INVALID, "pt_batch_create: bad film size");
        pt_status st = validate_scene(&scenes[s]);
        if (st != PT_OK) return st;
        for (int m = 0; m < scenes[s].n_materials; m++)
            if (scenes[s].materials[m].has_medium || scenes[s].materials[m].model == PT_MAT_NULL)
                return fail(PT_E_UNSUPPORTED, "pt_batch_create: batches run scenes without participating media");
        if (scenes[s].instances && scenes[s].n_instances > 0)
            return fail(PT_E_UNSUPPORTED, "pt_batch_create: batches run scenes without instance batches");
    }
    HIP_TRY(hipSetDevice(ctx->device));
    auto* b = new pt_batch();
    b->ctx = ctx;
    std::vector<DevScene> devs;
    for (int s = 0; s < n; s++) {
        auto store = std::make_unique<SceneStore>();
        pt_status st = build_scene(&scenes[s], ctx->stream, *store);
        if (st != PT_OK) { delete b; return st; }
        b->min_stack = std::max(b->min_stack, store->stack);
        b->any_ggx = b->any_ggx || store->has_ggx;
        b->any_mirror = b->any_mirror || store->has_mirror;
        b->any_rough = b->any_rough || store->has_rough;
        b->any_emissive_solid = b->any_emissive_solid || store->dev.emissive_solid != 0;
        b->any_dielectric = b->any_dielectric || store->dev.has_transmission;
        devs.push_back(store->dev);
        b->stores.push_back(std::move(store));
        b->progs.push_back(programs[s]);
        b->cams.push_back(make_camera(&params[s]));
        b->film_off.push_back(b->film_pixels);
        b->film_pixels += (long long)programs[s].width * programs[s].height;
        b->max_npix = std::max(b->max_npix, programs[s].width * programs[s].height);
    }
    struct { DevBuf* d; const void* h; size_t bytes; } up[] = {
        {&b->d_scenes, devs.data(), devs.size() * sizeof(DevScene)}, {&b->d_cams, b->cams.data(), b->cams.size() * sizeof(Camera)}};
    for (auto& u : up) {
        u.d->bytes = u.bytes;
        if (hipMalloc(&u.d->p, u.bytes) != hipSuccess || hipMemcpy(u.d->p, u.h, u.bytes, hipMemcpyHostToDevice) != hipSuccess) {
            delete b; return fail(PT_E_OOM, "pt_batch_create: device allocation failed");
        }
    }
    b->films.bytes = (size_t)b->film_pixels * sizeof(float4);
    if (hipMalloc(&b->films.p, b->films.bytes) != hipSuccess) { delete b; return fail(PT_E_OOM, "pt_batch_create: film allocation failed"); }
    b->r = new pt_renderer();
    b->r->ctx = ctx;
    b->r->prog = p0;
    b->cfg.wave_paths = std::max(b->max_npix, 1 << 23);
    b->cfg.trace_block = 128;
    b->cfg.trace_blocks_per_cu = 8;
    b->cfg.shade_blocks_per_cu = 8;
    b->cfg.stack = b->min_stack;
    b->cfg.use_graph = 1;
    *out = b;
    return PT_OK;
}

void pt_batch_destroy(pt_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
}

pt_status pt_batch_set_config(pt_batch* b, const pt_batch_config* cfg) {
    if (!b || !cfg) return fail(PT_E_INVALID, "pt_batch_set_config: null argument");
    pt_status st = batch_check_config(b, *cfg);
    if (st != PT_OK) return st;
    b->cfg = *cfg;
    return PT_OK;
}

pt_status pt_batch_get_config(const pt_batch* b, pt_batch_config* cfg) {
    if (!b || !cfg) return fail(PT_E_INVALID, "pt_batch_get_config: null argument");
    *cfg = b->cfg;
    return PT_OK;
}

pt_status pt_batch_render(pt_batch* b, const uint32_t* spp) {
    if (!b || !spp) return fail(PT_E_INVALID, "pt_batch_render: null argument");
    pt_ctx* ctx = b->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<uint32_t> sv(spp, spp + b->stores.size());
    const bool reuse = b->cfg.use_graph && b->exec && sv == b->graph_spp && same_cfg(b->cfg, b->graph_cfg);
    if (!reuse) {
        b->drop_graph();
        pt_status st = batch_schedule(b, spp);
        if (st != PT_OK) return st;
        if ((st = ensure_paths(b->r, b->cfg.wave_paths)) != PT_OK) return st;
        b->r->paths.scenes = (const DevScene*)b->d_scenes.p;
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (b->d_entries.p) { HIP_TRY(hipFree(b->d_entries.p)); b->d_entries.p = nullptr; }
        b->d_entries.bytes = std::max<size_t>(b->entries.size(), 1) * sizeof(BatchEntry);
        HIP_TRY(hipMalloc(&b->d_entries.p, b->d_entries.bytes));
        if (!b->entries.empty())
            HIP_TRY(hipMemcpy(b->d_entries.p, b->entries.data(), b->entries.size() * sizeof(BatchEntry), hipMemcpyHostToDevice));
    }
    const bool timing = ctx->timing;
    if (b->cfg.use_graph) {
        if (!reuse) {
            ctx->timing = false;   // no event records inside the captured graph
            HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            batch_launch(b);
            hipError_t e = hipStreamEndCapture(ctx->stream, &b->graph);
            ctx->timing = timing;
            if (e != hipSuccess) return fail(PT_E_DEVICE, std::string("pt_batch_render: graph capture: ") + hipGetErrorString(e));
            HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->graph_spp = sv;
            b->graph_cfg = b->cfg;
        }
        HIP_TRY(hipGraphLaunch(b->exec, ctx->stream));
    } else {
        batch_launch(b);
    }
    HIP_TRY(hipGetLastError());
    return PT_OK;
}

pt_status pt_batch_read_export(pt_batch* b, int32_t scene, float* out, size_t out_floats) {
    if (!b || !out) return fail(PT_E_INVALID, "pt_batch_read_export: null argument");
    if (scene < 0 || scene >= (int32_t)b->stores.size()) return fail(PT_E_INVALID, "pt_batch_read_export: no such scene");
    const size_t need = (size_t)b->progs[scene].width * b->progs[scene].height * 4;
    if (out_floats < need) return fail(PT_E_INVALID, "pt_batch_read_export: output buffer too small");
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    HIP_TRY(hipMemcpy(out, (const float4*)b->films.p + b->film_off[scene], need * sizeof(float), hipMemcpyDeviceToHost));
    return PT_OK;
}

pt_status pt_batch_autotune(pt_batch* b, const uint32_t* spp, uint32_t probe_div, pt_batch_config* best,
                            pt_batch_config* results, float* ms, int32_t max_results, int32_t* n_results) {
    if (!b || !spp) return fail(PT_E_INVALID, "pt_batch_autotune: null argument");
    pt_ctx* ctx = b->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<uint32_t> probe(b->stores.size());
    for (size_t s = 0; s < probe.size(); s++) probe[s] = std::max<uint32_t>(1, spp[s] / std::max<uint32_t>(probe_div, 1));
    const pt_batch_config start = b->cfg;
    pt_batch_config winner = start;
    float best_ms = INFINITY;
    int count = 0;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const int waves[] = {1 << 21, 1 << 22, 1 << 23};
    const int tbs[] = {64, 128, 256};
    const int bpcs[] = {4, 8};
    const int stacks[] = {16, 32};
    for (int wp : waves)
        for (int tb : tbs)
            for (int bpc : bpcs)
                for (int sk : stacks) {
                    pt_batch_config c = start;
                    c.wave_paths = std::max(wp, b->max_npix);
                    c.trace_block = tb;
                    c.trace_blocks_per_cu = std::max(1, bpc * 128 / tb);   // bpc x 128 lanes per CU
                    c.stack = sk;
                    c.use_graph = 1;
                    if (batch_check_config(b, c) != PT_OK) continue;
                    b->cfg = c;
                    pt_status st = pt_batch_render(b, probe.data());       // build + warm
                    if (st != PT_OK) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return st; }
                    HIP_TRY(hipEventRecord(e0, ctx->stream));
                    st = pt_batch_render(b, probe.data());
                    if (st != PT_OK) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return st; }
                    HIP_TRY(hipEventRecord(e1, ctx->stream));
                    HIP_TRY(hipEventSynchronize(e1));
                    float t = 0.0f;
                    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
                    if (results && ms && count < max_results) { results[count] = c; ms[count] = t; }
                    count++;
                    if (t < best_ms) { best_ms = t; winner = c; }
                }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->cfg = winner;
    b->drop_graph();
    if (best) *best = winner;
    if (n_results) *n_results = count;
    return PT_OK;
}

pt_status pt_scene_update_materials(pt_ctx* ctx, const pt_material* materials, int32_t n) {
    if (!ctx || !materials) return fail(PT_E_INVALID, "pt_scene_update_materials: null argument");
    if (!ctx->has_scene) return fail(PT_E_MISSING_INPUT, "missing input: no scene uploaded (pt_scene_upload)");
    if (n != (int32_t)ctx->host_materials.size()) return fail(PT_E_INVALID, "pt_scene_update_materials: material count differs from the scene");
    for (int m = 0; m < n; m++) {
        const pt_material& a = ctx->host_materials[m];
        const pt_material& b = materials[m];
        if (a.model != b.model || a.has_medium != b.has_medium || a.emissive != b.emissive || a.medium_arm != b.medium_arm)
            return fail(PT_E_INVALID, "pt_scene_update_materials: model, has_medium, medium_arm and emissive must not change");
    }
    std::vector<pt_material> next(materials, materials + n);
    for (pt_material& m : next) derive_majorants(m);   // the majorant follows live coefficients
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemcpy((void*)ctx->dev.materials, next.data(), sizeof(pt_material) * (size_t)n, hipMemcpyHostToDevice));
    ctx->host_materials = std::move(next);
    return PT_OK;
}

pt_status pt_read_export(pt_renderer* r, const char* name, float* out, size_t out_floats) {
    if (!r || !name || !out) return fail(PT_E_INVALID, "pt_read_export: null argument");
    const bool var = std::strcmp(name, "variance") == 0;
    const int aov = std::strcmp(name, "albedo") == 0 ? 0 : std::strcmp(name, "normal") == 0 ? 1 : -1;
    if (std::strcmp(name, "radiance") != 0 && !var && aov < 0) return fail(PT_E_INVALID, std::string("unknown export '") + name + "'");
    if (var && !r->var_film.p) return fail(PT_E_MISSING_INPUT, "missing export: 'variance' (compile the program with variance on)");
    if (aov >= 0 && !r->aov_film[aov].p)
        return fail(PT_E_MISSING_INPUT, std::string("missing export: '") + name + "' (list it in the program's aovs)");
    size_t need = (size_t)r->prog.width * r->prog.height * 4;
    if (out_floats < need) return fail(PT_E_INVALID, "pt_read_export: output buffer too small");
    HIP_TRY(hipSetDevice(r->ctx->device));
    const void* src = aov >= 0 ? r->aov_film[aov].p : var ? r->var_film.p : r->film.p;
    HIP_TRY(hipMemcpyAsync(out, src, need * sizeof(float), hipMemcpyDeviceToHost, r->ctx->stream));
    HIP_TRY(hipStreamSynchronize(r->ctx->stream));
    return PT_OK;
}

pt_status pt_clear_accumulation(pt_renderer* r) {
    if (!r) return fail(PT_E_INVALID, "pt_clear_accumulation: null renderer");
    HIP_TRY(hipSetDevice(r->ctx->device));
    HIP_TRY(hipMemsetAsync(r->film.p, 0, r->film.bytes, r->ctx->stream));
    if (r->var_film.p) HIP_TRY(hipMemsetAsync(r->var_film.p, 0, r->var_film.bytes, r->ctx->stream));
    for (int a = 0; a < 2; a++)
        if (r->aov_film[a].p) HIP_TRY(hipMemsetAsync(r->aov_film[a].p, 0, r->aov_film[a].bytes, r->ctx->stream));
    r->sample_count = 0;
    return PT_OK;
}

uint32_t pt_sample_count(const pt_renderer* r) { return r ? r->sample_count : 0; }

pt_status pt_trace_closest_device(pt_ctx* ctx, const float* d_rays, uint32_t n, float* d_t, int32_t* d_prim) {
    if (!ctx || (!d_rays && n)) return fail(PT_E_INVALID, "pt_trace_closest_device: null argument");
    if (!ctx->has_scene) return fail(PT_E_MISSING_INPUT, "missing input: no scene uploaded (pt_scene_upload)");
    HIP_TRY(hipSetDevice(ctx->device));
    if (n == 0) return PT_OK;
    unsigned long long* trav = (unsigned long long*)ctx->trav.p;
    if (ctx->stack == 16)
        TIMED(ctx, "trace_closest", n, hipLaunchKernelGGL((k_trace_closest<16, false>), dim3(grid_for(n, TBLOCK)), dim3(TBLOCK), 0, ctx->stream, ctx->dev, d_rays, (int)n, d_t, d_prim, trav));
    else
        TIMED(ctx, "trace_closest", n, hipLaunchKernelGGL((k_trace_closest<32, false>), dim3(grid_for(n, TBLOCK)), dim3(TBLOCK), 0, ctx->stream, ctx->dev, d_rays, (int)n, d_t, d_prim, trav));
    HIP_TRY(hipGetLastError());
    return PT_OK;
}

pt_status pt_trace_any_device(pt_ctx* ctx, const float* d_rays, uint32_t n, uint8_t* d_hit) {
    if (!ctx || (!d_rays && n)) return fail(PT_E_INVALID, "pt_trace_any_device: null argument");
    if (!ctx->has_scene) return fail(PT_E_MISSING_INPUT, "missing input: no scene uploaded (pt_scene_upload)");
    HIP_TRY(hipSetDevice(ctx->device));
    if (n == 0) return PT_OK;
    if (ctx->stack == 16)
        TIMED(ctx, "trace_any", n, hipLaunchKernelGGL((k_trace_any<16>), dim3(grid_for(n, TBLOCK)), dim3(TBLOCK), 0, ctx->stream, ctx->dev, d_rays, (int)n, d_hit));
    else
        TIMED(ctx, "trace_any", n, hipLaunchKernelGGL((k_trace_any<32>), dim3(grid_for(n, TBLOCK)), dim3(TBLOCK), 0, ctx->stream, ctx->dev, d_rays, (int)n, d_hit));
    HIP_TRY(hipGetLastError());
    return PT_OK;
}

}  // extern "C"

static pt_status stage_rays(pt_ctx* ctx, const pt_rays_soa* rays, uint32_t n) {
    const float* src[7] = {rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmax};
    for (int k = 0; k < 7; k++) if (!src[k]) return fail(PT_E_INVALID, "ray array " + std::to_string(k) + " is null");
    size_t bytes = (size_t)n * 7 * sizeof(float);
    if (ctx->rays_tmp.bytes < bytes) {
        if (ctx->rays_tmp.p) HIP_TRY(hipFree(ctx->rays_tmp.p));
        ctx->rays_tmp.p = nullptr;
        HIP_TRY(hipMalloc(&ctx->rays_tmp.p, bytes));
        ctx->rays_tmp.bytes = bytes;
    }
    for (int k = 0; k < 7; k++)
        HIP_TRY(hipMemcpyAsync((float*)ctx->rays_tmp.p + (size_t)k * n, src[k], (size_t)n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    return PT_OK;
}
static pt_status ensure_buf(DevBuf& b, size_t bytes) {
    if (b.bytes >= bytes) return PT_OK;
    if (b.p) HIP_TRY(hipFree(b.p));
    b.p = nullptr;
    HIP_TRY(hipMalloc(&b.p, bytes));
    b.bytes = bytes;
    return PT_OK;
}

extern "C" {

pt_status pt_trace_closest(pt_ctx* ctx, const pt_rays_soa* rays, uint32_t n, float* t_out, int32_t* prim_out) {
    if (!ctx || !rays || ((!t_out || !prim_out) && n)) return fail(PT_E_INVALID, "pt_trace_closest: null argument");
    if (!ctx->has_scene) return fail(PT_E_MISSING_INPUT, "missing input: no scene uploaded (pt_scene_upload)");
    if (n == 0) return PT_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    pt_status st;
    if ((st = stage_rays(ctx, rays, n)) != PT_OK) return st;
    if ((st = ensure_buf(ctx->t_tmp, (size_t)n * 4)) != PT_OK) return st;
    if ((st = ensure_buf(ctx->p_tmp, (size_t)n * 4)) != PT_OK) return st;
    if ((st = pt_trace_closest_device(ctx, (const float*)ctx->rays_tmp.p, n, (float*)ctx->t_tmp.p, (int32_t*)ctx->p_tmp.p)) != PT_OK) return st;
    HIP_TRY(hipMemcpyAsync(t_out, ctx->t_tmp.p, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(prim_out, ctx->p_tmp.p, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PT_OK;
}

pt_status pt_trace_any(pt_ctx* ctx, const pt_rays_soa* rays, uint32_t n, uint8_t* hit_out) {
    if (!ctx || !rays || (!hit_out && n)) return fail(PT_E_INVALID, "pt_trace_any: null argument");
    if (!ctx->has_scene) return fail(PT_E_MISSING_INPUT, "missing input: no scene uploaded (pt_scene_upload)");
    if (n == 0) return PT_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    pt_status st;
    if ((st = stage_rays(ctx, rays, n)) != PT_OK) return st;
    if ((st = ensure_buf(ctx->p_tmp, (size_t)n * 4)) != PT_OK) return st;
    if ((st = pt_trace_any_device(ctx, (const float*)ctx->rays_tmp.p, n, (uint8_t*)ctx->p_tmp.p)) != PT_OK) return st;
    HIP_TRY(hipMemcpyAsync(hit_out, ctx->p_tmp.p, n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PT_OK;
}

pt_status pt_set_timing(pt_ctx* ctx, int enabled) {
    if (!ctx) return fail(PT_E_INVALID, "pt_set_timing: null ctx");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (auto& kv : ctx->timings)
        for (auto& pr : kv.second.ev) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
    ctx->timings.clear();
    ctx->timing_order.clear();
    ctx->timing = enabled != 0;
    return PT_OK;
}

int pt_get_timings(pt_ctx* ctx, pt_kernel_timing* out, int n_slots) {
    if (!ctx) return -1;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return -1;
    int k = 0;
    for (const std::string& name : ctx->timing_order) {
        TimingSlot& s = ctx->timings[name];
        for (auto& pr : s.ev) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) s.ms += ms;
            (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second);
        }
        s.ev.clear();
        if (out && k < n_slots) {
            std::memset(&out[k], 0, sizeof(pt_kernel_timing));
            std::strncpy(out[k].name, name.c_str(), sizeof(out[k].name) - 1);
            out[k].total_ms = s.ms;
            out[k].launches = s.launches;
            out[k].items = s.items;
        }
        k++;
    }
    return k;
}

pt_status pt_last_ray_counts(pt_renderer* r, uint64_t* closest, uint64_t* any) {
    if (!r) return fail(PT_E_INVALID, "pt_last_ray_counts: null renderer");
    unsigned long long h[2] = {0, 0};
    if (r->b_hist.p) {
        HIP_TRY(hipMemcpyAsync(h, r->b_hist.p, sizeof h, hipMemcpyDeviceToHost, r->ctx->stream));
        HIP_TRY(hipStreamSynchronize(r->ctx->stream));
    }
    if (closest) *closest = h[0];
    if (any) *any = h[1];
    return PT_OK;
}

// ---- instrumentation / introspection used by tests and the bench (not in the reference) ----

// Replay a render's extend kernels with node/triangle counters (identical rays, untimed).
pt_status pt_renderer_set_traversal_counting(pt_renderer* r, int enabled) {
    if (!r) return fail(PT_E_INVALID, "null renderer");
    r->count_traversal = enabled != 0;
    return PT_OK;
}
pt_status pt_last_traversal_counts(pt_renderer* r, uint64_t* nodes, uint64_t* tris) {
    if (!r) return fail(PT_E_INVALID, "null renderer");
    unsigned long long h[2] = {0, 0};
    if (r->b_trav.p) {
        HIP_TRY(hipMemcpyAsync(h, r->b_trav.p, sizeof h, hipMemcpyDeviceToHost, r->ctx->stream));
        HIP_TRY(hipStreamSynchronize(r->ctx->stream));
    }
    *nodes = h[0]; *tris = h[1];
    return PT_OK;
}
// Closest-hit trace with node/triangle counters, device-resident rays.
pt_status pt_trace_closest_device_counted(pt_ctx* ctx, const float* d_rays, uint32_t n, float* d_t, int32_t* d_prim,
                                          uint64_t* nodes, uint64_t* tris) {
    if (!ctx || !ctx->has_scene) return fail(PT_E_MISSING_INPUT, "missing input: no scene uploaded");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipMemsetAsync(ctx->trav.p, 0, ctx->trav.bytes, ctx->stream));
    unsigned long long* trav = (unsigned long long*)ctx->trav.p;
    if (ctx->stack == 16)
        hipLaunchKernelGGL((k_trace_closest<16, true>), dim3(grid_for(n, TBLOCK)), dim3(TBLOCK), 0, ctx->stream, ctx->dev, d_rays, (int)n, d_t, d_prim, trav);
    else
        hipLaunchKernelGGL((k_trace_closest<32, true>), dim3(grid_for(n, TBLOCK)), dim3(TBLOCK), 0, ctx->stream, ctx->dev, d_rays, (int)n, d_t, d_prim, trav);
    HIP_TRY(hipGetLastError());
    unsigned long long h[2];
    HIP_TRY(hipMemcpyAsync(h, trav, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *nodes = h[0]; *tris = h[1];
    return PT_OK;
}
pt_status pt_scene_stats(pt_ctx* ctx, uint32_t* wide_nodes, uint32_t* wide_depth, uint32_t* stack) {
    if (!ctx || !ctx->has_scene) return fail(PT_E_MISSING_INPUT, "missing input: no scene uploaded");
    *wide_nodes = (uint32_t)ctx->dev.n_cw_nodes;
    *wide_depth = ctx->wide_depth;
    *stack = (uint32_t)ctx->stack;
    return PT_OK;
}
// Struct sizes as compiled, for the ABI mirror check.
size_t pt_abi_sizeof(int which) {
    switch (which) {
        case 0: return sizeof(pt_object);
        case 1: return sizeof(pt_material);
        case 2: return sizeof(pt_light);
        case 3: return sizeof(pt_env);
        case 4: return sizeof(pt_scene_desc);
        case 5: return sizeof(pt_program);
        case 6: return sizeof(pt_params);
        case 7: return sizeof(pt_inst_table);
        default: return 0;
    }
}

}  // extern "C"

extern "C" {
// Host-only: build the wide BVH for tri_vertices and copy it out (nodes: 20 u32 each; tris: 12
// floats each). Returns the node count; *n_tris_out receives the leaf triangle count. Buffers
// may be null to query sizes. Used by the CPU-side structural checks of the builder.
long long pt_debug_wide_bvh(const float* tri_vertices, int n_tris, uint32_t* nodes, long long node_cap,
                            float* tris, long long tri_cap, long long* n_tris_out, uint32_t* depth) {
    WideBVH wb;
    try { build_wide_bvh(tri_vertices, n_tris, wb); } catch (...) { return -1; }
    long long nn = (long long)wb.n_nodes, nt = (long long)(wb.tris.size() / 12);
    if (nodes && node_cap >= nn) std::memcpy(nodes, wb.nodes.data(), (size_t)nn * 20 * 4);
    if (tris && tri_cap >= nt) std::memcpy(tris, wb.tris.data(), (size_t)nt * 48);
    if (n_tris_out) *n_tris_out = nt;
    if (depth) *depth = wb.max_depth;
    return nn;
}
}  // extern "C"

namespace {
__global__ void k_debug_math(int fn, const float* x, const float* y, float* out, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a = x[i], b = y[i], r;
    switch (fn) {
        case 0: r = ptm_sin(a); break;
        case 1: r = ptm_cos(a); break;
        case 2: r = ptm_exp(a); break;
        case 3: r = ptm_log(a); break;
        case 4: r = ptm_atan2(b, a); break;
        case 5: r = ptm_tan(a); break;
        case 6: r = ptm_pow(a, b); break;
        case 7: r = ptm_atan(a); break;
        case 8: r = ptm_acos(a); break;
        case 9: r = ptm_tanh(a); break;
        case 10: r = ptm_atanh(a); break;
        case 11: r = ptm_asin(a); break;
        case 20: r = a / b; break;
        case 21: r = sqrtf(a); break;
        case 22: r = 1.0f / sqrtf(a); break;
        case 23: r = sqrtf(a); break;
        case 24: r = 1.0f / a; break;
        case 25: r = a * b + a; break;
        case 26: r = gmax(a, b); break;
        default: r = 0.0f;
    }
    out[i] = r;
}
}  // namespace

extern "C" {
// Evaluate one declared-arithmetic function (or a basic op) on the GPU; host arrays.
pt_status pt_debug_math(int fn, const float* x, const float* y, float* out, int n) {
    float *dx, *dy, *dout;
    HIP_TRY(hipMalloc(&dx, n * 4)); HIP_TRY(hipMalloc(&dy, n * 4)); HIP_TRY(hipMalloc(&dout, n * 4));
    HIP_TRY(hipMemcpy(dx, x, n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dy, y, n * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_debug_math, dim3((n + 255) / 256), dim3(256), 0, 0, fn, dx, dy, dout, n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dout, n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipFree(dx)); HIP_TRY(hipFree(dy)); HIP_TRY(hipFree(dout));
    return PT_OK;
}
}  // extern "C"

namespace {
__global__ void __launch_bounds__(256) k_calib_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}
}  // namespace

extern "C" {
// Plain device copy of `bytes` (multiple of 16) repeated `iters` times: the reference kernel of
// known traffic that scales the FETCH_SIZE / WRITE_SIZE counters of the profiled kernels.
pt_status pt_calib_copy(pt_ctx* ctx, size_t bytes, int iters) {
    if (!ctx || bytes % 16 != 0 || iters < 1) return fail(PT_E_INVALID, "pt_calib_copy: bad argument");
    HIP_TRY(hipSetDevice(ctx->device));
    void *a, *b;
    HIP_TRY(hipMalloc(&a, bytes)); HIP_TRY(hipMalloc(&b, bytes));
    HIP_TRY(hipMemsetAsync(a, 0, bytes, ctx->stream));
    for (int i = 0; i < iters; i++)
        hipLaunchKernelGGL(k_calib_copy, dim3(4096), dim3(256), 0, ctx->stream, (const float4*)a, (float4*)b, bytes / 16);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipFree(a)); HIP_TRY(hipFree(b));
    return PT_OK;
}
}  // extern "C"

extern "C" {
// Host-only: the env importance table libptmi builds for `env` (cond W*H, marg H floats).
// Device bake of `env`'s importance table on `device` (the path pt_scene_upload takes): copies
// the tables and total out and writes the kernels' time (HIP events) to *ms. Returns a pt_status.
int pt_debug_env_bake(int device, const pt_env* env, float* cond, float* marg, double* total, float* ms) {
    if (!env || env->type != PT_ENV_SKY) return fail(PT_E_INVALID, "pt_debug_env_bake: not a tabulated environment");
    HIP_TRY(hipSetDevice(device));
    hipStream_t stream;
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    pt_status st;
    {
        SceneStore SS;
        SS.stream = stream;
        DevScene D{};
        st = bake_env_table(*env, SS, D, e0, e1);
        if (st == PT_OK) {
            const size_t H = env->table_h > 0 ? env->table_h : 256,
                         W = env->chart == PT_CHART_OCTAHEDRAL ? H : (env->table_w > 0 ? env->table_w : 512);
            HIP_TRY(hipMemcpyAsync(cond, D.env_cond, W * H * 4, hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipMemcpyAsync(marg, D.env_marg, H * 4, hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipMemcpyAsync(total, D.env_total, 8, hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipStreamSynchronize(stream));
            HIP_TRY(hipEventElapsedTime(ms, e0, e1));
        }
        (void)hipStreamSynchronize(stream);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(stream);
    return st;
}

int pt_debug_env_table(const pt_env* env, float* cond, float* marg, double* total) {
    std::vector<float> c, m; double t = 0;
    build_env_table(*env, c, m, t);
    std::memcpy(cond, c.data(), c.size() * 4);
    std::memcpy(marg, m.data(), m.size() * 4);
    *total = t;
    return 0;
}
}  // extern "C"
