// kernels.h — the wavefront kernels (raygen, extend, bin, shade, shadow, accumulate, the
// batch's raygen / accumulate, the adjoint reductions and the ray-query entry kernels).
//
// Every kernel has internal linkage: each translation unit that launches one instantiates its
// own copy, so the pipeline variants compile in parallel units.
#pragma once
#include "dev_common.h"
#include "shading.h"
#include "traversal.h"
#include "pipeline.h"

using namespace ptmi;

namespace {

// The scene a path belongs to: the kernel argument, or its entry in the batch.
template <bool BATCH> PT_DEV const DevScene& scene_of(const DevScene& Sk, const Paths& P, int path) {
    return BATCH ? P.scenes[P.path_scene[path]] : Sk;
}

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

// fisheye.glsl FISHEYE_THETA: the radial map of each projection.
PT_DEV float fisheye_theta(int projection, float rho, float K) {
    if (projection == PT_FISHEYE_EQUISOLID) return 2.0f * ptm_asin(gclamp(rho * K, 0.0f, 1.0f));
    if (projection == PT_FISHEYE_STEREOGRAPHIC) return 2.0f * ptm_atan(rho * K);
    if (projection == PT_FISHEYE_ORTHOGRAPHIC) return ptm_asin(gclamp(rho * K, 0.0f, 1.0f));
    return rho * K;
}

// path_state_init's current_medium = scene_region_at(camera origin), once per wave.
__global__ void __launch_bounds__(64) k_camera_medium(DevScene S, Paths P, Camera C) {
    if (threadIdx.x == 0) qcount(P, C_CAM_MEDIUM) = region_at(S, mk3(C.ox, C.oy, C.oz));
}
// The same per path when camera rays do not share an origin (orthographic), after raygen.
__global__ void __launch_bounds__(SBLOCK) k_ray_medium(DevScene S, Paths P, int n) {
    for (int base = blockIdx.x * SBLOCK; base < n; base += gridDim.x * SBLOCK) {
        const int i = base + threadIdx.x;
        if (i < n) P.medium[i] = region_at(S, xyz(P.ray_o[i]));
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
    if (P.prev_qn) P.prev_qn[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);   // kernel_init's prev_query: no normal
    P.flags[i] = mk_flags(true, 0, 0);
    if (media) P.medium[i] = qcount(P, C_CAM_MEDIUM);
    P.q[0][i] = i;
    if (ADJ) {
#pragma unroll
        for (int k = 0; k < ADJ_N; k++) { P.adj_D[(size_t)k * P.cap + i] = 0.0f; P.adj_G[(size_t)k * P.cap + i] = 0.0f; }
    }
}

template <bool SPEC, bool ADJ>
__global__ void __launch_bounds__(SBLOCK) k_raygen(Paths P, Camera C, int W, int H, int n_paths, uint32_t sample0, int media, int wl_uniform) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) qcount(P, C_RAYQ0) = n_paths;
    if (i >= n_paths) return;
    int pix = i % (W * H), frame = i / (W * H);
    raygen_path<SPEC, ADJ>(P, C, W, H, i, pix % W, pix / W, sample0 + (uint32_t)frame, media, wl_uniform);
}

__global__ void __launch_bounds__(SBLOCK) k_raygen_batch(Paths P, const BatchEntry* E, int ne, const Camera* cams, int n_paths) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) qcount(P, C_RAYQ0) = n_paths;
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

// A queue of at most SMALL_QUEUE entries per lane of the grid is walked grid-stride, one ray per
// lane at a time: the persistent lanes' queue claims are same-address atomics, and on a short
// queue a grid's worth of them costs more than the imbalance they remove (the 256x256 Cornell
// box: 2.0 ms per render at 2 entries per lane, 1.26 ms at 4 and at 16).
constexpr uint32_t SMALL_QUEUE = 4;
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

constexpr int FEED_CHUNK = 128;   // queue entries a wave claims per atomic (64: 2.8 % slower; the claims serialise on one counter)
constexpr int FEED_MIN_IDLE = 8;  // idle lanes before a wave refills (4: slower; 16: the same)

// extend: closest hit for every ray of the current queue, then bin the path by shading category.
// INST: the scene has instance batches. The TLAS walk keeps a private stack, so only scenes
// with placements compile it in (the plain kernels then need no scratch memory).
template <int STACK, bool COUNT, bool BATCH = false, int TB = TBLOCK, bool INST = false>
__global__ void __launch_bounds__(TB) k_extend(DevScene Sk, Paths P, int qsel) {
    __shared__ uint2 stk[STACK * TB];
    const uint32_t n = (uint32_t)qcount(P, qsel);
    const int* q = P.q[qsel];
    TraceCounters cnt = {0, 0};
    if (n <= SMALL_QUEUE * gridDim.x * TB) {
        for (uint32_t k = blockIdx.x * TB + threadIdx.x; k < n; k += gridDim.x * TB) {
            const int path = q[k];
            const DevScene& S = scene_of<BATCH>(Sk, P, path);
            const float4 o4 = P.ray_o[path], d4 = P.ray_d[path];
            float tmax = MAX_DIST; int prim = -1;
            trace_wide<false, STACK, TB, COUNT>(S, xyz(o4), xyz(d4), tmax, prim, stk, cnt);
            trace_analytic<false>(S, xyz(o4), xyz(d4), tmax, prim);
            if (INST) trace_instances<false, TB, COUNT>(S, xyz(o4), xyz(d4), tmax, prim, stk, cnt);
            P.hit[path] = make_float2(tmax, __int_as_float(prim));
        }
        if (COUNT) {
            atomicAdd(&P.trav[0], (unsigned long long)cnt.nodes);
            atomicAdd(&P.trav[1], (unsigned long long)cnt.tris);
        }
        return;
    }
    LaneFeeder feed; feed.init();
    TravState st;
    bool active = false;
    int path = -1;
    const DevScene* Sp = &Sk;
    // the lane's scene's BVH, read once per ray (a batch's scenes live in device memory)
    const uint4* nodes = Sk.cw_nodes;
    const float4* tris = Sk.cw_tris;
    bool walk = Sk.n_cw_nodes > 0;
    for (;;) {
        int k = feed.feed<FEED_CHUNK>(&qcount(P, C_EXT_FETCH), n, !active, FEED_MIN_IDLE);
        if (k >= 0) {
            path = q[k];
            Sp = &scene_of<BATCH>(Sk, P, path);
            if (BATCH) { nodes = Sp->cw_nodes; tris = Sp->cw_tris; walk = Sp->n_cw_nodes > 0; }
            float4 o4 = P.ray_o[path], d4 = P.ray_d[path];
            trace_init(st, xyz(o4), xyz(d4), MAX_DIST);
            active = true;
        }
        if (__ballot(active) == 0) {
            if (feed.exhausted) break;
            continue;
        }
        if (active && (!walk || trace_step<false, TB, COUNT>(nodes, tris, st, stk, cnt))) {
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
template <bool SPEC, bool ADJ>
PT_DEV void shadow_add(Paths& P, int k, float4 d4) {
    const int i = P.s_path[k];
    const float4 c = P.s_c[k];
    const float w = SPEC ? P.s_l[k].w : c.w;
    const sp4 a = SPEC ? ld4(c) : mk4(c.x, c.y, c.z, 0.0f);
    const sp4 v = (a * w) / d4.w;
    store_rad<SPEC>(P, i, load_rad<SPEC>(P, i) + v);
    if (ADJ) {
        float D[ADJ_N];
#pragma unroll
        for (int kk = 0; kk < ADJ_N; kk++) D[kk] = P.adj_sD[(size_t)kk * P.cap + k];
        adj_add_G(P, i, D, adj_e(P, i, v));
    }
}
// Persistent lanes, as k_extend (a short queue is walked grid-stride): a lane whose ray is done
// takes the next one while its wave-mates keep walking (any-hit walks end at very different
// depths). The analytic objects come first (a
// hit there needs no walk), then the wide BVH step by step, then the instance batches.
template <int STACK, bool SPEC, bool ADJ, bool BATCH = false, int TB = TBLOCK, bool INST = false>
__global__ void __launch_bounds__(TB) k_shadow(DevScene Sk, Paths P) {
    __shared__ uint2 stk[STACK * TB];
    const uint32_t n = (uint32_t)qcount(P, C_SHADOW);
    TraceCounters cnt = {0, 0};
    if (n <= SMALL_QUEUE * gridDim.x * TB) {
        for (uint32_t k = blockIdx.x * TB + threadIdx.x; k < n; k += gridDim.x * TB) {
            const DevScene& S = scene_of<BATCH>(Sk, P, P.s_path[k]);
            const float4 o4 = P.s_o[k], d4 = P.s_d[k];
            float maxd = o4.w; int prim = -1;
            bool blocked = trace_analytic<true>(S, xyz(o4), xyz(d4), maxd, prim);
            if (!blocked) blocked = trace_wide<true, STACK, TB, false>(S, xyz(o4), xyz(d4), maxd, prim, stk, cnt);
            if (INST && !blocked) blocked = trace_instances<true, TB, false>(S, xyz(o4), xyz(d4), maxd, prim, stk, cnt);
            if (!blocked) shadow_add<SPEC, ADJ>(P, (int)k, d4);
        }
        return;
    }
    LaneFeeder feed; feed.init();
    TravState st;
    bool active = false;
    int k = -1;
    float4 d4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const DevScene* Sp = &Sk;
    const uint4* nodes = Sk.cw_nodes;   // the lane's scene's BVH, read once per ray
    const float4* tris = Sk.cw_tris;
    for (;;) {
        const int kk = feed.feed<FEED_CHUNK>(&qcount(P, C_SHD_FETCH), n, !active, FEED_MIN_IDLE);
        if (kk >= 0) {
            k = kk;
            Sp = &scene_of<BATCH>(Sk, P, P.s_path[k]);
            if (BATCH) { nodes = Sp->cw_nodes; tris = Sp->cw_tris; }
            const float4 o4 = P.s_o[k];
            d4 = P.s_d[k];
            float maxd = o4.w; int prim = -1;
            if (!trace_analytic<true>(*Sp, xyz(o4), xyz(d4), maxd, prim)) {
                if (Sp->n_cw_nodes > 0) {
                    trace_init(st, xyz(o4), xyz(d4), o4.w);
                    active = true;
                } else if (!(INST && trace_instances<true, TB, false>(*Sp, xyz(o4), xyz(d4), maxd, prim, stk, cnt))) {
                    shadow_add<SPEC, ADJ>(P, k, d4);
                }
            }
        }
        if (__ballot(active) == 0) {
            if (feed.exhausted) break;
            continue;
        }
        if (active && trace_step<true, TB, false>(nodes, tris, st, stk, cnt)) {
            active = false;
            bool blocked = st.prim >= 0;
            float maxd = st.tmax; int prim = -1;
            if (INST && !blocked) blocked = trace_instances<true, TB, false>(*Sp, st.o, st.d, maxd, prim, stk, cnt);
            if (!blocked) shadow_add<SPEC, ADJ>(P, k, d4);
        }
    }
}

// shadow (media scenes): the segment walker of transport/shadow/media/media.glsl:29-62 —
// closest hits up to the light point, Beer-Lambert per segment in the current medium, pass
// through null interfaces, anything else blocks (dielectrics are opaque to shadow rays).
template <int STACK, bool SPEC, bool ADJ, bool INST = false>
__global__ void __launch_bounds__(TBLOCK) k_shadow_media(DevScene S, Paths P) {
    __shared__ uint2 stk[STACK * TBLOCK];
    const int n = qcount(P, C_SHADOW);
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
        int medium = region_at(S, o);
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
            classify_hit(S, h, back, true);
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
// The same for IT entries per thread (slot[i] for which[i], -1 where which[i] < 0).
template <int K, int IT>
PT_DEV void block_append_n(int* const* counters, const int* which, int* slot, int* s_cnt, int* s_base) {
    if (threadIdx.x < K) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    int local[IT];
#pragma unroll
    for (int i = 0; i < IT; i++) local[i] = which[i] >= 0 ? atomicAdd(&s_cnt[which[i]], 1) : -1;
    __syncthreads();
    if (threadIdx.x < K) s_base[threadIdx.x] = s_cnt[threadIdx.x] ? atomicAdd(counters[threadIdx.x], s_cnt[threadIdx.x]) : 0;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < IT; i++) slot[i] = which[i] >= 0 ? s_base[which[i]] + local[i] : -1;
}

// The shading category of a vertex on primitive prim (prim < 0: a miss).
PT_DEV int surface_category(const DevScene& S, int prim) {
    if (prim < 0) return CAT_MISS;
    const int model = S.materials[S.objects[prim_owner(S, prim)].material].model;
    return model == PT_MAT_LAMBERT || model == PT_MAT_CHECKER ? CAT_LAMBERT
         : model == PT_MAT_GGX ? CAT_GGX
         : model == PT_MAT_DIELECTRIC ? CAT_DIELECTRIC
         : model == PT_MAT_MIRROR ? CAT_MIRROR
         : model == PT_MAT_ROUGH_DIELECTRIC ? CAT_ROUGH : CAT_NULL;
}
// Paths per thread per pass of the opaque-medium bin: its work per path is a chain of dependent
// loads (queue, hit, owner, object, material), so several chains per thread keep more in flight
// (and the block's queue atomics are shared by more paths): 11.9 ms per 64-spp step of the
// 1M-triangle cloud at 1 path, 3.9 at 4, 3.3 at 8.
constexpr int BIN_ITEMS = 8;

// bin: per path of the iteration, the medium segment first (media scenes: the analytic
// channel-MIS distance sample of pt.ts's media branch, which may end the segment in a
// scattering event), then the shading category of the vertex.
template <bool MEDIA, bool SPEC, bool ADJ, bool BATCH = false>
__global__ void __launch_bounds__(SBLOCK) k_bin(DevScene Sk, Paths P, int qsel, int equiangular, int max_bounces, int max_null) {
    __shared__ int s_cnt[N_CAT], s_base[N_CAT];
    int* ctr[N_CAT];
    for (int c = 0; c < N_CAT; c++) ctr[c] = &qcount(P, C_CAT0 + c);
    int* sctr[1] = {&qcount(P, C_SHADOW)};
    const int n = qcount(P, qsel);
    const int* q = P.q[qsel];
    if (!MEDIA) {
        for (int base = blockIdx.x * SBLOCK * BIN_ITEMS; base < n; base += gridDim.x * SBLOCK * BIN_ITEMS) {
            int cat[BIN_ITEMS], path[BIN_ITEMS], slot[BIN_ITEMS];
#pragma unroll
            for (int i = 0; i < BIN_ITEMS; i++) {
                const int k = base + i * SBLOCK + threadIdx.x;
                cat[i] = -1; path[i] = 0;
                if (k < n) {
                    path[i] = q[k];
                    cat[i] = surface_category(scene_of<BATCH>(Sk, P, path[i]), __float_as_int(P.hit[path[i]].y));
                }
            }
            block_append_n<N_CAT, BIN_ITEMS>(ctr, cat, slot, s_cnt, s_base);
#pragma unroll
            for (int i = 0; i < BIN_ITEMS; i++) if (cat[i] >= 0) P.catq[cat[i]][slot[i]] = path[i];
        }
        return;
    }
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
                        const float f = phase_value(Mm.phase, eq_wi, -d, Mm.hg_g);
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
            if (cat < 0) cat = surface_category(S, prim);
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
        P.hist[1] += (unsigned long long)qcount(P, C_SHADOW);
        qcount(P, C_SHADOW) = 0;
    }
}

// Queue bookkeeping before an iteration (single thread).
__global__ void k_bookkeep(Paths P, int cur, int next, int first) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (!first) P.hist[1] += (unsigned long long)qcount(P, C_SHADOW);
        if (cur >= 0) P.hist[0] += (unsigned long long)qcount(P, cur);
        if (next >= 0) qcount(P, next) = 0;
        qcount(P, C_SHADOW) = 0;
        for (int c = 0; c < N_CAT; c++) qcount(P, C_CAT0 + c) = 0;
        qcount(P, C_EXT_FETCH) = 0;
        qcount(P, C_SHD_FETCH) = 0;
    }
}

// probe: the regions on both sides of every transmissive hit of the iteration (the dielectric
// and rough-dielectric queues; scene_region_at, features/intersection.ts:1121-1265), ahead of
// their shading. A probe is a chain of dependent loads (the hit, its object, the containment
// grid cell, the probe-map cell and its candidate triangles) with little arithmetic, and an
// iteration's queue is too short to fill the GPU with one probe per lane; so each probe runs on
// a group of PROBE_GROUP lanes that share its candidates, and a small kernel of its own keeps
// its waves resident where the shading kernel's registers would not.
constexpr int PROBE_GROUP = 4;
template <bool BATCH = false>
__global__ void __launch_bounds__(SBLOCK) k_probe(DevScene Sk, Paths P) {
    constexpr int G = PROBE_GROUP;
    const int n0 = qcount(P, C_CAT0 + CAT_DIELECTRIC);
    const int n = n0 + qcount(P, C_CAT0 + CAT_ROUGH);
    const int g = threadIdx.x % G;
    // the stride is a multiple of G, so a group's lanes share k on every pass (and leave together)
    for (int w = blockIdx.x * SBLOCK + threadIdx.x; w / G < n; w += gridDim.x * SBLOCK) {
        const int k = w / G;
        const int i = k < n0 ? P.catq[CAT_DIELECTRIC][k] : P.catq[CAT_ROUGH][k - n0];
        const DevScene& S = scene_of<BATCH>(Sk, P, i);
        const float2 hv = P.hit[i];
        const f3 d = xyz(P.ray_d[i]);
        Hit h;
        hit_geometry(S, xyz(P.ray_o[i]), d, hv.x, __float_as_int(hv.y), h.p, h.ng, h.owner);
        classify_regions<G>(S, h, !(dot(d, h.ng) < 0.0f), true, g);
        if (g == 0) P.region[i] = make_int2(h.region_from, h.region_to);
    }
}

// The per-vertex body of transport_trace (integrators/pt/pt.ts:299-464) with the combiner
// (transport/combiner.ts:23-102) and the kernel/light techniques (techniques/kernel/kernel.glsl,
// kernel_phase.glsl, techniques/light/light.glsl, light_medium.glsl), for the paths of one
// category. MEDIA selects the media program (region tracking, medium NEE shadow walker).
template <int EST, int CAT, bool MEDIA, bool SPEC, bool ADJ, bool BATCH = false>
__global__ void __launch_bounds__(SBLOCK) k_shade(DevScene Sk, Paths P, int qsel, int max_bounces,
                                                  int rr_start, float rr_max, int max_null, int equiangular = 0) {
    constexpr bool PROBE = cat_probes(CAT, MEDIA);
    __shared__ int s_cnt[2], s_base[2];
    int* ctr[2] = {&qcount(P, qsel ^ 1), &qcount(P, C_SHADOW)};
    const int n = qcount(P, C_CAT0 + CAT);
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
            const bool tree = P.prev_qn != nullptr;     // light-tree selection
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
                    const int phase = Mm.phase;
                    // the phase's own parameter: g of a Henyey-Greenstein medium (Rayleigh has none)
                    const bool param_medium = ADJ && mmat == P.adj_medium;
                    const bool param_g = param_medium && phase == PT_PHASE_HG;
                    f3 ro = xyz(P.ray_o[i]);
                    const float t = P.hit[i].x;
                    f3 p_evt = ro + rd * t;
                    f3 wo_med = -rd;
                    uint2 rs = P.rng[i];
                    Rng rng; rng.base = rs.x; rng.dim = rs.y;
                    if (NEE && !equiangular) {   // light_sample_direct_medium
                        float y0 = rng.next(), y1 = rng.next();
                        const LQuery qm{p_evt, mk3(0.0f, 0.0f, 0.0f), false};   // light_query_medium
                        LSample ls = lighting_sample<SPEC>(S, qm, y0, y1, wl, rng, tree);
                        if (ls.pdf > 0.0f) {
                            s_l = p_evt + ls.wi * ls.distance;
                            s_o = p_evt; s_d = ls.wi; s_cross = crossings_left;
                            float f = phase_value(phase, ls.wi, wo_med, g);
                            s_w = 1.0f;
                            if (MIS && !(ls.flags & 1)) s_w = power_heuristic(ls.pdf, phase_pdf<SPEC>(phase, ls.wi, wo_med, g));
                            s_a = (T * ls.radiance) * sgrey(f);
                            s_pdf = ls.pdf;
                            shadow = true;
                            if (ADJ) {
#pragma unroll
                                for (int kk = 0; kk < ADJ_N; kk++) sD[kk] = A.D[kk];
                                if (param_g) sD[6] += dlog_hg(g, dot(ls.wi, -wo_med));
                            }
                        }
                    }
                    // kernel_sample_phase
                    float z0 = rng.next(), z1 = rng.next();
                    float ppdf;
                    f3 wi = phase_sample<SPEC>(phase, wo_med, g, z0, z1, ppdf);
                    if (param_g) { A.D[6] += dlog_hg(g, dot(wi, -wo_med)); A.dirty = true; }
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
                        if (tree) P.prev_qn[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
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
                if (cat_preprobed(CAT, MEDIA)) {
                    const int2 rg = P.region[i];
                    h.region_from = rg.x;
                    h.region_to = rg.y;
                    orient_hit(h, back);
                } else {
                    classify_hit(S, h, back, PROBE || (S.emissive_solid && back));
                }
                // first-hit guides (pt.ts:422-433): the first non-null surface with no event before
                if (!SPEC && CAT != CAT_NULL && bounce == 0 && (P.aov[0] || P.aov[1])) {
                    const sp4 Ta = T;                            // the path's weight on arrival
                    if (P.aov[0]) {                                  // aov/albedo/albedo.glsl
                        float cell[3];
                        const float* row = MODEL == PT_MAT_LAMBERT ? M.albedo
                                         : (MODEL == PT_MAT_GGX || MODEL == PT_MAT_MIRROR) ? M.f0 : M.transmittance;
                        if (MODEL == PT_MAT_LAMBERT && M.model == PT_MAT_CHECKER) {   // checker_albedo at the hit
                            const sp4 c = lambert_albedo<false>(S, M, h, wl);
                            cell[0] = c.x; cell[1] = c.y; cell[2] = c.z;
                            row = cell;
                        }
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
                                else if (MIS) {
                                    LQuery pq{prev_qp, mk3(0.0f, 0.0f, 0.0f), false};
                                    if (tree) { const float4 qn = P.prev_qn[i]; pq.n = xyz(qn); pq.two_sided = qn.w != 0.0f; }
                                    w = power_heuristic(prev_pdf, lighting_pdf(S, pq, rd, lid, h.p, tree));
                                }
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
                            const LQuery qs{h.p, h.frame.n, MODEL == PT_MAT_ROUGH_DIELECTRIC};   // light_query_surface
                            LSample ls = lighting_sample<SPEC>(S, qs, x0, x1, wl, rng, tree);
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
                                if (tree) P.prev_qn[i] = f4(h.frame.n, MODEL == PT_MAT_ROUGH_DIELECTRIC ? 1.0f : 0.0f);
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
                                                       float4* aov1, int oneshot) {
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
            if (n == 0 || oneshot) g = make_float4(x.x, x.y, x.z, 1.0f);
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
        if (n == 0 || oneshot) acc = make_float4(x.x, x.y, x.z, 1.0f);   // oneshot_update: the latest sample
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
