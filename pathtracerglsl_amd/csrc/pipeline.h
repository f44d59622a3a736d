// pipeline.h — the wavefront pipeline's shared records: launch shapes, shading categories,
// queue counter slots, the SoA path state, the camera and the batch's wave entries. Host and
// device code of every translation unit see the same layouts.
#pragma once
#include "dev_common.h"

namespace ptmi {

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
// Categories whose regions k_probe classifies ahead of their shading (opaque-medium programs: the
// transmissive models); their shade kernels read them from the path state.
constexpr bool cat_preprobed(int cat, bool media) { return !media && (cat == CAT_DIELECTRIC || cat == CAT_ROUGH); }
// counts[] slots
enum { C_RAYQ0 = 0, C_RAYQ1 = 1, C_SHADOW = 2, C_CAT0 = 3, C_EXT_FETCH = C_CAT0 + N_CAT, C_SHD_FETCH, C_CAM_MEDIUM,
       N_COUNTS = 16 };

// Queue counters sit CNT_STRIDE ints apart, each in its own 128-byte line, so the atomics of
// different queues do not queue behind each other.
constexpr int CNT_STRIDE = 32;
// Bytes ahead of a path state's arrays: the queue counters, ray and traversal statistics.
constexpr size_t STATE_HEAD_BYTES = ((N_COUNTS * CNT_STRIDE * 4 + 255) & ~(size_t)255) + 512;

struct Paths {
    float4* ray_o;   // origin.xyz
    float4* ray_d;   // direction.xyz
    float4* thr;     // throughput.xyz, eta_scale (RGB) | throughput, 4 lanes (spectral)
    float4* rad;     // radiance.xyz (RGB) | 4 lanes (spectral)
    uint2* rng;      // rng_base, rng_dim
    float4* prev;    // prev_query.p.xyz, prev_bsdf_pdf
    float4* prev_qn; // prev_query.n.xyz, two_sided (1 / 0): light-tree programs only, null otherwise
    int* flags;      // bit 0 prev_was_delta, bits 8..15 null crossings, bits 16..31 bounce
    int* medium;     // current_medium region (media scenes)
    float2* hit;     // t (medium event t after binning), primitive id (bits)
    int2* region;    // (region_from, region_to) of a transmissive hit, from k_probe
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
constexpr int ADJ_N = PT_ADJ_N_PARAMS;
__host__ __device__ inline int& qcount(const Paths& P, int slot) { return P.counts[slot * CNT_STRIDE]; }

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


// Batched wavefront: a wave is a list of (scene, first frame, frames) chunks laid out back to
// back in the path slots; every scene keeps its own camera, film and sample order.
struct BatchEntry {
    int scene, W, H, npix;
    int f0, nf;
    long long base;       // first path slot of the chunk
    long long film;       // first pixel of the scene's film
};

}  // namespace ptmi
