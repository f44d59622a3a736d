// dev_common.h — device-side data views, fp32 vector helpers and the pcg4d sampler.
//
// Vector helpers keep GLSL's evaluation order (dot = (x*x' + y*y') + z*z', left-to-right
// sums, normalize(v) = v * (1/sqrt(dot(v,v)))) and the whole library is compiled with
// -ffp-contract=off, so every expression rounds exactly like the CPU restatement.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/ptmi.h"
#include "../../include/ptmi_math.h"

#define PT_DEV __device__ __forceinline__

namespace ptmi {

constexpr float MAX_DIST = 1000.0f;          // glsl/core/math.glsl:11
constexpr float EPS_INTERFACE = 0.001f;      // math.glsl:12
constexpr float SHADOW_BACKOFF = 0.002f;     // math.glsl:39
constexpr float BVH_TFAR_PAD = 1.00000024f;  // accel/bvh/bvh.ts:48
constexpr float INV_PI = 1.0f / PTM_PI;

struct f3 { float x, y, z; };
PT_DEV f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
PT_DEV f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
PT_DEV f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
PT_DEV f3 operator*(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
PT_DEV f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
PT_DEV f3 operator/(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
PT_DEV f3 operator-(f3 a) { return mk3(-a.x, -a.y, -a.z); }
PT_DEV float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PT_DEV f3 cross(f3 a, f3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
PT_DEV float length(f3 a) { return sqrtf(dot(a, a)); }
PT_DEV f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
// concentric_disk (glsl/core/math.glsl:71-78): the Shirley-Chiu concentric square-to-disk map.
PT_DEV float2 concentric_disk(float u0, float u1) {
    const float ox = 2.0f * u0 - 1.0f, oy = 2.0f * u1 - 1.0f;
    if (ox == 0.0f && oy == 0.0f) return make_float2(0.0f, 0.0f);
    float r, th;
    if (fabsf(ox) > fabsf(oy)) { r = ox; th = (PTM_PI / 4.0f) * (oy / ox); }
    else { r = oy; th = PTM_PI / 2.0f - (PTM_PI / 4.0f) * (ox / oy); }
    return make_float2(r * ptm_cos(th), r * ptm_sin(th));
}
PT_DEV float gmax(float a, float b) { return a > b ? a : b; }   // GLSL max on non-NaN input
PT_DEV float gmin(float a, float b) { return a < b ? a : b; }
PT_DEV float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
PT_DEV float gsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
PT_DEV f3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }
PT_DEV f3 xyz(float4 v) { return mk3(v.x, v.y, v.z); }
PT_DEV float4 f4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// Spectrum: four lanes. RGB programs use lanes 0-2 (lane 3 is never stored or reduced, so
// it costs nothing after dead-code elimination); spectral programs use all four hero
// wavelengths. Reductions take the lane count from SPEC (components/color/spectrum.glsl:5-24).
struct sp4 { float x, y, z, w; };
PT_DEV sp4 mk4(float x, float y, float z, float w) { sp4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
PT_DEV sp4 sgrey(float v) { return mk4(v, v, v, v); }
PT_DEV sp4 operator+(sp4 a, sp4 b) { return mk4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
PT_DEV sp4 operator*(sp4 a, sp4 b) { return mk4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
PT_DEV sp4 operator*(sp4 a, float s) { return mk4(a.x * s, a.y * s, a.z * s, a.w * s); }
PT_DEV sp4 operator/(sp4 a, float s) { return mk4(a.x / s, a.y / s, a.z / s, a.w / s); }
PT_DEV sp4 operator-(sp4 a) { return mk4(-a.x, -a.y, -a.z, -a.w); }
PT_DEV sp4 ld4(float4 v) { return mk4(v.x, v.y, v.z, v.w); }
PT_DEV float4 st4(sp4 v) { return make_float4(v.x, v.y, v.z, v.w); }
PT_DEV sp4 rgb4(const float* p) { return mk4(p[0], p[1], p[2], 0.0f); }
PT_DEV sp4 rgb4(f3 v) { return mk4(v.x, v.y, v.z, 0.0f); }
template <bool SPEC> PT_DEV bool is_black(sp4 s) { return s.x <= 0.0f && s.y <= 0.0f && s.z <= 0.0f && (!SPEC || s.w <= 0.0f); }
template <bool SPEC> PT_DEV float spectrum_max(sp4 s) {
    float m = SPEC ? s.w : s.z;
    if (SPEC) m = gmax(s.z, m);
    m = gmax(s.y, m); m = gmax(s.x, m);
    return m;
}
template <bool SPEC> PT_DEV float spectrum_average(sp4 s) {
    float sum = s.x; sum += s.y; sum += s.z;
    if (SPEC) { sum += s.w; return sum * 0.25f; }
    return sum * (1.0f / 3.0f);
}
PT_DEV float spectrum_comp(sp4 s, int c) { return c == 0 ? s.x : (c == 1 ? s.y : (c == 2 ? s.z : s.w)); }

// pcg4d (components/sampler/pcg4d/pcg4d.glsl:12-40); only .x is consumed.
PT_DEV uint32_t pcg4d_x(uint32_t x, uint32_t y, uint32_t z, uint32_t w) {
    x = x * 1664525u + 1013904223u; y = y * 1664525u + 1013904223u;
    z = z * 1664525u + 1013904223u; w = w * 1664525u + 1013904223u;
    x += y * w; y += z * x; z += x * y; w += y * z;
    x ^= x >> 16u; y ^= y >> 16u; z ^= z >> 16u; w ^= w >> 16u;
    x += y * w;
    return x;
}
struct Rng {
    uint32_t base, dim;
    PT_DEV float next() {
        uint32_t u = pcg4d_x(base, dim++, 0u, 0u);
        return (float)(u >> 8) * (1.0f / 16777216.0f);
    }
};

// Scene as the kernels see it (all device pointers).
struct DevScene {
    const uint4* cw_nodes;      // 5 uint4 per wide node
    const float4* cw_tris;      // 3 float4 per triangle, leaf order
    const float4* tri_shade;    // 3 float4 per triangle, original order: a | b | c
    const int* tri_object;
    int n_tris;
    int n_cw_nodes;
    int cw_depth;               // levels of the wide BVH (bounds the node groups a walk stacks)
    const pt_object* objects;
    int n_objects;
    const int* analytic;        // object ids of quads / spheres
    int n_analytic;
    const pt_material* materials;
    const pt_light* lights;
    int n_lights;
    int env_type, env_w, env_h, env_compensated, env_chart;
    int env_body;               // PT_ENV_SKY: PT_SKY_* formula body
    const float* env_img;       // PT_ENV_IMAGE: the equirect picture, RGB f32, env_img_w x env_img_h
    int env_img_w, env_img_h;
    int env_samplable;          // PT_ENV_CONSTANT sampled as a light
    float4 env_color;           // PT_ENV_CONSTANT radiance (RGB)
    pt_spectrum env_s_color;    // PT_ENV_CONSTANT radiance (spectral)
    float env_intensity, env_rotation, env_rot_c, env_rot_s, env_select_prob;
    const float* env_cond;
    const float* env_marg;
    const double* env_total;   // the table's mass (totalWeight), on the device
    // cut-point guides of the table's CDFs (Chen and Asau 1974): for bucket b of K equal parts of
    // [0, 1), the index the sampler's search returns at u = b / K; the marginal's K_m + 1 entries,
    // then K_c + 1 per row. K is a power of two, so floor(u K) is exact.
    const int* env_guide_marg;
    const int* env_guide_cond;
    int env_km, env_kc;
    // containment (closed meshes)
    const float4* obj_nodes;    // binary nodes, 2 float4 each: (min, A) (max, B)
    const int* obj_node_base;   // per object
    const int* obj_order;       // global triangle ids in per-object leaf order, at first_tri
    const float* tri_nrm;       // [shade triangle][9] per-corner normals of smooth objects (null: none)
    const uint4* ow_nodes;      // per closed mesh wide BVH (same format as cw_nodes), concatenated
    const float4* ow_tris;      // their leaf triangles (ids are global triangle ids)
    const int* ow_root;         // per object: root node in ow_nodes, -1 if none
    // containment probe maps (per closed mesh / instance batch): the plane basis across the probe
    // direction, per object (first cell, cells across u, cells across v, 0) and (u0, v0, 1/du, 1/dv),
    // cell offsets into the items (gu * gv + 1 per object), the items (indices into ow_tris) and
    // their footprints
    float pm_basis[6];
    const int4* pm_head;
    const float4* pm_grid;
    const int* pm_cells;
    const int* pm_items;
    const float4* pm_boxes;     // per item its widened footprint (umin, umax, vmin, vmax)
    // uniform grid over closed-mesh boxes: candidates of scene_region_at's containment test
    float grid_lo[3], grid_inv[3];
    int grid_n[3];
    const int2* grid_cells;     // (first item, count) per cell
    const int* grid_items;      // object ids
    int has_grid;
    int has_transmission;
    int emissive_solid;
    int ambient_material;       // material of region -1 (ambientMedium), -1 = none
    int weighted_absorption;    // some medium is on the analytic scattering arm (interior roulette runs)
    int has_media, has_null;
    // instance batches: binary TLAS over placement world boxes, placements (world -> local
    // quaternion, (t_rigid, s)), owning batch, first primitive-id offset; prototype triangles
    // sit in tri_shade after the world triangles and have per-batch wide BVHs in ow_*
    const float4* inst_nodes;
    int n_inst_nodes;
    const int* inst_order;
    const float4* inst_q;
    const float4* inst_ts;
    const int* inst_batch;
    const int* inst_base;
    int n_instances;
    int inst_prim0;             // n_tris + n_objects: the first instance primitive id
    // spectral color model: CIE X, Y, Z, D65 at 360 + 5 i nm (95 rows) and the cie1931 sensor
    const float4* spec_curves;
    float xyz_rgb[9];
    float cie_y;
    // the Jakob-Hanika coefficient atlas ([3 * 64^3] texels) and its 64 z nodes: resident only for
    // a spectral program whose procedural sky body has colour (null otherwise)
    const float4* spec_atlas;
    const float* spec_znodes;
    // the light tree over the lights (2 float4 per node: (min, phi) (max, link)) and each light's
    // trail (low, high 24-bit halves); null without the tree inputs
    const float4* lt_nodes;
    const uint2* lt_trails;
};

// Wave-aggregated queue append: every active lane calls it (convergent), lanes with
// pred == false get -1. One atomic per wave.
PT_DEV int wave_append(int* counter, bool pred) {
    uint64_t mask = __ballot(pred);
    if (mask == 0) return -1;
    int lane = (int)__lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, (int)__popcll(mask));
    base = __shfl(base, leader);
    int prefix = (int)__popcll(mask & ((1ull << lane) - 1ull));
    return pred ? base + prefix : -1;
}

}  // namespace ptmi
