// builder.h — host-side acceleration and table builders of libptmi.
//
// Wide BVH: an 8-wide compressed BVH over world-space triangles (Ylitie, Karras, Laine,
// "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs", HPG 2017),
// the layout the reference uses for instance batches (components/accel/cwbvh/cwbvh.ts:36-50),
// extended here to triangle leaves. Node = 80 bytes = 5 x uint4:
//   w0: origin p.xyz (f32), [e.x | e.y<<8 | e.z<<16 | imask<<24]   (e = biased exponent)
//   w1: child_base, tri_base, meta[0..3], meta[4..7]
//   w2: qlo.x[0..3], qlo.x[4..7], qlo.y[0..3], qlo.y[4..7]
//   w3: qlo.z[0..3], qlo.z[4..7], qhi.x[0..3], qhi.x[4..7]
//   w4: qhi.y[0..3], qhi.y[4..7], qhi.z[0..3], qhi.z[4..7]
// Child slot i box = p + q * 2^(e-127), decoded with the same fp32 ops the kernel uses and
// checked here to contain the child's geometry (so the kernel's slab test is never less
// conservative than an exact-box test). meta[i]: 0 = empty; internal child: 0x20 | i with
// imask bit i set (child node = child_base + popcount(imask below i)); leaf: unary count in
// bits 5..7 (1/3/7 = 1/2/3 triangles), triangle offset from tri_base in bits 0..4.
// Slots are assigned by centroid octant so a ray with direction octant o visits slot s in
// order s ^ (o ^ 7) (front to back).
#pragma once
#include <cstdint>
#include <vector>
#include <string>
#include "../../include/ptmi.h"

namespace ptmi {

struct WideBVH {
    std::vector<uint32_t> nodes;      // 20 u32 per node
    std::vector<float> tris;          // 12 floats per triangle in leaf order: a.xyz,id | e1.xyz,0 | e2.xyz,0
    uint32_t n_nodes = 0;
    uint32_t max_depth = 0;
};

// Binary BVH in the reference's node layout (accel/bvh/bvh.ts:21-25): 8 floats per node
// (min.xyz, A, max.xyz, B), A >= 0 leaf (count A at offset B of item order), A < 0 internal
// (axis -A-1, right child B, left child = this + 1). Used for containment queries.
struct BinaryBVH {
    std::vector<float> nodes;
    std::vector<uint32_t> order;
};

// Build the wide BVH over tri_vertices [n][9] (world space).
void build_wide_bvh(const float* tri_vertices, int n_tris, WideBVH& out);

// Binned-SAH binary BVH over boxes [n][6] (f64 min.xyz, max.xyz), leaf size <= leaf_size.
void build_binary_bvh(const double* boxes, int n, int leaf_size, BinaryBVH& out);

// The containment probe (mesh_inside_bvh, intersection/mesh/mesh.glsl:202-296) casts every ray
// along one fixed direction, so the triangles a probe ray can hit are those whose projections
// onto the plane across that direction cover the ray's point. The probe map is a uniform grid
// over that plane per object: each cell lists the triangles whose projected footprint, widened
// by a margin far above the rounding of the projection and of the Moller-Trumbore test,
// overlaps it. The cell of a probe's point therefore holds every triangle the probe ray can
// hit, and the closest of the cell's hits under the (t, id) rule is the walk's closest hit.
constexpr float PROBE_DIR[3] = {0.5320544f, 0.7396997f, 0.4114743f};
struct ProbeMap {
    float basis[6];                 // u = basis[0..2] . p, v = basis[3..5] . p (fp32 on the device)
    float u0 = 0, v0 = 0, inv_u = 0, inv_v = 0;
    int gu = 0, gv = 0;
    std::vector<int> cell_start;    // gu * gv + 1 offsets into items
    std::vector<int> items;         // triangle indices (first + k) into the concatenated leaf triangles
    std::vector<float> boxes;       // per item its widened footprint (umin, umax, vmin, vmax), rounded outward
};
// The basis of the plane across PROBE_DIR (f64, rounded once to fp32).
void probe_basis(float basis[6]);
// The map of n triangles in the wide-BVH leaf layout (12 floats each: a.xyz,id | e1 | e2),
// which sit at index `first` of the concatenated triangle array.
void build_probe_map(const float* tris12, int n, int first, ProbeMap& m);

// Env importance table for the procedural sky: fp32 bake on the equirect grid, then the
// f64 luminance / blur / CDF build, stored f32 (EnvironmentBake.ts:62-67, env/importance.ts).
void build_env_table(const pt_env& env, std::vector<float>& cond, std::vector<float>& marg, double& total);
// The chain after the luminance pass (importance.ts:44-127), f64, stored f32.
void env_table_from_luma(const std::vector<float>& Y, int W, int H, bool oct, bool compensated, std::vector<float>& cond,
                         std::vector<float>& marg, double& total);
// resampleEquirectToOctahedral (env/octahedral/octahedral.ts:72-100): the W x H equirect RGB picture
// at the N x N equal-area texel centres, bilinear, u wrapping, v clamped, f64, stored f32.
void resample_equirect_to_octahedral(const float* rgb, int W, int H, int N, std::vector<float>& out);
// An image environment's table on env.chart (the picture, or its octahedral resample at N = table_h).
void build_env_table_image(const pt_env& env, const float* rgb, int w, int h, std::vector<float>& cond,
                           std::vector<float>& marg, double& total);

}  // namespace ptmi
