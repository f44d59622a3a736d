// rail.h — the reference's host-side scene-data builders, restated in C++ so the App's packing
// (app/sceneData.ts) can run natively:
//   * the binned-SAH binary BVH core (components/accel/bvh/bvh.ts buildBVHCore /
//     buildBVHNodesFlat / buildBVH), byte for byte: f64 boxes, f32 centroid and area stores,
//     12 bins per axis, strict-< axis-then-bin tie order, the in-place swap partition and the
//     median (or sorted-median) fallback, nodes as 8 f32 (min.xyz, A, max.xyz, B);
//   * the data ledger (components/data/ledger.ts planDataLayout): every tenant's region base in
//     every channel from counts alone;
//   * the mesh part of the scene-data packing: vertex / normal / uv / index / BLAS-node texels
//     written at the ledger's bases (app/sceneData.ts:36-47, components/data/pack.ts).
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/ptmi.h"

namespace ptmi {

struct FlatBVH {
    std::vector<float> nodes;       // 8 floats per node
    std::vector<uint32_t> order;    // item permutation; leaves are contiguous slices of it
    int node_count = 0;
    int max_depth = 0;
};

// buildBVHNodesFlat(boxes, n, leafSize, alwaysSplit) with the 'median' fallback, or
// buildBVHCore's 'sorted-median' fallback when sorted_median is set. boxes: n x 6 f64.
void build_bvh_flat(const double* boxes, int n, int leaf_size, bool always_split, bool sorted_median, FlatBVH& out);
// buildLightTree (accel/light_tree/light_tree.ts:46-77): 2n - 1 nodes of 8 floats, (min, phi)
// (max, link: right child, or -(light + 1) at a leaf), left child implicit at node + 1; trails
// 4 floats per light (the 24-bit low and high halves of its root-to-leaf bits, LSB first).
constexpr int LIGHT_TREE_MAX_DEPTH = 48;
void build_light_tree(const double* boxes, const double* powers, int n, std::vector<float>& nodes,
                      std::vector<float>& trails, int& node_count, int& max_depth);

// buildBVH(positions, indices): per-triangle f64 boxes, the flat build at leaf size 2, and the
// triangle indices re-emitted in leaf order (3 per triangle).
void build_mesh_bvh(const float* positions, const uint32_t* indices, int n_tris, FlatBVH& out,
                    std::vector<uint32_t>& reindexed);

}  // namespace ptmi
