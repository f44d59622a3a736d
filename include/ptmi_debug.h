/* ptmi_debug.h — instrumentation and introspection entry points of libptmi.
 *
 * None of these has a counterpart in the reference: they exist for the parity tests and the
 * benchmark (traversal counters for the roofline's algorithmic bytes, the copy kernel that
 * calibrates hardware byte counters, device-vs-host checks of the declared arithmetic and of
 * the environment table bake, record sizes for the ABI mirrors). A host application that
 * only renders needs nothing from this header; every symbol is exported by the same
 * libptmi.so as include/ptmi.h's.
 */
#ifndef PTMI_DEBUG_H
#define PTMI_DEBUG_H

#include <stddef.h>
#include <stdint.h>

#include "ptmi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replay the renderer's next renders with node / triangle counters in the extend kernel
 * (identical rays; the counters cost time, so timed runs keep this off). */
pt_status pt_renderer_set_traversal_counting(pt_renderer* r, int enabled);
/* Wide-BVH nodes visited and triangles tested by the last counted render. */
pt_status pt_last_traversal_counts(pt_renderer* r, uint64_t* nodes, uint64_t* tris);
/* pt_trace_closest_device with the same counters (rays resident on the device). */
pt_status pt_trace_closest_device_counted(pt_ctx* ctx, const float* d_rays, uint32_t n, float* d_t, int32_t* d_prim,
                                          uint64_t* nodes, uint64_t* tris);
/* The resident scene's wide-BVH node count, depth and the traversal stack it selected. */
pt_status pt_scene_stats(pt_ctx* ctx, uint32_t* wide_nodes, uint32_t* wide_depth, uint32_t* stack);
/* sizeof of the ABI records as compiled: 0 pt_object, 1 pt_material, 2 pt_light, 3 pt_env,
 * 4 pt_scene_desc, 5 pt_program, 6 pt_params, 7 pt_inst_table; 0 for any other value. */
size_t pt_abi_sizeof(int which);
/* The wide BVH the library builds over n_tris world triangles (9 floats each): copies the
 * 80-byte nodes (20 words each) when node_cap holds them and the leaf-ordered triangles (12
 * floats each) when tri_cap does; returns the node count, or -1 when the build fails. */
long long pt_debug_wide_bvh(const float* tri_vertices, int n_tris, uint32_t* nodes, long long node_cap,
                            float* tris, long long tri_cap, long long* n_tris_out, uint32_t* depth);
/* One function of the declared fp32 arithmetic (include/ptmi_math.h) evaluated on the GPU
 * over n host values: out[i] = f(x[i], y[i]). */
pt_status pt_debug_math(int fn, const float* x, const float* y, float* out, int n);
/* scene_region_at (features/intersection.ts:1121-1265) on the GPU at n points (xyz, host
 * memory) of the resident scene: out[i] = the innermost region holding point i, -1 outside all. */
pt_status pt_debug_region_at(pt_ctx* ctx, const float* pts, int n, int32_t* out);
/* environment_sample (components/env/sampler_cdf.glsl:61-80) of the resident scene's environment
 * on the GPU at n (x0, x1) pairs (host memory, 2 floats each): out[4i..4i+3] = the sampled
 * direction and its pdf. */
pt_status pt_debug_env_sample(pt_ctx* ctx, const float* xi, int n, float* out);
/* A plain device copy of `bytes` (a multiple of 16), repeated `iters` times: the kernel of
 * known traffic that scales the FETCH_SIZE / WRITE_SIZE counters of profiled kernels. */
pt_status pt_calib_copy(pt_ctx* ctx, size_t bytes, int iters);
/* The procedural sky's importance table as the device bake builds it on `device` (the path
 * pt_scene_upload takes), with the bake's time (HIP events) in *ms. Returns a pt_status. */
int pt_debug_env_bake(int device, const pt_env* env, float* cond, float* marg, double* total, float* ms);
/* The same table from the host builder (no device). Returns a pt_status. */
int pt_debug_env_table(const pt_env* env, float* cond, float* marg, double* total);

#ifdef __cplusplus
}
#endif

#endif /* PTMI_DEBUG_H */
