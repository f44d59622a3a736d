// ptmi.hip — libptmi: the C ABI of the wavefront path-tracing pipeline for gfx950 (the
// kernels are in kernels.h, the per-variant launch sequences in wave.h / wave_*.hip).
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
#include <mutex>
#include <string>
#include <vector>

#include "builder.h"
#include "description.h"
#include "formats.h"
#include "kernels.h"
#include "host.h"

namespace ptmi {
thread_local std::string g_err;
pt_status fail(pt_status s, const std::string& msg) { g_err = msg; return s; }
}  // namespace ptmi

namespace {

// ---------------------------------------------------------------- environment bake
// The procedural sky's importance table on the device. The reference renders the environment
// into a W x H equirect target (compiler/EnvironmentBake.ts:29-80) and builds the CDFs on the CPU
// (components/env/importance.ts:16-127). Here every step runs on the GPU, in the order and
// precision of the host builder (builder.cpp build_env_table), so the tables are bit-identical:
// texels and blur taps are independent (one thread each), and the sums the reference
// accumulates sequentially in f64 stay sequential (one thread per row for the conditional CDFs,
// one thread for the marginal CDF and for the compensation mean).
// Per-texel kernels are grid-stride (the launch grid is capped at MAX_GRID blocks).
__global__ void __launch_bounds__(256) k_env_luma(float* Y, int W, int H, int body, int oct) {
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
        float c[3];
        pt_sky_body(body, dx, dy, dz, c);
        Y[k] = (float)(0.2126 * (double)c[0] + 0.7152 * (double)c[1] + 0.0722 * (double)c[2]);
    }
}

// Luminance of a caller's RGB picture (importance.ts:38-45): the f64 weighted sum of the f32
// channels, stored as f32.
__global__ void __launch_bounds__(256) k_env_luma_rgb(const float* __restrict__ rgb, float* __restrict__ Y, long long n) {
    for (long long k = blockIdx.x * 256LL + threadIdx.x; k < n; k += (long long)gridDim.x * 256) {
        const float r = rgb[3 * k], g = rgb[3 * k + 1], b = rgb[3 * k + 2];
        Y[k] = (float)(0.2126 * (double)r + 0.7152 * (double)g + 0.0722 * (double)b);
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
}  // namespace

extern "C" {

const char* pt_last_error(void) { return g_err.c_str(); }

double pt_env_image_total_weight(const pt_env* env, const float* rgb, int32_t w, int32_t h) {
    if (!env || !rgb || w < 1 || h < 1 || w > PT_ENV_MAX_W) { fail(PT_E_INVALID, "pt_env_image_total_weight: bad argument"); return 0.0; }
    std::vector<float> cond, marg;
    double total = 0.0;
    build_env_table_image(*env, rgb, w, h, cond, marg, total);
    return total;
}

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

}  // extern "C"

// ---------------------------------------------------------------- allocation-free path state
// The path state of pt_render_frames_noalloc: one array per GPU, the library's own, held until
// the process ends (PT_NOALLOC_MAX_PATHS slots with both guide pictures, plus the counter
// head). Host code never passes the array's name to a kernel or a copy: its device address on
// each GPU comes from hipGetSymbolAddress, taken once per device with that device current (at
// the first pt_ctx_create on it) and kept in noalloc_addr.
constexpr size_t NOALLOC_BYTES_PER_PATH = 280;   // the carve of one slot: 212 B of state, 8 queues, 2 guides
constexpr size_t NOALLOC_BYTES = STATE_HEAD_BYTES + NOALLOC_BYTES_PER_PATH * (size_t)PT_NOALLOC_MAX_PATHS;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
constexpr int NOALLOC_MAX_DEVICES = 64;
static void* noalloc_addr[NOALLOC_MAX_DEVICES];
static std::mutex noalloc_mutex;

static pt_status noalloc_setup(int device, void** out) {
    if (device >= NOALLOC_MAX_DEVICES) return fail(PT_E_UNSUPPORTED, "allocation-free path state: device index too large");
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    if (!noalloc_addr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PT_E_DEVICE, std::string("allocation-free path state: hipGetSymbolAddress: ") + hipGetErrorString(e));
        noalloc_addr[device] = p;
    }
    *out = noalloc_addr[device];
    return PT_OK;
}

extern "C" {

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
        c->cu_count = prop.multiProcessorCount;
        if (prop.maxSharedMemoryPerMultiProcessor > 0) c->lds_per_cu = (int)prop.maxSharedMemoryPerMultiProcessor;
        c->shade_blocks = prop.multiProcessorCount * 8;
    }
    c->trav.bytes = 2 * sizeof(unsigned long long);
    if (hipMalloc(&c->trav.p, c->trav.bytes) != hipSuccess) { delete c; return fail(PT_E_OOM, "counter allocation failed"); }
    if (pt_status st = noalloc_setup(device_index, &c->noalloc_state); st != PT_OK) { delete c; return st; }
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
    if (sd->env.type == PT_ENV_IMAGE) {
        if (!sd->env_image || sd->env_image_w < 1 || sd->env_image_h < 1 || sd->env_image_w > PT_ENV_MAX_W)
            return fail(PT_E_MISSING_INPUT, "missing input: env_image (PT_ENV_IMAGE needs its picture, 1 <= width <= 4096)");
        for (long long k = 0; k < 3LL * sd->env_image_w * sd->env_image_h; k++)
            if (!std::isfinite(sd->env_image[k]) || sd->env_image[k] < 0.0f)
                return fail(PT_E_INVALID, "Environment RGB samples must be finite and nonnegative");
    }
    if ((sd->env.type == PT_ENV_SKY || sd->env.type == PT_ENV_IMAGE) && sd->env.chart != PT_CHART_EQUIRECT &&
        sd->env.chart != PT_CHART_OCTAHEDRAL)
        return fail(PT_E_INVALID, "pt_env.chart must be PT_CHART_EQUIRECT or PT_CHART_OCTAHEDRAL");
    if (sd->env.type == PT_ENV_SKY && (sd->env.body < PT_SKY_SUN || sd->env.body > PT_SKY_WHITE))
        return fail(PT_E_INVALID, "pt_env.body must be one of PT_SKY_*");
    if ((sd->spectral_atlas == nullptr) != (sd->spectral_z_nodes == nullptr))
        return fail(PT_E_INVALID, "pt_scene_desc: spectral_atlas and spectral_z_nodes come together");
    if ((sd->light_boxes == nullptr) != (sd->light_powers == nullptr))
        return fail(PT_E_INVALID, "pt_scene_desc: light_boxes and light_powers come together");
    return PT_OK;
}

// Environments sampled through an importance table: the procedural sky and the picture.
static bool env_tabulated(int type) { return type == PT_ENV_SKY || type == PT_ENV_IMAGE; }

// Build every device structure of one scene (wide BVH, containment, env table) into SS.
// The importance-table chain after the luminance pass (components/env/importance.ts:44-127):
// tent blur, MIS compensation when asked, row and marginal CDFs. The table (cond, marg and the
// mean / total pair) is allocated in SS; the passes' scratch (blurred luminance, row masses,
// the f64 row weights sin(theta_j) and texel solid angles, the host builder's expressions
// evaluated on the host, H values each) in tmp, which the caller keeps until the stream has
// run the chain.
// The guides of the table's CDFs: block 0 the marginal, block 1 + j row j. Entry b of a guide is
// the first index whose CDF value exceeds b / K (n when none does), by the sampler's own search.
__global__ void k_env_guide(const float* __restrict__ cond, const float* __restrict__ marg, int W, int H, int km, int kc,
                            int* __restrict__ gm, int* __restrict__ gc) {
    const int r = blockIdx.x;
    const float* c = r == 0 ? marg : cond + (size_t)(r - 1) * W;
    const int n = r == 0 ? H : W, k = r == 0 ? km : kc;
    int* g = r == 0 ? gm : gc + (size_t)(r - 1) * (kc + 1);
    for (int b = threadIdx.x; b <= k; b += blockDim.x) {
        const float x = (float)b / (float)k;
        int lo = 0, hi = n;
        while (lo < hi) { const int mid = (lo + hi) >> 1; if (x >= c[mid]) lo = mid + 1; else hi = mid; }
        g[b] = lo;
    }
}
static int guide_buckets(int n) { int k = 1; while (k < n) k <<= 1; return k; }

static pt_status env_table_chain(float* Y, int W, int H, bool oct, bool compensated, SceneStore& SS, SceneStore& tmp,
                                 DevScene& D) {
    const long long n = (long long)W * H;
    std::vector<double> sin_t((size_t)H), d_omega((size_t)H);
    for (int j = 0; j < H; j++) {
        sin_t[j] = oct ? 1.0 : std::sin(M_PI * (j + 0.5) / H);
        d_omega[j] = oct ? (4 * M_PI) / (W * H) : (2 * M_PI / W) * (M_PI / H) * std::sin(M_PI * (j + 0.5) / H);
    }
    pt_status st;
    float *B, *cond, *marg;
    double *rows, *aux;
    const double *d_sin, *d_dom;
    if ((st = tmp.alloc((size_t)n, &B)) != PT_OK || (st = SS.alloc((size_t)n, &cond)) != PT_OK ||
        (st = SS.alloc((size_t)H, &marg)) != PT_OK || (st = tmp.alloc((size_t)H, &rows)) != PT_OK ||
        (st = SS.alloc(2, &aux)) != PT_OK || (st = tmp.upload(sin_t.data(), sin_t.size(), &d_sin)) != PT_OK ||
        (st = tmp.upload(d_omega.data(), d_omega.size(), &d_dom)) != PT_OK)
        return st;
    hipLaunchKernelGGL(k_env_blur, dim3(grid_for(n, 256)), dim3(256), 0, SS.stream, (const float*)Y, B, W, H, oct ? 1 : 0);
    if (compensated) {
        hipLaunchKernelGGL(k_env_mean, dim3(1), dim3(256), 0, SS.stream, (const float*)B, d_dom, W, H, aux);
        hipLaunchKernelGGL(k_env_compensate, dim3(grid_for(n, 256)), dim3(256), 0, SS.stream, B, (int)n,
                           (const double*)aux);
    }
    hipLaunchKernelGGL(k_env_rows, dim3(grid_for(H, 64)), dim3(64), 0, SS.stream, (const float*)B, d_sin, W, H, cond, rows);
    hipLaunchKernelGGL(k_env_marginal, dim3(1), dim3(64), 0, SS.stream, (const double*)rows, H, marg, aux + 1);
    const int km = guide_buckets(H), kc = guide_buckets(W);
    int *gm, *gc;
    if ((st = SS.alloc((size_t)km + 1, &gm)) != PT_OK || (st = SS.alloc((size_t)H * (kc + 1), &gc)) != PT_OK) return st;
    hipLaunchKernelGGL(k_env_guide, dim3(H + 1), dim3(256), 0, SS.stream, (const float*)cond, (const float*)marg, W, H, km, kc, gm, gc);
    HIP_TRY(hipGetLastError());
    D.env_cond = cond;
    D.env_marg = marg;
    D.env_total = aux + 1;
    D.env_guide_marg = gm;
    D.env_guide_cond = gc;
    D.env_km = km;
    D.env_kc = kc;
    return PT_OK;
}

// The procedural sky's importance table, baked on the device into SS (k_env_* above): the
// sky's luminance at every texel centre, then the chain. ev0 / ev1, when given, bracket the
// kernels on SS.stream. Synchronises the stream before the scratch is released.
static pt_status bake_env_table(const pt_env& env, SceneStore& SS, DevScene& D, hipEvent_t ev0 = nullptr,
                                hipEvent_t ev1 = nullptr) {
    const bool oct = env.chart == PT_CHART_OCTAHEDRAL;
    const int H = env.table_h > 0 ? env.table_h : 256, W = oct ? H : (env.table_w > 0 ? env.table_w : 512);
    if (W > PT_ENV_MAX_W) return fail(PT_E_UNSUPPORTED, "environment table wider than 4096 texels");
    const long long n = (long long)W * H;
    SceneStore tmp;
    tmp.stream = SS.stream;
    pt_status st;
    float* Y;
    if ((st = tmp.alloc((size_t)n, &Y)) != PT_OK) return st;
    if (ev0) HIP_TRY(hipEventRecord(ev0, SS.stream));
    hipLaunchKernelGGL(k_env_luma, dim3(grid_for(n, 256)), dim3(256), 0, SS.stream, Y, W, H, env.body, oct ? 1 : 0);
    if ((st = env_table_chain(Y, W, H, oct, env.compensated != 0, SS, tmp, D)) != PT_OK) return st;
    if (ev1) HIP_TRY(hipEventRecord(ev1, SS.stream));
    HIP_TRY(hipStreamSynchronize(SS.stream));
    return PT_OK;
}

// The environment's live values on the device record: the uniforms env.intensity, env.selectProb,
// the sky's env.rotation with its derived (cos, sin) pair (features/environment.ts:36-45) and a
// constant sky's colour.
static void set_env_values(DevScene& D, const pt_env& e) {
    D.env_intensity = e.intensity;
    D.env_select_prob = e.select_prob;
    if (e.type == PT_ENV_SKY || e.type == PT_ENV_IMAGE) {
        D.env_rotation = e.rotation;
        D.env_rot_c = (float)std::cos((double)e.rotation);
        D.env_rot_s = (float)std::sin((double)e.rotation);
    } else if (e.type == PT_ENV_CONSTANT) {
        D.env_color = make_float4(e.color[0], e.color[1], e.color[2], 0.0f);
        D.env_s_color = e.s_color;
    }
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
    SS.host_lights.assign(sd->lights, sd->lights + sd->n_lights);
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
    if (sd->light_boxes && sd->n_lights > 0) {
        // the light tree (sceneData.ts:99-143 -> buildLightTree), built on the host
        std::vector<float> lt_nodes((size_t)8 * (2 * sd->n_lights - 1)), lt_trails((size_t)4 * sd->n_lights);
        int32_t nc = 0;
        if ((st = pt_light_tree_build(sd->light_boxes, sd->light_powers, sd->n_lights, lt_nodes.data(), lt_trails.data(),
                                      &nc, nullptr)) != PT_OK) return st;
        std::vector<uint2> trails((size_t)sd->n_lights);
        for (int l = 0; l < sd->n_lights; l++) trails[l] = make_uint2((uint32_t)lt_trails[4 * (size_t)l], (uint32_t)lt_trails[4 * (size_t)l + 1]);
        if ((st = SS.upload((const float4*)lt_nodes.data(), (size_t)2 * nc, &D.lt_nodes)) != PT_OK) return st;
        if ((st = SS.upload(trails.data(), trails.size(), &D.lt_trails)) != PT_OK) return st;
    }
    if (sd->spectral_atlas) {
        if ((st = SS.upload((const float4*)sd->spectral_atlas, (size_t)3 * 64 * 64 * 64, &D.spec_atlas)) != PT_OK) return st;
        if ((st = SS.upload(sd->spectral_z_nodes, 64, &D.spec_znodes)) != PT_OK) return st;
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
    std::vector<int4> pm_head((size_t)std::max(sd->n_objects, 1), make_int4(0, 0, 0, 0));
    std::vector<float4> pm_grid((size_t)std::max(sd->n_objects, 1), make_float4(0, 0, 0, 0));
    std::vector<int> pm_cells, pm_items;
    std::vector<float> pm_boxes;
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
        // the containment probe map over the same leaf triangles
        ProbeMap pm;
        build_probe_map(w.tris.data(), (int)(w.tris.size() / 12), (int)toff, pm);
        pm_head[o] = make_int4((int)pm_cells.size(), pm.gu, pm.gv, 0);
        pm_grid[o] = make_float4(pm.u0, pm.v0, pm.inv_u, pm.inv_v);
        const int item0 = (int)pm_items.size();
        for (int c : pm.cell_start) pm_cells.push_back(item0 + c);
        pm_items.insert(pm_items.end(), pm.items.begin(), pm.items.end());
        pm_boxes.insert(pm_boxes.end(), pm.boxes.begin(), pm.boxes.end());
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
    probe_basis(D.pm_basis);
    if ((st = SS.upload(pm_head.data(), pm_head.size(), &D.pm_head)) != PT_OK ||
        (st = SS.upload(pm_grid.data(), pm_grid.size(), &D.pm_grid)) != PT_OK ||
        (st = SS.upload(pm_cells.data(), pm_cells.size(), &D.pm_cells)) != PT_OK ||
        (st = SS.upload(pm_items.data(), pm_items.size(), &D.pm_items)) != PT_OK ||
        (st = SS.upload((const float4*)pm_boxes.data(), pm_boxes.size() / 4, &D.pm_boxes)) != PT_OK)
        return st;
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
        SS.env_first = SS.bufs.size();
        if ((st = bake_env_table(sd->env, SS, D)) != PT_OK) return st;
        SS.env_end = SS.bufs.size();
        D.env_chart = sd->env.chart == PT_CHART_OCTAHEDRAL ? PT_CHART_OCTAHEDRAL : PT_CHART_EQUIRECT;
        D.env_h = sd->env.table_h > 0 ? sd->env.table_h : 256;
        D.env_w = D.env_chart == PT_CHART_OCTAHEDRAL ? D.env_h : (sd->env.table_w > 0 ? sd->env.table_w : 512);
        D.env_compensated = sd->env.compensated;
        D.env_body = sd->env.body;
        set_env_values(D, sd->env);
    } else if (sd->env.type == PT_ENV_IMAGE) {
        // the picture (NEAREST lookups) and its importance table on the chart
        if ((st = SS.upload(sd->env_image, (size_t)3 * sd->env_image_w * sd->env_image_h, &D.env_img)) != PT_OK) return st;
        D.env_img_w = sd->env_image_w;
        D.env_img_h = sd->env_image_h;
        const bool oct = sd->env.chart == PT_CHART_OCTAHEDRAL;
        std::vector<float> resampled;
        const float* tex = sd->env_image;
        int W = sd->env_image_w, H = sd->env_image_h;
        if (oct) {
            W = H = sd->env.table_h > 0 ? sd->env.table_h : 256;
            if (W > PT_ENV_MAX_W) return fail(PT_E_UNSUPPORTED, "environment table wider than 4096 texels");
            resample_equirect_to_octahedral(sd->env_image, sd->env_image_w, sd->env_image_h, W, resampled);
            tex = resampled.data();
        }
        SS.env_first = SS.bufs.size();
        SceneStore tmp;
        tmp.stream = SS.stream;
        const long long n = (long long)W * H;
        const float* d_rgb;
        float* Y;
        if ((st = tmp.upload(tex, (size_t)(3 * n), &d_rgb)) != PT_OK || (st = tmp.alloc((size_t)n, &Y)) != PT_OK) return st;
        hipLaunchKernelGGL(k_env_luma_rgb, dim3(grid_for(n, 256)), dim3(256), 0, SS.stream, d_rgb, Y, n);
        if ((st = env_table_chain(Y, W, H, oct, sd->env.compensated != 0, SS, tmp, D)) != PT_OK) return st;
        HIP_TRY(hipStreamSynchronize(SS.stream));
        SS.env_end = SS.bufs.size();
        D.env_chart = oct ? PT_CHART_OCTAHEDRAL : PT_CHART_EQUIRECT;
        D.env_w = W;
        D.env_h = H;
        D.env_compensated = sd->env.compensated;
        set_env_values(D, sd->env);
    } else if (sd->env.type == PT_ENV_CONSTANT) {
        D.env_samplable = sd->env.samplable ? 1 : 0;
        set_env_values(D, sd->env);
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
    ctx->env_bufs.clear();
    ctx->has_scene = false;
    ctx->dev = DevScene{};
    SceneStore SS;
    if ((st = build_scene(sd, ctx->stream, SS)) != PT_OK) return st;
    for (size_t k = 0; k < SS.bufs.size(); k++)   // the env table apart: pt_env_table_upload replaces it
        (k >= SS.env_first && k < SS.env_end ? ctx->env_bufs : ctx->scene_bufs).push_back(std::move(SS.bufs[k]));
    ctx->dev = SS.dev;
    ctx->host_env = sd->env;
    ctx->stack = SS.stack;
    ctx->wide_depth = SS.wide_depth;
    ctx->has_ggx = SS.has_ggx;
    ctx->has_mirror = SS.has_mirror;
    ctx->has_rough = SS.has_rough;
    ctx->host_materials = std::move(SS.host_materials);
    ctx->host_lights = std::move(SS.host_lights);
    ctx->has_scene = true;
    return PT_OK;
}

}  // extern "C"

// The uniform selection table (computeSelectPdf with weight 1 per light + cumulative,
// features/lighting.ts:375-380, 537-542; envSelectionProbability's 1 / (N + 1), :488-518),
// each value the f64 expression rounded to fp32 as the reference's baked literal is.
static bool holds_uniform_selection(const DevScene& dev, const std::vector<pt_light>& lights) {
    const int n = (int)lights.size();
    if (n >= 2) {
        double acc = 0.0;
        for (int i = 0; i < n; i++) {
            const double p = 1.0 / (double)n;
            acc += p;
            if (lights[i].select_pdf != (float)p || lights[i].cdf != (float)acc) return false;
        }
    }
    const bool env_live = (env_tabulated(dev.env_type) || (dev.env_type == PT_ENV_CONSTANT && dev.env_samplable)) && n > 0;
    return !env_live || dev.env_select_prob == (float)std::min(0.99, std::max(0.01, 1.0 / (double)(n + 1)));
}

// The program-vs-scene rules, checked at renderer load when a scene is resident and again by
// every render, forward and backward call (a later pt_scene_upload may have replaced the
// scene): a spectral program lifts a coloured sky body through the coefficient atlas, which the
// scene must then carry, a program that names its
// envSampler.chart must sample the chart the sky table was built on, and a program with
// uniform light selection needs the scene's selection table to be the uniform one.
static pt_status check_program_scene(const pt_program& prog, const DevScene& dev, const std::vector<pt_light>& lights) {
    if (prog.light_selection != PT_SEL_POWER && prog.light_selection != PT_SEL_UNIFORM && prog.light_selection != PT_SEL_BVH)
        return fail(PT_E_INVALID, "unknown light selection");
    if (prog.light_selection == PT_SEL_BVH) {
        if (!lights.empty() && !dev.lt_nodes)
            return fail(PT_E_MISSING_INPUT, "missing input: light_boxes / light_powers (the light tree's inputs in pt_scene_desc)");
        const bool env_live = (env_tabulated(dev.env_type) || (dev.env_type == PT_ENV_CONSTANT && dev.env_samplable)) && !lights.empty();
        if (env_live && dev.env_select_prob != 0.5f)
            return fail(PT_E_INVALID, "the program selects lights through the light tree but the scene's pt_env.select_prob "
                                      "is not the tree's 1/2");
        if (prog.medium_light_sampling == PT_MLS_EQUIANGULAR)
            return fail(PT_E_UNSUPPORTED, "mediumLightSampling 'equiangular' with the light tree is not provided by this build");
    }
    if (prog.light_selection == PT_SEL_UNIFORM && !holds_uniform_selection(dev, lights))
        return fail(PT_E_INVALID, "the program selects lights uniformly but the scene's selection table (pt_light.select_pdf / "
                                  "cdf, pt_env.select_prob) is not the uniform one");
    if (prog.color == PT_COLOR_SPECTRAL && ((dev.env_type == PT_ENV_SKY && !pt_sky_body_grey(dev.env_body)) ||
                                            dev.env_type == PT_ENV_IMAGE) && !dev.spec_atlas)
        return fail(PT_E_MISSING_INPUT, "missing input: spectral_atlas (a spectral program lifts the procedural sky's "
                                        "colours through the coefficient atlas in pt_scene_desc)");
    if (prog.env_chart && env_tabulated(dev.env_type) && prog.env_chart - 1 != dev.env_chart)
        return fail(PT_E_INVALID, std::string("the program's envSampler.chart is '") +
                                      (prog.env_chart - 1 == PT_CHART_OCTAHEDRAL ? "octahedral" : "equirect") +
                                      "' but the scene's sky table was built on the other chart (pt_env.chart)");
    return PT_OK;
}

extern "C" {

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
    if (prog->width <= 0 || prog->height <= 0 || prog->max_bounces < 0) return fail(PT_E_INVALID, "bad film size or bounce budget");
    if (prog->aovs & ~(PT_AOV_ALBEDO | PT_AOV_NORMAL)) return fail(PT_E_INVALID, "unknown aov bits");
    if (prog->aovs && prog->color != PT_COLOR_RGB)
        return fail(PT_E_UNSUPPORTED, "first-hit guides (aovs) are provided for RGB programs only");
    if (prog->env_chart < 0 || prog->env_chart > PT_CHART_OCTAHEDRAL + 1) return fail(PT_E_INVALID, "unknown env chart");
    if (prog->accumulation != PT_ACC_AVERAGE && prog->accumulation != PT_ACC_ONESHOT) return fail(PT_E_INVALID, "unknown accumulation");
    if (prog->accumulation == PT_ACC_ONESHOT && prog->variance)
        return fail(PT_E_INVALID, "accumulation 'oneshot' does not support variance (it keeps no statistic)");
    if (ctx->has_scene)
        if (pt_status sc = check_program_scene(*prog, ctx->dev, ctx->host_lights); sc != PT_OK) return sc;
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
    ctx->renderers.push_back(r);
    *out = r;
    return PT_OK;
}

void pt_renderer_unload(pt_renderer* r) {
    if (!r) return;
    pt_ctx* ctx = r->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->renderers.erase(std::remove(ctx->renderers.begin(), ctx->renderers.end(), r), ctx->renderers.end());
    delete r;
}

}  // extern "C"

// The path state of n slots carved from one device block, each piece 256-byte aligned: the
// queue counters and the ray / traversal statistics first, then the SoA arrays (ray, throughput,
// radiance, rng, MIS state, flags, hit, two ray queues, the shadow queue, medium, wavelengths,
// eta, the batch's scene ids, one queue per shading category and the guide samples the program
// asks for). Returns the bytes needed; with a null base only the size is computed.
constexpr size_t STATE_HEAD = STATE_HEAD_BYTES;
static size_t carve_paths(const pt_program& prog, char* base, long long n, Paths& P) {
    size_t off = 0;
    auto take = [&](size_t bytes) -> void* {
        void* p = base ? base + off : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    const size_t m = (size_t)n;
    P.counts = (int*)take(N_COUNTS * CNT_STRIDE * sizeof(int));
    P.hist = (unsigned long long*)take(2 * sizeof(unsigned long long));
    P.trav = (unsigned long long*)take(2 * sizeof(unsigned long long));
    P.ray_o = (float4*)take(16 * m); P.ray_d = (float4*)take(16 * m); P.thr = (float4*)take(16 * m);
    P.rad = (float4*)take(16 * m); P.rng = (uint2*)take(8 * m); P.prev = (float4*)take(16 * m);
    P.flags = (int*)take(4 * m); P.hit = (float2*)take(8 * m); P.region = (int2*)take(8 * m);
    P.q[0] = (int*)take(4 * m); P.q[1] = (int*)take(4 * m);
    P.s_path = (int*)take(4 * m); P.s_o = (float4*)take(16 * m); P.s_d = (float4*)take(16 * m);
    P.s_c = (float4*)take(16 * m); P.s_l = (float4*)take(16 * m); P.medium = (int*)take(4 * m);
    P.wl = (float4*)take(16 * m); P.eta = (float*)take(4 * m); P.path_scene = (int*)take(4 * m);
    for (int c = 0; c < N_CAT; c++) P.catq[c] = (int*)take(4 * m);       // one queue per shading category
    for (int a = 0; a < 2; a++) P.aov[a] = (prog.aovs & (1 << a)) ? (float4*)take(16 * m) : nullptr;
    P.prev_qn = prog.light_selection == PT_SEL_BVH ? (float4*)take(16 * m) : nullptr;   // the stored query's normal
    P.cap = (int)n;
    P.npix = prog.width * prog.height;
    return off;
}

static pt_status ensure_paths(pt_renderer* r, long long n) {
    if (r->cap >= n) return PT_OK;
    Paths P = r->paths;
    const size_t bytes = carve_paths(r->prog, nullptr, n, P);
    if (r->b_state.p) { HIP_TRY(hipFree(r->b_state.p)); r->b_state.p = nullptr; r->b_state.bytes = 0; }
    HIP_TRY(hipMalloc(&r->b_state.p, bytes));
    r->b_state.bytes = bytes;
    carve_paths(r->prog, (char*)r->b_state.p, n, P);
    HIP_TRY(hipMemset(r->b_state.p, 0, STATE_HEAD));
    r->paths = P;
    r->cap = n;
    return PT_OK;
}

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
static pt_status run_waves(pt_renderer* r, const pt_params* prm, uint32_t nframes, uint32_t sample0, bool adj,
                           int max_frames = 0, bool clear_stats = true) {
    pt_ctx* ctx = r->ctx;
    const pt_program& pg = r->prog;
    int F = std::max(1, std::min<int>(pg.frames_in_flight, (int)std::max<uint32_t>(nframes, 1)));
    if (max_frames > 0) F = std::min(F, max_frames);
    const bool spectral = pg.color == PT_COLOR_SPECTRAL;
    if (pt_status sc = check_program_scene(pg, ctx->dev, ctx->host_lights); sc != PT_OK) return sc;
    if (adj && pg.light_selection == PT_SEL_BVH)
        return fail(PT_E_UNSUPPORTED, "the adjoint replays power or uniform light selection only (not the light tree)");
    if (spectral && !ctx->dev.spec_curves)
        return fail(PT_E_MISSING_INPUT, "missing input: spectral_curves (the spectral program needs the CIE/D65 curves in pt_scene_desc)");
    if (clear_stats) {
        HIP_TRY(hipMemsetAsync(r->paths.hist, 0, 2 * sizeof(unsigned long long), ctx->stream));
        HIP_TRY(hipMemsetAsync(r->paths.trav, 0, 2 * sizeof(unsigned long long), ctx->stream));
    }
    const Camera cam = make_camera(prm);
    uint32_t done = 0;
    while (done < nframes) {
        int f = (int)std::min<uint32_t>((uint32_t)F, nframes - done);
        const uint32_t s0 = sample0 + done;
        const int v = (ctx->stack == 16 ? 0 : 4) | (ctx->dev.has_media ? 2 : 0) | (spectral ? 1 : 0);
        if (adj) {
            // the media, spectral and adjoint programs run with the 32-entry stack, which holds
            // every accepted BVH (depth <= 31); only the plain RGB path keeps two depths
            switch (v) {
                case 0: case 4: wave_adjoint(r, cam, f, s0); break;
                case 2: case 6: wave_adjoint_media(r, cam, f, s0); break;
                default: return fail(PT_E_UNSUPPORTED, "the adjoint runs RGB programs only");
            }
        } else {
            switch (v) {
                case 0: wave_rgb(r, cam, f, s0, 16); break;
                case 4: wave_rgb(r, cam, f, s0, 32); break;
                case 1: case 5: wave_spectral(r, cam, f, s0); break;
                case 2: case 6: wave_media(r, cam, f, s0); break;
                default: wave_media_spectral(r, cam, f, s0);
            }
        }
        HIP_TRY(hipGetLastError());
        done += (uint32_t)f;
    }
    return PT_OK;
}

// Calls on a renderer or batch whose device resources went with pt_ctx_reset.
static pt_status lost_error(const char* fn) {
    return fail(PT_E_CONTEXT_LOST, std::string(fn) + ": the context was reset (pt_ctx_reset); unload the handle and load it again");
}

extern "C" {

pt_status pt_render_frames(pt_renderer* r, const pt_params* prm, uint32_t nframes) {
    if (!r || !prm) return fail(PT_E_INVALID, "pt_render_frames: null argument");
    if (r->lost) return lost_error("pt_render_frames");
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

pt_status pt_render_frames_noalloc(pt_renderer* r, const pt_params* prm, uint32_t nframes) {
    if (!r || !prm) return fail(PT_E_INVALID, "pt_render_frames_noalloc: null argument");
    if (r->lost) return lost_error("pt_render_frames_noalloc");
    pt_ctx* ctx = r->ctx;
    if (!ctx->has_scene) return fail(PT_E_MISSING_INPUT, "missing input: no scene uploaded (pt_scene_upload)");
    if (pt_status sc = check_camera(prm); sc != PT_OK) return sc;
    const pt_program& pg = r->prog;
    const long long npix = (long long)pg.width * pg.height;
    if (npix > PT_NOALLOC_MAX_PATHS)
        return fail(PT_E_INVALID, "pt_render_frames_noalloc: the film has " + std::to_string(npix) +
                                      " pixels; the allocation-free path state holds " + std::to_string(PT_NOALLOC_MAX_PATHS));
    if (!ctx->noalloc_state) return fail(PT_E_DEVICE, "pt_render_frames_noalloc: no allocation-free path state on this device");
    const int F = (int)std::min<long long>(std::max(1, pg.frames_in_flight), PT_NOALLOC_MAX_PATHS / npix);
    Paths P = r->paths;
    if (carve_paths(pg, nullptr, npix * F, P) > NOALLOC_BYTES)
        return fail(PT_E_INVALID, "pt_render_frames_noalloc: the path state exceeds the allocation-free array");
    carve_paths(pg, (char*)ctx->noalloc_state, npix * F, P);
    HIP_TRY(hipSetDevice(ctx->device));
    // counters and ray statistics start at zero: a fill kernel, so the call may be captured
    hipLaunchKernelGGL(k_fill_f4, dim3(1), dim3(SBLOCK), 0, ctx->stream, (float4*)ctx->noalloc_state,
                       (long long)(STATE_HEAD / sizeof(float4)));
    const Paths own = r->paths;
    const long long own_cap = r->cap;
    const bool timing = ctx->timing;
    r->paths = P;
    r->cap = npix * F;
    ctx->timing = false;                  // no event objects inside the call
    pt_status st = run_waves(r, prm, nframes, r->sample_count, false, F, false);
    ctx->timing = timing;
    r->paths = own;
    r->cap = own_cap;
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
    if (r->lost) return lost_error("pt_backward");
    pt_ctx* ctx = r->ctx;
    if (!ctx->has_scene) return fail(PT_E_MISSING_INPUT, "missing input: no scene uploaded (pt_scene_upload)");
    const pt_program& pg = r->prog;
    if (pg.color != PT_COLOR_RGB) return fail(PT_E_UNSUPPORTED, "pt_backward: the adjoint runs RGB programs only");
    if (pg.light_selection == PT_SEL_BVH)
        return fail(PT_E_UNSUPPORTED, "pt_backward: the adjoint replays power or uniform light selection only (not the light tree)");
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

static pt_status batch_check_config(const pt_batch* b, const pt_batch_config& c) {
    if (c.trace_block != 64 && c.trace_block != 128 && c.trace_block != 256) return fail(PT_E_INVALID, "pt_batch: trace_block must be 64, 128 or 256");
    if (c.stack != 16 && c.stack != 32) return fail(PT_E_INVALID, "pt_batch: stack must be 16 or 32");
    if (c.stack < b->min_stack) return fail(PT_E_INVALID, "pt_batch: stack shallower than a scene's BVH needs");
    if (c.wave_paths < b->max_npix || c.wave_paths > (1 << 28)) return fail(PT_E_INVALID, "pt_batch: wave_paths must hold the largest film and be <= 2^28");
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
    // every record's ABI version before any field that a newer ABI added is read
    for (int s = 0; s < n; s++)
        if (programs[s].abi_version != PT_ABI_VERSION || scenes[s].abi_version != PT_ABI_VERSION)
            return fail(PT_E_INVALID, "pt_batch_create: ABI version mismatch");
    for (int s = 0; s < n; s++) {
        if (programs[s].aovs) return fail(PT_E_UNSUPPORTED, "pt_batch_create: first-hit guides (aovs) are not batched");
        if (programs[s].accumulation != PT_ACC_AVERAGE)
            return fail(PT_E_UNSUPPORTED, "pt_batch_create: batches average their samples (accumulation 'average')");
        if (programs[s].env_chart < 0 || programs[s].env_chart > PT_CHART_OCTAHEDRAL + 1)
            return fail(PT_E_INVALID, "pt_batch_create: unknown env chart");
        if (programs[s].env_chart && env_tabulated(scenes[s].env.type) && programs[s].env_chart - 1 != scenes[s].env.chart)
            return fail(PT_E_INVALID, "pt_batch_create: a program's envSampler.chart differs from its scene's sky table chart");
    }
    const pt_program& p0 = programs[0];
    for (int s = 0; s < n; s++) {
        const pt_program& p = programs[s];
        if (p.estimator != p0.estimator || p.max_bounces != p0.max_bounces || p.rr_start != p0.rr_start || p.rr_max_survival != p0.rr_max_survival)
            return fail(PT_E_UNSUPPORTED, "pt_batch_create: every program must share estimator, maxBounces and Russian roulette");
        if (p.color != PT_COLOR_RGB) return fail(PT_E_UNSUPPORTED, "pt_batch_create: batches run RGB programs");
        if (p.light_selection == PT_SEL_BVH) return fail(PT_E_UNSUPPORTED, "pt_batch_create: batches select lights by power or uniformly");
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
        if (st == PT_OK) st = check_program_scene(programs[s], store->dev, store->host_lights);
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
    b->cfg.wave_paths = std::max(b->max_npix, 1 << 26);
    b->cfg.trace_block = 128;
    b->cfg.trace_blocks_per_cu = 8;
    b->cfg.shade_blocks_per_cu = 8;
    b->cfg.stack = b->min_stack;
    b->cfg.use_graph = 1;
    ctx->batches.push_back(b);
    *out = b;
    return PT_OK;
}

void pt_batch_destroy(pt_batch* b) {
    if (!b) return;
    pt_ctx* ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->batches.erase(std::remove(ctx->batches.begin(), ctx->batches.end(), b), ctx->batches.end());
    delete b;
}

pt_status pt_batch_set_config(pt_batch* b, const pt_batch_config* cfg) {
    if (!b || !cfg) return fail(PT_E_INVALID, "pt_batch_set_config: null argument");
    if (b->lost) return lost_error("pt_batch_set_config");
    pt_status st = batch_check_config(b, *cfg);
    if (st != PT_OK) return st;
    b->cfg = *cfg;
    return PT_OK;
}

pt_status pt_batch_get_config(const pt_batch* b, pt_batch_config* cfg) {
    if (!b || !cfg) return fail(PT_E_INVALID, "pt_batch_get_config: null argument");
    if (b->lost) return lost_error("pt_batch_get_config");
    *cfg = b->cfg;
    return PT_OK;
}

pt_status pt_batch_render(pt_batch* b, const uint32_t* spp) {
    if (!b || !spp) return fail(PT_E_INVALID, "pt_batch_render: null argument");
    if (b->lost) return lost_error("pt_batch_render");
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
    if (b->lost) return lost_error("pt_batch_read_export");
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
    if (b->lost) return lost_error("pt_batch_autotune");
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
    // two stages: the wave size at the starting launch shape, then the launch shape (trace block,
    // blocks per CU, stack depth) at the best wave size
    auto trial = [&](const pt_batch_config& c) -> pt_status {
        if (batch_check_config(b, c) != PT_OK) return PT_OK;
        b->cfg = c;
        pt_status st = pt_batch_render(b, probe.data());       // build + warm
        if (st != PT_OK) return st;
        HIP_TRY(hipEventRecord(e0, ctx->stream));
        if ((st = pt_batch_render(b, probe.data())) != PT_OK) return st;
        HIP_TRY(hipEventRecord(e1, ctx->stream));
        HIP_TRY(hipEventSynchronize(e1));
        float t = 0.0f;
        HIP_TRY(hipEventElapsedTime(&t, e0, e1));
        if (results && ms && count < max_results) { results[count] = c; ms[count] = t; }
        count++;
        if (t < best_ms) { best_ms = t; winner = c; }
        return PT_OK;
    };
    pt_status st = PT_OK;
    const int waves[] = {1 << 23, 1 << 24, 1 << 25, 1 << 26, 1 << 27, 1 << 28};
    for (int wp : waves) {
        pt_batch_config c = start;
        c.wave_paths = std::max(wp, b->max_npix);
        c.use_graph = 1;
        if ((st = trial(c)) != PT_OK) break;
    }
    const pt_batch_config stage1 = winner;
    const int tbs[] = {64, 128, 256};
    const int bpcs[] = {4, 8, 16};
    const int stacks[] = {16, 32};
    for (int tb : tbs)
        for (int bpc : bpcs)
            for (int sk : stacks) {
                if (st != PT_OK) break;
                pt_batch_config c = stage1;
                c.trace_block = tb;
                c.trace_blocks_per_cu = std::max(1, bpc * 128 / tb);   // bpc x 128 lanes per CU
                c.stack = sk;
                st = trial(c);
            }
    if (st != PT_OK) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return st; }
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

pt_status pt_scene_update_lights(pt_ctx* ctx, const pt_light* lights, int32_t n) {
    if (!ctx || (!lights && n)) return fail(PT_E_INVALID, "pt_scene_update_lights: null argument");
    if (!ctx->has_scene) return fail(PT_E_MISSING_INPUT, "missing input: no scene uploaded (pt_scene_upload)");
    if (n != (int32_t)ctx->host_lights.size()) return fail(PT_E_INVALID, "pt_scene_update_lights: light count differs from the scene");
    float prev_cdf = 0.0f;
    for (int i = 0; i < n; i++) {
        const pt_light& a = ctx->host_lights[i];
        const pt_light& b = lights[i];
        if (a.kind != b.kind || a.region != b.region)
            return fail(PT_E_INVALID, "pt_scene_update_lights: kind and region must not change");
        // geometry rows are constant by contract (resolveLightValues, lighting.ts:337-348)
        if (std::memcmp(a.corner, b.corner, sizeof a.corner) || std::memcmp(a.edge1, b.edge1, sizeof a.edge1) ||
            std::memcmp(a.edge2, b.edge2, sizeof a.edge2) || std::memcmp(a.normal, b.normal, sizeof a.normal) ||
            std::memcmp(&a.area, &b.area, sizeof a.area))
            return fail(PT_E_INVALID, "pt_scene_update_lights: a light's geometry must not change (upload the scene again)");
        for (int c = 0; c < 3; c++)
            if (!std::isfinite(b.radiance[c]) || b.radiance[c] < 0.0f)
                return fail(PT_E_INVALID, "pt_scene_update_lights: emission must be finite and >= 0");
        if (!(b.select_pdf > 0.0f && b.select_pdf <= 1.0f) || !(b.cdf >= prev_cdf && b.cdf <= 1.0f + 1e-5f))
            return fail(PT_E_INVALID, "pt_scene_update_lights: select_pdf must lie in (0, 1] and cdf must rise");
        prev_cdf = b.cdf;
    }
    if (n > 0 && std::fabs(prev_cdf - 1.0f) > 1e-5f) return fail(PT_E_INVALID, "pt_scene_update_lights: the selection cdf must end at 1");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n) HIP_TRY(hipMemcpy((void*)ctx->dev.lights, lights, sizeof(pt_light) * (size_t)n, hipMemcpyHostToDevice));
    ctx->host_lights.assign(lights, lights + n);
    return PT_OK;
}

pt_status pt_scene_update_env(pt_ctx* ctx, const pt_env* env) {
    if (!ctx || !env) return fail(PT_E_INVALID, "pt_scene_update_env: null argument");
    if (!ctx->has_scene) return fail(PT_E_MISSING_INPUT, "missing input: no scene uploaded (pt_scene_upload)");
    const pt_env& a = ctx->host_env;
    if (env->type != a.type || env->samplable != a.samplable)
        return fail(PT_E_INVALID, "pt_scene_update_env: type and samplable must not change");
    if (env_tabulated(a.type) && (env->chart != a.chart || env->compensated != a.compensated || env->body != a.body ||
                                 env->table_h != a.table_h || (a.chart != PT_CHART_OCTAHEDRAL && env->table_w != a.table_w)))
        return fail(PT_E_INVALID, "pt_scene_update_env: the sky's body and its table's chart, size and compensation must not change "
                                  "(pt_env_table_upload rebuilds the table)");
    if (!std::isfinite(env->intensity) || env->intensity < 0.0f || !std::isfinite(env->rotation) ||
        !(env->select_prob >= 0.0f && env->select_prob <= 1.0f))
        return fail(PT_E_INVALID, "pt_scene_update_env: intensity >= 0, a finite rotation and select_prob in [0, 1]");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));   // kernels take the scene record by value at launch
    set_env_values(ctx->dev, *env);
    const int32_t chart = a.chart, tw = a.table_w, th = a.table_h, comp = a.compensated;
    ctx->host_env = *env;
    ctx->host_env.chart = chart;    // the resident table's (a table upload may have changed them)
    ctx->host_env.table_w = tw;
    ctx->host_env.table_h = th;
    ctx->host_env.compensated = comp;
    return PT_OK;
}

pt_status pt_env_table_upload(pt_ctx* ctx, const float* rgb, int32_t W, int32_t H, int32_t chart, int32_t compensated,
                              double* total_weight) {
    if (!ctx || !rgb) return fail(PT_E_INVALID, "pt_env_table_upload: null argument");
    if (!ctx->has_scene) return fail(PT_E_MISSING_INPUT, "missing input: no scene uploaded (pt_scene_upload)");
    if (!env_tabulated(ctx->host_env.type))
        return fail(PT_E_INVALID, "pt_env_table_upload: the resident environment has no importance table (PT_ENV_SKY or PT_ENV_IMAGE)");
    if (chart != PT_CHART_EQUIRECT && chart != PT_CHART_OCTAHEDRAL)
        return fail(PT_E_INVALID, "pt_env_table_upload: chart must be PT_CHART_EQUIRECT or PT_CHART_OCTAHEDRAL");
    if (W < 1 || H < 1 || W > PT_ENV_MAX_W) return fail(PT_E_INVALID, "pt_env_table_upload: 1 <= width <= 4096 and height >= 1");
    if (chart == PT_CHART_OCTAHEDRAL && W != H) return fail(PT_E_INVALID, "pt_env_table_upload: the octahedral chart is square");
    const long long n = (long long)W * H;
    for (long long k = 0; k < 3 * n; k++)
        if (!std::isfinite(rgb[k]) || rgb[k] < 0.0f)
            return fail(PT_E_INVALID, "Environment RGB samples must be finite and nonnegative");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    SceneStore SS, tmp;
    SS.stream = tmp.stream = ctx->stream;
    DevScene D = ctx->dev;
    pt_status st;
    const float* d_rgb;
    float* Y;
    if ((st = tmp.upload(rgb, (size_t)(3 * n), &d_rgb)) != PT_OK || (st = tmp.alloc((size_t)n, &Y)) != PT_OK) return st;
    hipLaunchKernelGGL(k_env_luma_rgb, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, d_rgb, Y, n);
    if ((st = env_table_chain(Y, W, H, chart == PT_CHART_OCTAHEDRAL, compensated != 0, SS, tmp, D)) != PT_OK) return st;
    double total = 0.0;
    HIP_TRY(hipMemcpyAsync(&total, D.env_total, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    // every kernel that read the previous table has finished (the stream synchronised above)
    ctx->env_bufs = std::move(SS.bufs);
    ctx->dev.env_cond = D.env_cond;
    ctx->dev.env_marg = D.env_marg;
    ctx->dev.env_total = D.env_total;
    ctx->dev.env_guide_marg = D.env_guide_marg;
    ctx->dev.env_guide_cond = D.env_guide_cond;
    ctx->dev.env_km = D.env_km;
    ctx->dev.env_kc = D.env_kc;
    ctx->dev.env_chart = chart;
    ctx->dev.env_w = W;
    ctx->dev.env_h = H;
    ctx->dev.env_compensated = compensated ? 1 : 0;
    ctx->host_env.chart = chart;
    ctx->host_env.table_w = W;
    ctx->host_env.table_h = H;
    ctx->host_env.compensated = compensated ? 1 : 0;
    if (total_weight) *total_weight = total;
    return PT_OK;
}

// A renderer's device resources, released by a context reset.
static void release_renderer(pt_renderer* r) {
    r->film.release();
    r->var_film.release();
    for (DevBuf& b : r->aov_film) b.release();
    r->b_state.release();
    for (DevBuf* b : {&r->b_adjD, &r->b_adjG, &r->b_adjsD, &r->b_delta, &r->b_pixgrad, &r->b_grad}) b->release();
    r->cap = 0;
    r->adj_cap = 0;
    r->paths = Paths{};
    r->sample_count = 0;
    r->lost = true;
}

pt_status pt_ctx_reset(pt_ctx* ctx) {
    if (!ctx) return fail(PT_E_INVALID, "pt_ctx_reset: null context");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (pt_renderer* r : ctx->renderers) release_renderer(r);
    for (pt_batch* b : ctx->batches) {
        b->drop_graph();
        b->stores.clear();
        b->d_scenes.release();
        b->d_cams.release();
        b->d_entries.release();
        b->films.release();
        if (b->r) release_renderer(b->r);
        b->lost = true;
    }
    ctx->scene_bufs.clear();
    ctx->env_bufs.clear();
    ctx->dev = DevScene{};
    ctx->has_scene = false;
    ctx->host_materials.clear();
    ctx->host_lights.clear();
    ctx->host_env = pt_env{};
    for (DevBuf* b : {&ctx->trav, &ctx->rays_tmp, &ctx->t_tmp, &ctx->p_tmp}) b->release();
    for (auto& kv : ctx->timings)
        for (auto& pr : kv.second.ev) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
    ctx->timings.clear();
    ctx->timing_order.clear();
    return PT_OK;
}

pt_status pt_read_export(pt_renderer* r, const char* name, float* out, size_t out_floats) {
    if (!r || !name || !out) return fail(PT_E_INVALID, "pt_read_export: null argument");
    if (r->lost) return lost_error("pt_read_export");
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
    if (r->lost) return lost_error("pt_clear_accumulation");
    HIP_TRY(hipSetDevice(r->ctx->device));
    HIP_TRY(hipMemsetAsync(r->film.p, 0, r->film.bytes, r->ctx->stream));
    if (r->var_film.p) HIP_TRY(hipMemsetAsync(r->var_film.p, 0, r->var_film.bytes, r->ctx->stream));
    for (int a = 0; a < 2; a++)
        if (r->aov_film[a].p) HIP_TRY(hipMemsetAsync(r->aov_film[a].p, 0, r->aov_film[a].bytes, r->ctx->stream));
    r->sample_count = 0;
    return PT_OK;
}

uint32_t pt_sample_count(const pt_renderer* r) { return r && !r->lost ? r->sample_count : 0; }

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
    if (r->lost) return lost_error("pt_last_ray_counts");
    unsigned long long h[2] = {0, 0};
    if (r->b_state.p) {
        HIP_TRY(hipMemcpyAsync(h, r->paths.hist, sizeof h, hipMemcpyDeviceToHost, r->ctx->stream));
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
    if (r->b_state.p) {
        HIP_TRY(hipMemcpyAsync(h, r->paths.trav, sizeof h, hipMemcpyDeviceToHost, r->ctx->stream));
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
// scene_region_at over n points of the resident scene (the containment probe's parity hook).
__global__ void __launch_bounds__(256) k_debug_region(DevScene S, const float* __restrict__ pts, int n, int* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = region_at(S, mk3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
}
// environment_sample (sampler_cdf.glsl:61-80) at n (x0, x1) pairs of the resident scene's
// environment: the sampled direction and its pdf (the guided CDF search's parity hook).
__global__ void __launch_bounds__(256) k_debug_env_sample(DevScene S, const float* __restrict__ xi, int n, float* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    Wl wl{};
    const LSample ls = environment_sample<false>(S, xi[2 * i], xi[2 * i + 1], wl);
    out[4 * i] = ls.wi.x; out[4 * i + 1] = ls.wi.y; out[4 * i + 2] = ls.wi.z; out[4 * i + 3] = ls.pdf;
}
}  // namespace

extern "C" {
pt_status pt_debug_region_at(pt_ctx* ctx, const float* pts, int n, int32_t* out) {
    if (!ctx || !pts || !out || n < 0) return fail(PT_E_INVALID, "pt_debug_region_at: bad argument");
    if (!ctx->has_scene) return fail(PT_E_MISSING_INPUT, "missing input: no scene uploaded (pt_scene_upload)");
    if (n == 0) return PT_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    DevBuf dp, dout;
    HIP_TRY(hipMalloc(&dp.p, (size_t)n * 12));
    HIP_TRY(hipMalloc(&dout.p, (size_t)n * 4));
    HIP_TRY(hipMemcpyAsync(dp.p, pts, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_debug_region, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, ctx->dev, (const float*)dp.p, n,
                       (int*)dout.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, dout.p, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PT_OK;
}

pt_status pt_debug_env_sample(pt_ctx* ctx, const float* xi, int n, float* out) {
    if (!ctx || !xi || !out || n < 0) return fail(PT_E_INVALID, "pt_debug_env_sample: bad argument");
    if (!ctx->has_scene) return fail(PT_E_MISSING_INPUT, "missing input: no scene uploaded (pt_scene_upload)");
    if (!env_tabulated(ctx->dev.env_type) && ctx->dev.env_type != PT_ENV_CONSTANT)
        return fail(PT_E_INVALID, "pt_debug_env_sample: the scene has no environment to sample");
    if (n == 0) return PT_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    DevBuf dx, dout;
    HIP_TRY(hipMalloc(&dx.p, (size_t)n * 8));
    HIP_TRY(hipMalloc(&dout.p, (size_t)n * 16));
    HIP_TRY(hipMemcpyAsync(dx.p, xi, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_debug_env_sample, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, ctx->dev, (const float*)dx.p, n,
                       (float*)dout.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, dout.p, (size_t)n * 16, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
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
