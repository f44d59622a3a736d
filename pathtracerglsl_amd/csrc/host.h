// host.h — host-side state shared by the library's translation units: device buffers, the
// context / renderer / batch records, error reporting, launch timing and the per-variant
// wave launchers (each compiled in a unit of its own).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "dev_common.h"
#include "pipeline.h"

namespace ptmi {

// Records msg as the calling thread's last error (pt_last_error) and returns s.
pt_status fail(pt_status s, const std::string& msg);

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return ptmi::fail(e_ == hipErrorOutOfMemory ? PT_E_OOM : PT_E_DEVICE,            \
                              std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() { if (p) (void)hipFree(p); p = nullptr; bytes = 0; }
};

inline int grid_for(long long n, int block) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return (int)std::min<long long>(g, MAX_GRID);
}

}  // namespace ptmi

using namespace ptmi;

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
    std::vector<pt_light> host_lights;
    size_t env_first = 0, env_end = 0;   // bufs[env_first, env_end): the environment table
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
    void* noalloc_state = nullptr;   // device address of this GPU's allocation-free path state
    std::vector<std::unique_ptr<DevBuf>> scene_bufs;
    DevScene dev{};
    bool has_scene = false;
    int stack = 16;
    uint32_t wide_depth = 0;
    int cu_count = 256;
    int lds_per_cu = 160 * 1024;   // LDS bytes per CU (device property)
    int shade_blocks = 2048;
    bool has_ggx = false;
    bool has_mirror = false;
    bool has_rough = false;
    std::vector<pt_material> host_materials;   // the uploaded records (adjoint checks, updates)
    std::vector<pt_light> host_lights;         // the uploaded light records (selection-table check)
    pt_env host_env{};                         // the resident environment (live updates, table uploads)
    std::vector<std::unique_ptr<DevBuf>> env_bufs;   // the resident environment table
    std::vector<pt_renderer*> renderers;       // loaded renderers (released by pt_ctx_reset)
    std::vector<pt_batch*> batches;            // live batches (released by pt_ctx_reset)
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

// Persistent traversal grid: as many 2-wave blocks per CU as the LDS stacks leave room for, at
// most 10 (a 16-entry stack takes 16 KB per block, so 10 blocks fill a 160 KB CU).
inline int trace_grid(const pt_ctx* c, int stack) {
    const int per_block = stack * TBLOCK * (int)sizeof(uint2);
    const int per_cu = std::max(1, std::min(10, c->lds_per_cu / per_block));
    return c->cu_count * per_cu;
}

struct pt_renderer {
    pt_ctx* ctx = nullptr;
    pt_program prog{};
    DevBuf film, var_film, aov_film[2];
    uint32_t sample_count = 0;
    long long cap = 0;    // path slots allocated
    DevBuf b_state;       // the path state of cap slots, one block carved by carve_paths
    // adjoint buffers (allocated by the first pt_backward)
    DevBuf b_adjD, b_adjG, b_adjsD, b_delta, b_pixgrad, b_grad;
    long long adj_cap = 0;
    int adj_mode = PT_REDUCE_DETERMINISTIC;
    Paths paths{};
    bool count_traversal = false;
    bool lost = false;    // its device resources went with pt_ctx_reset
};

#define TIMED(ctx, name, items, launch)             \
    do {                                            \
        hipEvent_t ev_b_ = nullptr;                 \
        (ctx)->ev_begin(name, &ev_b_);              \
        launch;                                     \
        (ctx)->ev_end(name, ev_b_, items);          \
    } while (0)

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
    bool lost = false;    // its device resources went with pt_ctx_reset
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

namespace ptmi {
// One frame wave of a renderer's program (wave_*.hip): frames samples per pixel starting at
// sampleCount sample0, traced and accumulated (forward) or replayed and reduced (adjoint).
void wave_rgb(pt_renderer* r, const Camera& cam, int frames, uint32_t sample0, int stack);
void wave_spectral(pt_renderer* r, const Camera& cam, int frames, uint32_t sample0);
void wave_media(pt_renderer* r, const Camera& cam, int frames, uint32_t sample0);
void wave_media_spectral(pt_renderer* r, const Camera& cam, int frames, uint32_t sample0);
void wave_adjoint(pt_renderer* r, const Camera& cam, int frames, uint32_t sample0);
void wave_adjoint_media(pt_renderer* r, const Camera& cam, int frames, uint32_t sample0);
// Every wave of a batch's schedule, then the film updates (batch_kernels.hip).
void batch_launch(pt_batch* b);
}  // namespace ptmi
