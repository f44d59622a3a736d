// batch_kernels.hip — the launch sequence of a multi-scene batch: mixed-scene waves over the
// batch's path slots, in the traversal block size and stack depth of its configuration.
#include "wave.h"

template <int EST>
static void batch_shade(pt_batch* b, int cur) {
    pt_ctx* ctx = b->ctx;
    const pt_program& pg = b->progs[0];
    const DevScene& S0 = b->stores[0]->dev;
    const int g = ctx->cu_count * b->cfg.shade_blocks_per_cu;
    Paths& P = b->r->paths;
    hipStream_t st = ctx->stream;
#define PT_BSHADE(CATV) hipLaunchKernelGGL((k_shade<EST, CATV, false, false, false, true>), dim3(g), dim3(SBLOCK), 0, st, S0, P, cur, pg.max_bounces, pg.rr_start, pg.rr_max_survival, pg.max_null_crossings, 0)
    if (b->any_dielectric || b->any_rough) hipLaunchKernelGGL(k_probe<true>, dim3(g), dim3(SBLOCK), 0, st, S0, P);
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
                case PT_EST_PT: batch_shade<PT_EST_PT>(b, cur); break;
                case PT_EST_NEE: batch_shade<PT_EST_NEE>(b, cur); break;
                default: batch_shade<PT_EST_MIS>(b, cur);
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

void ptmi::batch_launch(pt_batch* b) {
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
