// wave.h — host-side launch sequence of one frame wave (raygen, then per iteration bookkeep,
// extend, bin, the shade categories and the shadow rays; accumulate or the adjoint reduction).
// Included by the wave_*.hip units, each of which instantiates one or two variants.
#pragma once
#include "kernels.h"
#include "host.h"

template <int EST, bool MEDIA, bool SPEC, bool ADJ>
static void launch_shade(pt_renderer* r, int cur) {
    pt_ctx* ctx = r->ctx;
    const pt_program& pg = r->prog;
    const DevScene& S = ctx->dev;
    hipStream_t st = ctx->stream;
    const int g = ctx->shade_blocks;
    const int mn = pg.max_null_crossings;
#define PT_SHADE(name, CATV) TIMED(ctx, name, 0, hipLaunchKernelGGL((k_shade<EST, CATV, MEDIA, SPEC, ADJ>), dim3(g), dim3(SBLOCK), 0, st, S, r->paths, cur, pg.max_bounces, pg.rr_start, pg.rr_max_survival, mn, pg.medium_light_sampling))
    if (!MEDIA && (S.has_transmission || ctx->has_rough))
        TIMED(ctx, "probe", 0, hipLaunchKernelGGL(k_probe<false>, dim3(g), dim3(SBLOCK), 0, st, S, r->paths));
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
    if (MEDIA) hipLaunchKernelGGL(k_camera_medium, dim3(1), dim3(64), 0, st, S, P, cam);
    TIMED(ctx, "raygen", n,  hipLaunchKernelGGL((k_raygen<SPEC, ADJ>), dim3((unsigned)((n + SBLOCK - 1) / SBLOCK)), dim3(SBLOCK), 0, st, P, cam, W, H, (int)n, sample0, MEDIA ? 1 : 0, pg.wavelengths));
    if (MEDIA && (cam.ortho || cam.thin)) hipLaunchKernelGGL(k_ray_medium, dim3(ctx->shade_blocks), dim3(SBLOCK), 0, st, S, P, (int)n);
    int cur = 0;
    const int tgrid = trace_grid(ctx, STACK);
    const int mgrid = ctx->cu_count * 8;   // the media walker is grid-stride: more blocks than fit balance it
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
            if (inst) TIMED(ctx, "shadow_media", 0, hipLaunchKernelGGL((k_shadow_media<STACK, SPEC, ADJ, true>), dim3(mgrid), dim3(TBLOCK), 0, st, S, P));
            else TIMED(ctx, "shadow_media", 0, hipLaunchKernelGGL((k_shadow_media<STACK, SPEC, ADJ>), dim3(mgrid), dim3(TBLOCK), 0, st, S, P));
            hipLaunchKernelGGL(k_shadow_reset, dim3(1), dim3(64), 0, st, P);
        }
        switch (pg.estimator) {
            case PT_EST_PT: launch_shade<PT_EST_PT, MEDIA, SPEC, ADJ>(r, cur); break;
            case PT_EST_NEE: launch_shade<PT_EST_NEE, MEDIA, SPEC, ADJ>(r, cur); break;
            default: launch_shade<PT_EST_MIS, MEDIA, SPEC, ADJ>(r, cur);
        }
        if (pg.estimator != PT_EST_PT) {
            if (MEDIA && inst) TIMED(ctx, "shadow_media", 0, hipLaunchKernelGGL((k_shadow_media<STACK, SPEC, ADJ, true>), dim3(mgrid), dim3(TBLOCK), 0, st, S, P));
            else if (MEDIA) TIMED(ctx, "shadow_media", 0, hipLaunchKernelGGL((k_shadow_media<STACK, SPEC, ADJ>), dim3(mgrid), dim3(TBLOCK), 0, st, S, P));
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
    TIMED(ctx, "accumulate", npix, hipLaunchKernelGGL((k_accumulate<SPEC>), dim3((npix + SBLOCK - 1) / SBLOCK), dim3(SBLOCK), 0, st, S, P, (float4*)r->film.p, npix, frames, sample0, pg.wavelengths, (float4*)r->var_film.p, (float4*)r->aov_film[0].p, (float4*)r->aov_film[1].p,
                                                        pg.accumulation == PT_ACC_ONESHOT ? 1 : 0));
}
