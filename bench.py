#!/usr/bin/env python3
"""Benchmark: Mpaths/s and Mrays/s of the transport loop on one MI355X.

Workload (the 1-GPU configuration the metric is quoted on): seeded 1M-triangle scene of 2000
noise-displaced detail-4 icospheres (60 % GGX, 20 % dielectric, 20 % Lambert) under the
procedural sky, pt-mis with env NEE, 1920x1080, 64 spp, depth 12, Russian roulette from
bounce 3, through the 8-wide compressed BVH. One step = one full 64-spp render of the frame
(pt_render_frames(64) + pt_finish); inputs (scene, BVH, env table) are resident in HBM
before the timed region. All 64 frames are in flight at once: one wave of 132.7 M paths
(about 37 GB of path state), so the late, short bounces launch 13 times per step rather than
208 times (each launch of a short bounce costs the latency of its longest traversal).

Adjoint leg (the differentiable media inverse render): the glass-medium Cornell box with 16 grey
Lambert patches at 512x512, 64 spp, depth 32, pt-mis; the adjoint image is the L2 loss gradient
between the perturbed-init render and the hidden-target render. Reported: pt_forward and
pt_backward wall times (their ratio is the adjoint/forward cost) and the gradient's relative
error against the FP64 CPU replay of the same paths.

Batch leg (the batched persistent wavefront): 64 seeded icosphere-cloud scenes with log-uniform
triangle count (1k-500k), resolution (128^2-1024^2) and spp (4-256), pt-mis depth 12, rendered
as one hipGraph-captured batch after the autotuner's sweep, against 64 sequential plain
launches (pt_render_frames per scene; uploads untimed); reported with the per-pixel identity
of all 64 films.

Configuration legs (same JSON line): the plain Cornell box (256x256, 16 spp, depth 8, pt-mis, RGB)
and the spectral glass-medium Cornell box (1024x1024, 256 spp, depth 32, pt-mis, 4 hero
wavelengths), each with Mpaths/s, Mrays/s and a bounded CPU sample on the oracle.

Also reported: the roofline of the dominant kernel (the extend / closest-hit traversal),
whose algorithmic bytes per query are 32 (ray) + 16 (hit) + 80 per wide node visited + 36 per
triangle tested, counted by an instrumented, untimed replay of the identical render; and a
CPU baseline: the CPU restatement of the reference walk on a bounded sample of the workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpaths/s & Mrays/s on 1 GPU; adjoint/forward cost ratio at 1% grad err"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(comp, spp, sample=None):
    """CPU restatement of the reference walk (the oracle), all host threads, bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    lib = oracle_lib.load()
    ors = oracle_lib.OracleScene(lib, comp.scene)
    t0 = time.perf_counter()
    _, _, counts = ors.render(comp.program, comp.params(), spp)
    dt = time.perf_counter() - t0
    paths = comp.width * comp.height * spp
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    flags_path = os.path.join(ROOT, "oracle", "_build", "liboracle.flags")
    flags = open(flags_path).read().strip() if os.path.exists(flags_path) else None
    return {"value": round(paths / dt / 1e6, 4), "unit": "Mpaths/s", "cores": cores, "kind": "port",
            "sample": sample or f"{comp.width}x{comp.height} frame, {spp} spp of the 64 (same scene, streams, depth)",
            "seconds": round(dt, 2), "mrays_per_s": round(float(counts.sum()) / dt / 1e6, 4), "build": flags}


def config_leg(Engine, comp, spp, reps=3, cpu=None):
    """Throughput of one more configuration at its own size: Mpaths/s and Mrays/s of full
    renders with the scene resident (best of reps after one warm render), and a bounded CPU
    sample of the same scene on the oracle."""
    eng = Engine(0)
    eng.load_renderer(comp)
    eng.clear_accumulation()
    eng.render_frames(spp)
    eng.finish()
    best = None
    for _ in range(reps):
        eng.clear_accumulation()
        eng.finish()
        t0 = time.perf_counter()
        eng.render_frames(spp)
        eng.finish()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    c, a = eng.ray_counts()
    film = eng.read_export()
    eng.close()
    paths = comp.width * comp.height * spp
    out = {"paths": paths, "ms": round(best * 1e3, 3), "mpaths_per_s": round(paths / best / 1e6, 3),
           "mrays_per_s": round((c + a) / best / 1e6, 3), "closest_rays": int(c), "any_rays": int(a),
           "film_mean": float(film[..., :3].mean()), "finite": bool(__import__("numpy").isfinite(film).all())}
    if cpu is not None:
        ccomp, cspp, label = cpu
        out["cpu_baseline"] = cpu_baseline(ccomp, cspp, label)
    return out


def adjoint_leg(Engine, compiler, scenes, abi, fp64=True, res=512, spp=64):
    """Forward / backward cost of the path-replay adjoint and its error vs the FP64 replay."""
    import numpy as np
    target, init = scenes.inverse_media_params()
    strategy = compiler.RenderStrategy("mis", 32, 3, 0.95, 64)
    eng = Engine(0)
    ct = compiler.compile_scene(scenes.inverse_media_scene(target), strategy, res, res)
    eng.load_renderer(ct)
    eng.forward(spp, ct.params(101))
    tgt = eng.read_export()
    comp = compiler.compile_scene(scenes.inverse_media_scene(init), strategy, res, res)
    med, alb = scenes.inverse_adjoint_materials(comp)
    adj = abi.PtAdjointParams.make(med, alb, abi.REDUCE_DETERMINISTIC)
    eng.load_renderer(comp)
    eng.forward(spp, comp.params(102))
    img = eng.read_export()
    dL = np.zeros(img.shape, np.float32)
    dL[..., :3] = 2.0 * (img[..., :3] - tgt[..., :3]) / (res * res)
    eng.backward(spp, adj, dL, comp.params(103))          # warm
    tf, tb = [], []
    for _ in range(3):
        eng.finish()
        t0 = time.perf_counter()
        eng.forward(spp, comp.params(103))
        eng.finish()
        tf.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        g = eng.backward(spp, adj, dL, comp.params(103))
        tb.append(time.perf_counter() - t0)
    out = {"workload": f"inverse-media cornell 16 albedos + medium {res}x{res} {spp}spp depth32 pt-mis",
           "paths": res * res * spp, "forward_ms": round(1e3 * min(tf), 2), "backward_ms": round(1e3 * min(tb), 2),
           "adjoint_forward_ratio": round(min(tb) / min(tf), 3), "n_params": abi.ADJ_N_PARAMS,
           "reduce_mode": "deterministic"}
    if fp64:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        ors = oracle_lib.OracleScene(oracle_lib.load(), comp.scene)
        t0 = time.perf_counter()
        ref = ors.adjoint(comp.program, comp.params(103), spp, adj, dL)
        out["fp64_replay_seconds"] = round(time.perf_counter() - t0, 2)
        out["grad_rel_err_vs_fp64"] = float(np.linalg.norm(g.astype(np.float64) - ref) / max(np.linalg.norm(ref), 1e-30))
    eng.close()
    return out


def batch_leg(Engine, compiler, scenes, abi):
    """64-scene batch vs 64 sequential plain launches, identical films required."""
    import numpy as np
    from pathtracerglsl_amd.batch import BatchRenderer
    items = scenes.batch_workload()
    strategy = compiler.RenderStrategy("mis", 12, 3, 0.95, 64)    # sequential launches: all frames in flight
    comps = [compiler.compile_scene(sc, strategy, w, h) for sc, w, h, _ in items]
    spp = [n for _, _, _, n in items]
    paths = sum(c.width * c.height * n for c, n in zip(comps, spp))
    eng = Engine(0)
    seq, films = 0.0, []
    for c, n in zip(comps, spp):
        rid = eng.load_renderer(c)
        eng.render_frames(min(n, c.program.frames_in_flight))   # path state allocated outside the clock
        eng.finish()
        t0 = time.perf_counter()
        eng.clear_accumulation()
        eng.render_frames(n)
        eng.finish()
        seq += time.perf_counter() - t0
        films.append(eng.read_export())
        eng.unload_renderer(rid)           # its path state goes before the next scene's
    log(f"[bench] batch: sequential {seq:.3f}s for {paths / 1e6:.1f} Mpaths")
    # the same sequential launches with 4 frames per wave (the library's earlier default)
    comps4 = [compiler.compile_scene(sc, compiler.RenderStrategy("mis", 12, 3, 0.95, 4), w, h) for sc, w, h, _ in items]
    seq4 = 0.0
    for c, n in zip(comps4, spp):
        rid = eng.load_renderer(c)
        eng.render_frames(min(n, 4))
        eng.finish()
        t0 = time.perf_counter()
        eng.clear_accumulation()
        eng.render_frames(n)
        eng.finish()
        seq4 += time.perf_counter() - t0
        eng.unload_renderer(rid)
    br = BatchRenderer(eng, comps)
    best, sweep = br.autotune(spp, probe_div=8)
    br.render(spp)                       # capture + first replay
    eng.finish()
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        br.render(spp)
        eng.finish()
        times.append(time.perf_counter() - t0)
    bt = min(times)
    bfilms = [br.read_export(s) for s in range(len(comps))]
    identical = all(np.array_equal(bfilms[s], films[s], equal_nan=True) for s in range(len(comps)))
    nonfinite = int(sum(int((~np.isfinite(f)).sum()) for f in bfilms))
    sweep.sort(key=lambda x: x[1])
    out = {"workload": "64 icosphere-cloud scenes, log-uniform 1k-500k tris, 128^2-1024^2, 4-256 spp, pt-mis depth12",
           "triangles": int(sum(c.scene.n_tris for c in comps)), "paths": int(paths),
           "sequential_s": round(seq, 4), "batched_s": round(bt, 4), "speedup": round(seq / bt, 3),
           "sequential_4_frames_per_wave_s": round(seq4, 4), "speedup_vs_4_frames_per_wave": round(seq4 / bt, 3),
           "batched_mpaths_per_s": round(paths / bt / 1e6, 2), "sequential_mpaths_per_s": round(paths / seq / 1e6, 2),
           "identical_films": bool(identical), "nonfinite_pixels": nonfinite, "autotuned": best.as_dict(),
           "sweep_top": [{"config": c, "ms": round(m, 2)} for c, m in sweep[:5]], "sweep_size": len(sweep)}
    br.close()
    eng.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--frames-in-flight", type=int, default=64)
    ap.add_argument("--cpu-spp", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the Cornell and spectral-media legs")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run a plain device copy of this size first (counter calibration runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for --gpus {args.gpus}); a multi-GPU run would "
              "shard the frame by image tiles, one process per GPU", file=sys.stderr)
        sys.exit(2)

    from pathtracerglsl_amd import compiler, scenes
    from pathtracerglsl_amd.engine import Engine

    log("[bench] building scene")
    sc = scenes.icosphere_cloud()
    comp = compiler.compile_scene(sc, compiler.RenderStrategy("mis", 12, 3, 0.95, args.frames_in_flight),
                                  args.width, args.height)
    eng = Engine(0)
    t0 = time.perf_counter()
    eng.load_renderer(comp)
    log(f"[bench] scene upload + wide BVH build {time.perf_counter() - t0:.1f}s {eng.scene_stats()}")

    if args.calib_copy_mib:
        import ctypes as C
        eng.lib.pt_calib_copy.argtypes = [C.c_void_p, C.c_size_t, C.c_int]
        from pathtracerglsl_amd import _lib
        _lib.check(eng.lib.pt_calib_copy(eng.ctx, args.calib_copy_mib << 20, 3))
        log(f"[bench] calibration copy {args.calib_copy_mib} MiB x3 done")

    for w in range(args.warmup):
        eng.clear_accumulation()
        eng.render_frames(args.spp)
        eng.finish()
        log(f"[bench] warmup {w} done")

    eng.set_timing(True)
    ray_closest = ray_any = 0
    eng.finish()
    t0 = time.perf_counter()
    for s in range(args.steps):
        eng.clear_accumulation()
        eng.render_frames(args.spp)
    eng.finish()
    elapsed = time.perf_counter() - t0
    timings = eng.timings()
    c, a = eng.ray_counts()            # counts of the last step
    ray_closest, ray_any = c * args.steps, a * args.steps
    log(f"[bench] timed {args.steps} steps in {elapsed:.3f}s; kernels {timings}")
    eng.set_timing(False)

    # instrumented replay (untimed): node / triangle visits of the extend kernel
    eng.set_traversal_counting(True)
    eng.clear_accumulation()
    eng.render_frames(args.spp)
    eng.finish()
    nodes, tris = eng.traversal_counts()
    rc_closest, _ = eng.ray_counts()
    eng.set_traversal_counting(False)

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "extend_traffic.json")
    if os.path.exists(tpath) and (args.width, args.height, args.spp) == (1920, 1080, 64):
        rec = json.load(open(tpath))
        if rec.get("frames_in_flight") == args.frames_in_flight:
            traffic = rec.get("hbm_bytes_per_launch")

    paths = args.width * args.height * args.spp * args.steps
    ms_per_step = elapsed * 1e3 / args.steps
    ext_ms, ext_launches, _ = timings.get("extend", (0.0, 0, 0))
    alg_bytes_per_render = 48.0 * rc_closest + 80.0 * nodes + 36.0 * tris
    ext_ms_per_render = ext_ms / args.steps
    achieved = alg_bytes_per_render / (ext_ms_per_render * 1e-3) / 1e9 if ext_ms_per_render > 0 else 0.0
    launches_per_render = ext_launches / args.steps if args.steps else 0
    result = {
        "metric": METRIC,
        "value": round(paths / elapsed / 1e6, 3),
        "unit": "Mpaths/s",
        "mrays_per_s": round((ray_closest + ray_any) / elapsed / 1e6, 3),
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded icosphere cloud, procedural sky)",
        "config": {"workload": "icosphere-cloud-1M-tri 1920x1080 64spp depth12 pt-mis sky",
                   "triangles": comp.scene.n_tris, "width": args.width, "height": args.height,
                   "spp": args.spp, "max_bounces": 12, "rr_start": 3,
                   "frames_in_flight": args.frames_in_flight, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_extend (closest-hit wide-BVH traversal)",
                     "alg_bytes_per_launch": round(alg_bytes_per_render / max(launches_per_render, 1)),
                     "avg_launch_ms": round(ext_ms_per_render / max(launches_per_render, 1), 4),
                     "nodes_per_query": round(nodes / max(rc_closest, 1), 2),
                     "tris_per_query": round(tris / max(rc_closest, 1), 2),
                     # the counted HBM bytes over the same launch time: the algorithmic bytes are
                     # mostly served by L2 and the Infinity Cache, so this is the DRAM share
                     "traffic_gbs": round(traffic / (ext_ms_per_render / max(launches_per_render, 1) * 1e-3) / 1e9, 1)
                     if traffic and ext_ms_per_render > 0 else None,
                     "traffic_frac": round(traffic / (ext_ms_per_render / max(launches_per_render, 1) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                     if traffic and ext_ms_per_render > 0 else None},
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in timings.items()},
        "shade_ms_per_step": round(sum(v[0] for k, v in timings.items() if k.startswith("shade")) / args.steps, 3),
    }
    if not args.no_cpu:
        log("[bench] CPU baseline")
        result["cpu_baseline"] = cpu_baseline(comp, args.cpu_spp)
    eng.close()
    if not args.no_configs:
        log("[bench] configuration legs")
        st0 = compiler.RenderStrategy("mis", 8, 3, 0.95, 16)
        c0 = compiler.compile_scene(scenes.cornell_box(), st0, 256, 256)
        result["cornell"] = dict(config_leg(Engine, c0, 16, cpu=None if args.no_cpu else (c0, 16, "the whole workload")),
                                 workload="cornell 256x256 16spp depth8 pt-mis rgb")
        log(f"[bench] cornell {result['cornell']}")
        st2 = compiler.RenderStrategy("mis", 32, 3, 0.95, 64, color="spectral")
        c2 = compiler.compile_scene(scenes.cornell_spectral_medium(), st2, 1024, 1024)
        c2s = compiler.compile_scene(scenes.cornell_spectral_medium(), st2, 128, 128)
        result["spectral_media"] = dict(config_leg(Engine, c2, 256, cpu=None if args.no_cpu else (c2s, 16, "128x128 16spp of the same scene")),
                                        workload="spectral glass-medium cornell 1024x1024 256spp depth32 pt-mis")
        log(f"[bench] spectral media {result['spectral_media']}")
    if not args.no_adjoint:
        log("[bench] adjoint leg")
        from pathtracerglsl_amd import abi
        result["adjoint"] = adjoint_leg(Engine, compiler, scenes, abi, fp64=not args.no_cpu)
        log(f"[bench] adjoint {result['adjoint']}")
    if not args.no_batch:
        log("[bench] batch leg")
        from pathtracerglsl_amd import abi
        result["batch"] = batch_leg(Engine, compiler, scenes, abi)
        log(f"[bench] batch {result['batch']}")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
