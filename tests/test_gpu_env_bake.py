"""The procedural sky's importance table baked on the GPU (k_env_luma / k_env_blur / k_env_mean /
k_env_compensate / k_env_rows / k_env_marginal, the path pt_scene_upload takes) against the
host builder and the CPU oracle: conditional CDFs, marginal CDF and total mass bit for bit, at
the default 512 x 256 table and at other sizes, with and without MIS compensation. The timing
of both is printed (HIP events around the bake's kernels; wall clock around the host build)."""
import ctypes as C
import sys
import time

import numpy as np
import pytest

import oracle_lib
from pathtracerglsl_amd import _lib, abi, compiler, scenes

pytestmark = pytest.mark.gpu


def _env(w, h, compensated):
    return abi.PtEnv(abi.ENV_SKY, w, h, int(compensated), 1.0, 0.0, 0.5, 1)


def _host(env):
    lib = _lib.load()
    lib.pt_debug_env_table.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]
    w = env.table_h if env.chart == abi.CHART_OCTAHEDRAL else env.table_w
    cond = np.zeros((env.table_h, w), np.float32)
    marg = np.zeros(env.table_h, np.float32)
    tot = C.c_double()
    t0 = time.perf_counter()
    lib.pt_debug_env_table(C.byref(env), cond.ctypes.data, marg.ctypes.data, C.byref(tot))
    return cond, marg, tot.value, (time.perf_counter() - t0) * 1e3


def _device(env):
    lib = _lib.load()
    lib.pt_debug_env_bake.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double),
                                      C.POINTER(C.c_float)]
    w = env.table_h if env.chart == abi.CHART_OCTAHEDRAL else env.table_w
    cond = np.zeros((env.table_h, w), np.float32)
    marg = np.zeros(env.table_h, np.float32)
    tot, ms = C.c_double(), C.c_float()
    _lib.check(lib.pt_debug_env_bake(0, C.byref(env), cond.ctypes.data, marg.ctypes.data, C.byref(tot), C.byref(ms)))
    return cond, marg, tot.value, ms.value


@pytest.mark.parametrize("w,h,comp", [(512, 256, 0), (512, 256, 1), (256, 128, 0), (1024, 512, 1), (2048, 1024, 0),
                                      (4096, 2048, 0), (300, 150, 1), (4096, 2048, 1)])
def test_device_bake_equals_host_builder(engine, w, h, comp):
    env = _env(w, h, comp)
    print(f"[gpu] env bake {w}x{h} compensated={comp}", flush=True)
    _device(env)                                          # first call loads the kernels
    gc, gm, gt, gms = _device(env)
    hc, hm, ht, hms = _host(env)
    print(f"[gpu] env bake {w}x{h} comp={comp}: device {gms:.3f} ms, host {hms:.1f} ms", flush=True)
    sys.stdout.flush()
    assert np.array_equal(gc.view(np.uint32), hc.view(np.uint32))
    assert np.array_equal(gm.view(np.uint32), hm.view(np.uint32))
    # rows the compensation empties keep an all-zero CDF; every other row ends at 1
    assert gt == ht and gm[-1] == np.float32(1.0) and np.all((gc[:, -1] == 1.0) | ((gc[:, -1] == 0.0) & (comp == 1)))


@pytest.mark.parametrize("w,h,comp", [(512, 256, 1), (256, 128, 0)])
def test_device_bake_equals_oracle(engine, oracle, w, h, comp):
    sc = scenes.cornell_box()
    sc["environment"] = {"type": "procedural", "intensity": 1.0, "compensated": bool(comp)}
    cs = compiler.compile_scene(sc, compiler.RenderStrategy("mis", 4, -1), 16, 16)
    cs.scene.env.table_w, cs.scene.env.table_h = w, h
    ors = oracle_lib.OracleScene(oracle, cs.scene)
    oc, om = ors.env_table(w, h)
    gc, gm, gt, _ = _device(cs.scene.env)
    assert np.array_equal(gc.view(np.uint32), oc.view(np.uint32)) and np.array_equal(gm.view(np.uint32), om.view(np.uint32))
    assert gt == oracle.or_env_total(ors.h)


@pytest.mark.parametrize("n,comp", [(256, 0), (256, 1), (512, 0), (1024, 1)])
def test_device_bake_octahedral_equals_host_and_oracle(engine, oracle, n, comp):
    """The N x N equal-area octahedral table (N = tableSize[1]): device bake == host builder bit
    for bit, and == the oracle's table at the default N = 256."""
    env = abi.PtEnv(abi.ENV_SKY, 512, n, comp, 1.0, 0.0, 0.5, 1)
    env.chart = abi.CHART_OCTAHEDRAL
    print(f"[gpu] octahedral env bake {n}x{n} compensated={comp}", flush=True)
    gc, gm, gt, gms = _device(env)
    hc, hm, ht, hms = _host(env)
    assert gc.shape == (n, n)
    print(f"[gpu] octahedral env bake {n}x{n} comp={comp}: device {gms:.3f} ms, host {hms:.1f} ms", flush=True)
    assert np.array_equal(gc.view(np.uint32), hc.view(np.uint32))
    assert np.array_equal(gm.view(np.uint32), hm.view(np.uint32))
    assert gt == ht and gm[-1] == np.float32(1.0)
    if n == 256:
        sc = scenes.cornell_box()
        sc["environment"] = {"type": "procedural", "intensity": 1.0, "compensated": bool(comp), "chart": "octahedral"}
        cs = compiler.compile_scene(sc, compiler.RenderStrategy("mis", 4, -1), 16, 16)
        ors = oracle_lib.OracleScene(oracle, cs.scene)
        oc, om = ors.env_table(n, n)
        assert np.array_equal(gc.view(np.uint32), oc.view(np.uint32)) and np.array_equal(gm.view(np.uint32), om.view(np.uint32))


def test_program_chart_must_match_the_scene_table(engine):
    """A program that names an envSampler.chart (a lowered ProgramDescription does) loads only
    over a sky whose table is on that chart; the matching chart loads and renders."""
    sc = scenes.cornell_box()
    sc["environment"] = {"type": "procedural", "intensity": 1.0}
    cs = compiler.compile_scene(sc, compiler.RenderStrategy("mis", 4, -1), 16, 16)
    print("[gpu] chart mismatch at renderer load", flush=True)
    cs.program.env_chart = abi.CHART_OCTAHEDRAL + 1
    with pytest.raises(_lib.PtError) as e:
        engine.load_renderer(cs)
    assert e.value.status == 1 and "envSampler.chart" in str(e.value)
    cs.program.env_chart = abi.CHART_EQUIRECT + 1
    engine.load_renderer(cs)
    engine.clear_accumulation()
    engine.render_frames(1)
    assert np.isfinite(engine.read_export()).all()
