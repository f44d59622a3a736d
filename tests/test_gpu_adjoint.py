"""The path-replay adjoint (pt_forward / pt_backward) on the HIP path, through the C ABI:
- against the oracle's FP64 replay of the same fp32 paths (gradient error well under 1 %);
- bitwise reproducibility of the deterministic reduction, and atomic == deterministic;
- the torch.autograd wrapper against central finite differences (albedos, RR off);
- Beer-Lambert d/dsigma_a = -exp(-sigma_a) through the F-SLAB ink;
- a short inverse render that recovers parameters from a target image."""
import sys

import numpy as np
import pytest

import oracle_lib
from pathtracerglsl_amd import abi, compiler, inverse, scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup(engine, theta, strategy, w, h, spp, mode=abi.REDUCE_DETERMINISTIC):
    comp = compiler.compile_scene(scenes.inverse_media_scene(theta), strategy, w, h)
    med, alb = scenes.inverse_adjoint_materials(comp)
    return comp, inverse.InverseRender(engine, comp, med, alb, spp, mode)


@pytest.mark.parametrize("depth", [12, 32])
def test_backward_matches_fp64_replay(engine, oracle, depth):
    """At the differentiable media workload's bounce depth 32 (and 12), on a small film."""
    _, theta = scenes.inverse_media_params()
    comp, inv = _setup(engine, theta, compiler.RenderStrategy("mis", depth, 3), 32, 32, 16)
    dL = np.random.default_rng(5).uniform(-1, 1, size=(32, 32, 4)).astype(np.float32)
    _say(f"[gpu] backward 32x32x16 depth {depth} inverse-media vs FP64 replay")
    g = inv.gradient(theta, dL, 7)
    comp.reset_salt = 7
    ors = oracle_lib.OracleScene(oracle, comp.scene)
    ref = ors.adjoint(comp.program, comp.params(7), 16, inv.adj, dL)
    err = inverse.relative_error(g, ref)
    _say(f"[gpu] adjoint relative error vs FP64 replay {err:.3e}")
    assert np.abs(ref).max() > 0.0
    assert err < 1e-3


def test_deterministic_reduction_is_reproducible(engine):
    _, theta = scenes.inverse_media_params()
    comp, inv = _setup(engine, theta, compiler.RenderStrategy("mis", 12, 3), 24, 24, 8)
    dL = np.random.default_rng(6).uniform(-1, 1, size=(24, 24, 4)).astype(np.float32)
    a = inv.gradient(theta, dL, 3)
    b = inv.gradient(theta, dL, 3)
    assert np.array_equal(a, b)
    inv.adj.reduce_mode = abi.REDUCE_ATOMIC
    c = inv.gradient(theta, dL, 3)
    assert inverse.relative_error(c, a) < 1e-5


def test_autograd_wrapper_matches_central_differences(engine):
    torch = pytest.importorskip("torch")
    from pathtracerglsl_amd import autograd
    _, theta0 = scenes.inverse_media_params()
    comp, inv = _setup(engine, theta0, compiler.RenderStrategy("mis", 12, -1), 24, 24, 32)
    w = torch.from_numpy(np.random.default_rng(8).uniform(-1, 1, size=(24, 24, 3)).astype(np.float32))
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    loss = (autograd.render(theta, inv, 5) * w).sum()
    loss.backward()
    g = theta.grad.numpy()
    for k in (7, 10, 16, 21, 22):
        h = 5e-3      # the film is fp32: a smaller step resolves its rounding, not the derivative
        tp, tm = theta0.copy(), theta0.copy()
        tp[k] += h
        tm[k] -= h
        with torch.no_grad():
            fd = float(((autograd.render(torch.tensor(tp), inv, 5) * w).sum()
                        - (autograd.render(torch.tensor(tm), inv, 5) * w).sum()) / (2 * h))
        _say(f"[gpu] d/dtheta[{k}] adjoint {g[k]:.6f} fd {fd:.6f}")
        assert abs(g[k] - fd) <= 1e-2 * max(abs(fd), 1e-2 * np.abs(g[7:]).max())


def test_beer_lambert_gradient_on_device(engine):
    comp = compiler.compile_scene(scenes.slab_witness(), compiler.RenderStrategy("none", 4, -1), 24, 24)
    ink = comp.material_names.index("ink")
    engine.load_renderer(comp)
    dL = np.zeros((24, 24, 4), np.float32)
    dL[8:16, 8:16, :3] = 1.0
    g = engine.backward(2048, abi.PtAdjointParams.make(ink, []), dL)
    np.testing.assert_allclose(g[3:6] / 64.0, -np.exp(-np.array([1.0, 2.0, 4.0])), rtol=0.03)


def test_inverse_render_recovers_parameters(engine):
    target, init = scenes.inverse_media_params()
    comp, inv = _setup(engine, target, compiler.RenderStrategy("mis", 12, 3), 64, 64, 1024)
    tgt_img = inv.render(target, 99999)
    inv.spp = 32
    theta, hist = inv.fit(init, tgt_img, steps=60, lr=0.02, log=_say)
    e0, e1 = inverse.param_error(init, target), inverse.param_error(theta, target)
    _say(f"[gpu] inverse: mean relative parameter error {e0:.3f} -> {e1:.3f} in 60 Adam steps")
    assert e1 < 0.7 * e0


def test_backward_matches_central_differences_all_parameters(engine, oracle):
    """pt_backward at the differentiable media workload's depth 32 against central differences
    (h = 1e-3) of the L2 loss for all 23 parameters, on the oracle's frozen-proposal estimator
    (test_adjoint_fd.py: its paths are the ones pt_backward replays): within 1 % per parameter
    (1e-4 of the largest gradient for the smallest ones)."""
    import test_adjoint_fd as F
    theta, comp, med, alb, dL, fd = F.fd_problem(oracle)
    _say("[gpu] backward vs central differences, 23 parameters, depth 32")
    engine.load_renderer(comp)
    g = engine.backward(F.SPP, abi.PtAdjointParams.make(med, alb), dL, comp.params(F.SALT)).astype(np.float64)
    err = F.report("[gpu] pt_backward vs FD", g, fd)
    assert err.max() < 0.01, err
