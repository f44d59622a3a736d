"""The oneshot accumulator (components/accumulator/oneshot/oneshot.glsl): the picture is the
latest sample alone, the sample count still advances, and it keeps no statistic.

On the oracle: after n samples the film equals the last sample's radiance exactly (RGB and
spectral, nee and mis), and the first-hit guides follow the same rule; an average program's
film is the running mean of the same samples, so the two differ. The reference's Validator
refuses oneshot with variance (validator.test.ts: "'oneshot' does not support variance"), and so
do the compiler and the ProgramDescription lowering; the plan record round-trips."""
import json

import numpy as np
import pytest

import oracle_lib
from pathtracerglsl_amd import _lib, abi, compiler, scenes


@pytest.mark.parametrize("est,color", [("nee", "rgb"), ("mis", "spectral")])
def test_oneshot_film_is_the_last_sample(oracle, est, color):
    st = compiler.RenderStrategy(est, 8, 3, color=color, accumulation="oneshot")
    comp = compiler.compile_scene(scenes.cornell_box(), st, 24, 20)
    assert comp.program.accumulation == abi.ACC_ONESHOT
    ors = oracle_lib.OracleScene(oracle, comp.scene)
    film, samples, _ = ors.render(comp.program, comp.params(), 5, keep_samples=True)
    assert np.array_equal(film[..., :3], samples[-1])
    assert film[..., :3].mean() > 0.05
    avg = compiler.compile_scene(scenes.cornell_box(), compiler.RenderStrategy(est, 8, 3, color=color), 24, 20)
    mean, _, _ = oracle_lib.OracleScene(oracle, avg.scene).render(avg.program, avg.params(), 5)
    assert not np.array_equal(mean[..., :3], film[..., :3])
    np.testing.assert_allclose(mean[..., :3], samples.mean(axis=0), rtol=1e-5, atol=1e-6)


def test_oneshot_guides_are_the_last_sample(oracle):
    st = compiler.RenderStrategy("nee", 6, 3, aovs=("albedo", "normal"), accumulation="oneshot")
    comp = compiler.compile_scene(scenes.cornell_box(), st, 16, 16)
    ors = oracle_lib.OracleScene(oracle, comp.scene)
    f3, a3, n3, _ = ors.render_aovs(comp.program, comp.params(), 3)
    f1, a1, n1, _ = ors.render_aovs(comp.program, comp.params(), 1, sample_count=2)   # sample 2 alone
    assert np.array_equal(f3, f1) and np.array_equal(a3, a1) and np.array_equal(n3, n1)


def test_oneshot_refuses_variance():
    with pytest.raises(compiler.CompileError, match="does not support variance"):
        compiler.compile_scene(scenes.cornell_box(), compiler.RenderStrategy("nee", 4, 3, variance=True, accumulation="oneshot"), 8, 8)
    with pytest.raises(compiler.CompileError, match="not yet supported"):
        compiler.compile_scene(scenes.cornell_box(), compiler.RenderStrategy("nee", 4, 3, accumulation="exponential"), 8, 8)
    d = compiler.describe(scenes.cornell_box(), compiler.RenderStrategy("nee", 4, 3, accumulation="oneshot"))
    assert d["estimator"]["accumulation"] == {"type": "oneshot"}
    assert compiler.program_from_description(d, 8, 8, 1).accumulation == abi.ACC_ONESHOT
    d["estimator"]["accumulation"]["variance"] = True
    with pytest.raises(_lib.PtError, match="oneshot"):
        compiler.program_from_description(d, 8, 8, 1)
