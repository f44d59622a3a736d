"""Live parameters on the host (Engine.setParameter, src/engine/Engine.ts:258): the dotted paths
map onto the authored scene without touching geometry, and the derived values follow them as
the reference recomputes its derived uniforms on the CPU (features/lighting.ts:166-183 the
selection arrays, :488-531 the env selection probability and its dependencies env.intensity,
env.totalWeight, env.size and the driven emission)."""
import math

import numpy as np
import pytest

from pathtracerglsl_amd import abi, compiler, scenes
from pathtracerglsl_amd.engine import with_parameter


def sky_lit():
    sc = scenes.cornell_multi_light(point=True)
    sc["environment"] = {"type": "procedural", "intensity": 1.0}
    return sc


def test_paths_touch_only_their_branch():
    sc = sky_lit()
    out = with_parameter(sc, "lights.0.radiance", [1.0, 2.0, 3.0])
    assert out["lights"][0]["radiance"] == [1.0, 2.0, 3.0] and sc["lights"][0]["radiance"] != [1.0, 2.0, 3.0]
    assert out["objects"] is sc["objects"] and out["materials"] is sc["materials"]
    out = with_parameter(sc, "env.selectProb", 0.3)
    assert out["environment"]["select_prob"] == 0.3 and "select_prob" not in sc["environment"]
    out = with_parameter(sc, "materials.white.albedo", [0.1, 0.2, 0.3])
    assert out["materials"]["white"]["albedo"] == [0.1, 0.2, 0.3] and out["materials"]["red"] is sc["materials"]["red"]
    out = with_parameter(sc, "camera.fov", 0.5)
    assert out["camera"]["fov"] == 0.5 and out["camera"]["position"] == sc["camera"]["position"]
    for bad in ["lights.0.corner", "lights.3.radiance", "lights.7.radiance", "materials.white.model",
                "materials.nope.albedo", "env.chart", "objects.0.corner", "camera"]:
        with pytest.raises(KeyError):
            with_parameter(sc, bad, 1.0)


def test_emission_change_recomputes_the_selection_table_and_env_probability():
    st = compiler.RenderStrategy("mis", 8, 3)
    a = compiler.compile_scene(sky_lit(), st)
    b = compiler.compile_scene(with_parameter(sky_lit(), "lights.0.radiance", [170.0, 120.0, 40.0]), st)
    # ten times the ceiling light's power: its selection pdf rises, the others' fall, the
    # table still closes at 1, and the sky's share of the NEE draws falls
    assert b.scene.lights[0]["select_pdf"] > a.scene.lights[0]["select_pdf"]
    assert all(b.scene.lights[i]["select_pdf"] < a.scene.lights[i]["select_pdf"] for i in (1, 2, 3))
    assert abs(float(b.scene.lights[-1]["cdf"]) - 1.0) < 1e-6
    assert b.scene.env.select_prob < a.scene.env.select_prob
    # power partition by hand: quads pi A mean(L), the point light 4 pi mean(I)
    pw = [math.pi * 0.25 * (170 + 120 + 40) / 3, math.pi * 0.09 * 7.0, math.pi * 0.08 * (10 / 3), 4 * math.pi * 0.5]
    np.testing.assert_allclose(b.scene.lights["select_pdf"], np.array(pw) / sum(pw), rtol=1e-6)
    # the emitter's material record follows the light
    em = b.scene.materials[int(b.scene.objects[int(b.scene.lights[0]["region"])]["material"])]
    assert list(em["emission"][:3]) == [170.0, 120.0, 40.0]
    # uniform selection ignores power
    u = compiler.compile_scene(with_parameter(sky_lit(), "lights.0.radiance", [170.0, 120.0, 40.0]),
                               compiler.RenderStrategy("mis", 8, 3, light_selection="uniform"))
    assert np.all(u.scene.lights["select_pdf"] == np.float32(0.25))


def test_env_dependencies_follow_intensity_total_weight_and_size():
    st = compiler.RenderStrategy("mis", 8, 3)
    a = compiler.compile_scene(sky_lit(), st)
    twice = compiler.compile_scene(with_parameter(sky_lit(), "env.intensity", 2.0), st)
    assert twice.scene.env.select_prob > a.scene.env.select_prob
    # a registered table's mass and size replace the baked table's in mean(L) = tw pi / (2 W H)
    src = with_parameter(sky_lit(), "env.totalWeight", 1000.0)
    src["environment"]["table_size"] = (100, 50)
    c = compiler.compile_scene(src, st)
    assert (c.scene.env.table_w, c.scene.env.table_h) == (100, 50)
    powers = [float(p) for p in (math.pi * 0.25 * 11.0, math.pi * 0.09 * 7.0, math.pi * 0.08 * (10 / 3), 4 * math.pi * 0.5)]
    r = compiler.scene_radius_estimate(src)
    phi_env = 4 * math.pi ** 2 * r * r * (1000.0 * math.pi / (2 * 100 * 50))
    expect = min(0.99, max(0.01, phi_env / (phi_env + sum(powers))))
    assert c.scene.env.select_prob == np.float32(expect)
