"""The reference's ProgramDescription (the Planner's link map, src/compiler/plan/types.ts:39-249)
lowered by the library (pt_program_from_description), over the 505 frozen link maps of
tests/compiler/__snapshots__/program-description.snapshot.test.ts.snap (tests/golden/
program_descriptions.json; fields the lowering reads, plus five complete records).

Pins: every accepted record lowers to the program its fields name (estimator, bounce budget,
Russian roulette, colour model, wavelength sampler, medium options, variance); every rejected
record is rejected with PT_E_UNSUPPORTED naming a field whose value selects a component outside
this backend; complete records (with the fields the lowering ignores) lower like their pruned
form; and this backend's own plans round-trip: the lowering of compiler.describe(scene,
strategy) is compile_scene's program."""
import json
import os

import pytest

from pathtracerglsl_amd import _lib, abi, compiler, scenes

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "program_descriptions.json")))
SUPPORTED_MODELS = {"lambert", "dielectric", "ggx", "none", "mirror", "rough_dielectric", "checker"}


def _lower(d, w=64, h=48, fif=1):
    return compiler.program_from_description(d, w, h, fif)


def _expected_estimator(d):
    lighting = d["estimator"].get("lighting")
    return abi.EST_PT if lighting is None else {"nee": abi.EST_NEE, "mis": abi.EST_MIS}[lighting["method"]]


# The number of the 505 frozen link maps this backend lowers: a lowering that starts rejecting
# a case (or accepting one it cannot render) changes it.
EXPECTED_ACCEPTED = 409


def test_library_lowers_or_names_every_snapshot_case():
    accepted, rejected = [], 0
    for name, d in FIX["cases"].items():
        try:
            p = _lower(d)
        except _lib.PtError as e:
            assert e.status == 5, (name, str(e))                      # PT_E_UNSUPPORTED
            msg = str(e)
            assert "is not supported by this backend" in msg, (name, msg)
            rejected += 1
            continue
        accepted.append(name)
        m, e = d["measurement"], d["estimator"]
        assert p.max_bounces == m["maxBounces"], name
        assert p.max_null_crossings == m.get("maxNullCrossings", 32), name
        assert p.color == (abi.COLOR_SPECTRAL if m["color"] == "spectral" else abi.COLOR_RGB), name
        assert p.estimator == _expected_estimator(d), name
        sel = (d["estimator"].get("lighting") or {}).get("selection", "power")
        assert p.light_selection == {"uniform": abi.SEL_UNIFORM, "bvh": abi.SEL_BVH}.get(sel, abi.SEL_POWER), name
        rr = e.get("russianRoulette")
        assert p.rr_start == (-1 if rr is None else rr["startDepth"]), name
        assert abs(p.rr_max_survival - ((rr or {}).get("maxSurvival") or 0.95)) < 1e-7, name
        assert p.wavelengths == (abi.WL_UNIFORM if e.get("wavelengths") == "uniform" else abi.WL_VISIBLE), name
        assert p.medium_light_sampling == (abi.MLS_EQUIANGULAR if e.get("mediumLightSampling") == "equiangular"
                                           else abi.MLS_VERTEX), name
        assert p.volume_sampling == (abi.VS_DELTA_TRACKING if e.get("volumeSampling") == "delta-tracking"
                                     else abi.VS_ANALYTIC), name
        assert p.variance == int(bool(e.get("accumulation", {}).get("variance"))), name
        assert p.accumulation == (abi.ACC_ONESHOT if e.get("accumulation", {}).get("type") == "oneshot"
                                  else abi.ACC_AVERAGE), name
        assert set(d["materials"]["models"]) <= SUPPORTED_MODELS, name
        chart = (e.get("envSampler") or {}).get("chart")
        assert p.env_chart == {None: 0, "equirect": abi.CHART_EQUIRECT + 1, "octahedral": abi.CHART_OCTAHEDRAL + 1}[chart], name
        assert (p.width, p.height, p.frames_in_flight, p.abi_version) == (64, 48, 1, abi.ABI_VERSION)
    print(f"lowered {len(accepted)} of {len(accepted) + rejected}")
    assert len(accepted) + rejected == 505
    assert len(accepted) == EXPECTED_ACCEPTED
    # frozen octahedral-chart cases on procedural skies and image environments
    assert {"spectral-image-constant-ref + oct", "sky-missed-cap + spectral-octahedral", "spectral-comp-empty + octahedral",
            "spectral-proc-sky + oct"} <= set(accepted)
    assert "two-light + uniform" in accepted                  # estimator.lighting.selection 'uniform'
    assert "spectral-rayleigh-furnace + visible" in accepted  # media.models ['rayleigh']
    assert {"spectral-gaussian-plane + spectral-mis", "spectral-gaussian-plane + spectral-none"} <= set(accepted)
    assert "cornell-oneshot + oneshot" in accepted                 # estimator.accumulation.type 'oneshot'
    # procedural sky bodies other than proc-sky's, and proc-sky under the spectral colour model
    assert {"proc-sky-rotated + pt-mis", "spectral-proc-sky + mis"} <= set(accepted)
    assert {"two-light + bvh", "hundred-spheres + mis-bvh", "spectral-packed + nee-bvh"} <= set(accepted)   # the light tree


def test_chart_and_albedo_flag_lowering():
    """envSampler.chart lowers to env_chart (0 = unnamed, 1 + PT_CHART_*); materials.albedo (the
    albedo guide's dispatch flag) is accepted and leaves the program as measurement.aovs sets it."""
    base = FIX["cases"]["proc-sky + pt-mis"]
    assert _lower(base).env_chart == abi.CHART_EQUIRECT + 1
    d = json.loads(json.dumps(base))
    del d["estimator"]["envSampler"]
    assert _lower(d).env_chart == 0
    for chart, want in (("equirect", abi.CHART_EQUIRECT + 1), ("octahedral", abi.CHART_OCTAHEDRAL + 1)):
        d = json.loads(json.dumps(base))
        d["estimator"]["envSampler"] = {"chart": chart, "compensation": False}
        assert _lower(d).env_chart == want
    d = json.loads(json.dumps(base))
    d["materials"]["albedo"] = True
    assert bytes(_lower(d)) == bytes(_lower(base))
    d["measurement"]["aovs"] = ["albedo", "normal"]
    assert _lower(d).aovs == abi.AOV_ALBEDO | abi.AOV_NORMAL


@pytest.mark.parametrize("name", sorted(FIX["full"]))
def test_complete_records_lower_like_their_pruned_form(name):
    full, pruned = FIX["full"][name], FIX["cases"][name]
    try:
        a = _lower(full)
    except _lib.PtError as e:
        with pytest.raises(_lib.PtError) as e2:
            _lower(pruned)
        assert str(e2.value) == str(e)
        return
    b = _lower(pruned)
    assert bytes(a) == bytes(b)


def test_rejection_names_the_field():
    d = json.loads(json.dumps(FIX["cases"]["analytic-glass + pathtracer"]))
    d["measurement"]["camera"]["type"] = "panoramic"
    with pytest.raises(_lib.PtError, match="measurement.camera.type 'panoramic'"):
        _lower(d)
    d = json.loads(json.dumps(FIX["cases"]["analytic-glass + pathtracer"]))
    d["materials"]["models"].append("toon")
    with pytest.raises(_lib.PtError, match="material model 'toon'"):
        _lower(d)
    with pytest.raises(_lib.PtError, match="malformed JSON"):
        _lower('{"measurement": {')


@pytest.mark.parametrize("make,strategy", [
    (scenes.cornell_box, compiler.RenderStrategy("mis", 8, 3)),
    (scenes.two_light, compiler.RenderStrategy("nee", 6, -1, light_selection="uniform")),
    (scenes.cornell_box, compiler.RenderStrategy("none", 5, -1, variance=True)),
    (scenes.haze, compiler.RenderStrategy("nee", 24, 4, medium_light_sampling="equiangular")),
    (scenes.emit_sat, compiler.RenderStrategy("nee", 32, -1, volume_sampling="delta-tracking")),
    (scenes.cornell_spectral_medium, compiler.RenderStrategy("mis", 8, 3, color="spectral", wavelengths="uniform")),
    (scenes.icosphere_cloud, compiler.RenderStrategy("mis", 12, 3, 0.95, 4)),
    (lambda: scenes.thinlens(scenes.cornell_box(), 0.05, 3.0), compiler.RenderStrategy("nee", 6, 2)),
])
def test_own_plans_round_trip(make, strategy):
    sc = make()
    comp = compiler.compile_scene(sc, strategy, 40, 30)
    p = _lower(compiler.describe(sc, strategy), 40, 30, strategy.frames_in_flight)
    assert bytes(p) == bytes(comp.program)
