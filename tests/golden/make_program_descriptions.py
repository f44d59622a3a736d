#!/usr/bin/env python3
"""Extract the reference's frozen ProgramDescription link maps
(tests/compiler/__snapshots__/program-description.snapshot.test.ts.snap, 505 cases) into
tests/golden/program_descriptions.json: for every case the fields a lowering reads (the rest of
each record pruned), plus five complete records. Runs where /root/reference exists; the output
is committed data."""
import json
import os
import re
import sys

SNAP = "tests/compiler/__snapshots__/program-description.snapshot.test.ts.snap"
KEEP = {
    "measurement": ["color", "sensor", "maxBounces", "maxNullCrossings", "response", "shadows", "ambient",
                    "scattering", "aovs", "camera"],
    "estimator": None, "materials": ["models", "twoSidedShading", "materialsReadUv", "albedo", "emissionCones"],
    "media": ["deflecting", "heterogeneousArms", "models"], "intersection": ["drivenPlacement"],
    "spectrumKinds": None, "environment": ["type", "glsl"],
}
FULL = ["analytic-glass + pathtracer", "proc-sky + pt-mis", "cornell + orthographic", "emit-sat + pathtracer",
        "spectral-dispersion-slab + spectral"]


def parse(ref):
    src = open(os.path.join(ref, SNAP)).read()
    out = {}
    for name, body in re.findall(r"exports\[`ProgramDescription structural snapshot \(the link map\) > (.*?) 1`\] = `\n(.*?)\n`;",
                                 src, re.S):
        text = re.sub(r":\s*undefined", ": null", body)      # JSON.stringify drops them; null reads the same
        text = re.sub(r",(\s*[}\]])", r"\1", text)
        out[name] = json.loads(text)
    return out


def prune(d):
    p = {}
    for k, keep in KEEP.items():
        if k not in d:
            continue
        v = d[k]
        if keep is None or not isinstance(v, dict):
            p[k] = v
        else:
            p[k] = {f: v[f] for f in keep if f in v}
            if k == "measurement" and isinstance(p[k].get("camera"), dict):
                p[k]["camera"] = {"type": p[k]["camera"].get("type")}
    return p


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    cases = parse(ref)
    full = {k: cases[k] for k in FULL if k in cases}
    out = {"source": SNAP, "cases": {k: prune(v) for k, v in sorted(cases.items())}, "full": full}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "program_descriptions.json")
    json.dump(out, open(path, "w"), separators=(",", ":"), sort_keys=True)
    print(f"wrote {len(out['cases'])} cases ({len(full)} full) to {path}")


if __name__ == "__main__":
    main()
