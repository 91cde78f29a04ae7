"""Extract the reference's own measurement records for the hot path into a fixture.

Reads results/gpu-verify.json of the reference checkout (data, not code) and
writes tests/golden/reference_held.json. Run only where the reference is mounted;
the committed fixture is what tests read.
"""
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_held.json")

d = json.load(open(os.path.join(REF, "results", "gpu-verify.json")))
fixture = {
    "source": "results/gpu-verify.json (tools/gputest.html, Intel Iris Xe, FP32 WGSL)",
    "grid": 64,
    "seeds": {"divergence": 991, "projection": 4242, "curves": 777},
    "projection": d["projection"],
    "poissonCurves": {k: [[p["iterations"], p["residual"]] for p in v] for k, v in d["poissonCurves"].items()},
    "schemeEnergy": d["schemeEnergy"],
    "advection": d["advection"],
    "divergence": d["divergence"],
    "fullStep": d["fullStep"],
    "macCormack": d["macCormack"],
}
json.dump(fixture, open(OUT, "w"), indent=1)
print(OUT)
