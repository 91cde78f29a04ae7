"""Extract the reference's own measurement records for the hot path into fixtures.

Reads results/gpu-verify.json and results/bench.json of the reference checkout
(data, not code) and writes tests/golden/reference_held.json (the numbers) and
tests/golden/reference_result_shape.json (the key structure of both files and
the names of the 16 checks). Run only where the reference is mounted; the
committed fixtures are what tests read.
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


def shape(x):
    """Keys of every dict, in order; a list is represented by its first element."""
    if isinstance(x, dict):
        return {k: shape(v) for k, v in x.items()}
    if isinstance(x, list):
        return [shape(x[0])] if x else []
    return type(x).__name__ if not isinstance(x, (int, float)) or isinstance(x, bool) else "number"


b = json.load(open(os.path.join(REF, "results", "bench.json")))
OUT2 = os.path.join(os.path.dirname(OUT), "reference_result_shape.json")
json.dump({"source": "results/gpu-verify.json, results/bench.json (key structure only)",
           "gpu_verify": shape(d), "bench": shape(b), "check_names": [c["name"] for c in d["checks"]]},
          open(OUT2, "w"), indent=1)
print(OUT2)
