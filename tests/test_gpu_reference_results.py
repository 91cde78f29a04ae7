"""The reference's result files restated on the MI355X (tests/reference_results.py):
gpu-verify.json (tools/gputest.html) and bench.json (tools/bench.html) come out with
the reference's key structure (tests/golden/reference_result_shape.json, extracted
from the reference's own results/*.json), and every one of the 16 checks passes.
"""
import json
import os

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def fs():
    import fluidsim
    fluidsim.createDevice()
    return fluidsim


@pytest.fixture(scope="module")
def ref_shape():
    with open(os.path.join(HERE, "golden", "reference_result_shape.json")) as f:
        return json.load(f)


def shape(x):
    if isinstance(x, dict):
        return {k: shape(v) for k, v in x.items()}
    if isinstance(x, list):
        return [shape(x[0])] if x else []
    return type(x).__name__ if not isinstance(x, (int, float)) or isinstance(x, bool) else "number"


def test_gpu_verify_has_the_reference_shape_and_every_check_passes(fs, O, ref_shape):
    import reference_results as rr
    v = rr.gpu_verify(fs, O)
    json.dumps(v)
    assert shape(v) == ref_shape["gpu_verify"]
    assert [c["name"] for c in v["checks"]] == ref_shape["check_names"]
    failed = [(c["name"], c["detail"]) for c in v["checks"] if not c["pass"]]
    assert not failed, failed


def test_bench_has_the_reference_shape(fs, O, ref_shape):
    import reference_results as rr
    b = rr.bench_results(fs, O, quick=True)
    json.dumps(b)
    got, want = shape(b), dict(ref_shape["bench"])
    # the host description fields are strings on both sides
    assert got == want
    for row in b["sameAlgorithm"]["rows"]:
        assert row["gpuMs"] > 0 and row["cpuMs"] > 0
    mg = b["equalQuality"]["rows"][0]["multigrid"]
    assert mg["reachedTarget"] and mg["div"] <= 1e-3
    assert b["interactive"]["gpu"] is not None
