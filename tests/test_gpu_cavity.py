"""Batched lid-driven cavities on the GPU against the f64 oracle's walled step
(oracle/fsoracle.c or_cav_*: advection with wall ghosts, explicit diffusion,
pinned-disk forcing, exact cosine-basis Neumann projection)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return math.sqrt(np.sum((a - b) ** 2) / max(np.sum(b * b), 1e-300))


@pytest.fixture(scope="module")
def fs():
    import fluidsim
    fluidsim.createDevice()
    return fluidsim


@pytest.mark.parametrize("scheme", ["semi-lagrangian", "maccormack"])
def test_single_cavity_matches_oracle(fs, O, scheme):
    # one 64^2 cavity with a pinned disk, 20 steps at the batch's own dt
    cav = dict(n=64, lid=1.0, nu=2e-3, disk=dict(x=0.55, y=0.45, r=0.12))
    o = O.Cavity(64, U=1.0, nu=2e-3, advection=scheme, disk=cav["disk"], passes=2)
    with fs.CavityBatch([cav], advection=scheme, tolerance=1e-9) as b:
        for _ in range(20):
            b.stepCFL(0.5)
            dt = b.lastDt()[0]
            assert abs(dt - o.dt_cfl(0.5)) <= 2e-5 * dt
            o.step(dt)
            assert b.stats()["last_residual"] <= 1e-9
        out = b.readVelocity(0)
    err = max(rel(out["u"], o.u), rel(out["v"], o.v))
    assert err < 1e-4, err
    uu = out["u"].reshape(64, 64)
    assert np.all(uu[:, 0] == 0) and np.all(out["v"].reshape(64, 64)[0, :] == 0)
    d = O.cav_divergence(64, out["u"].astype(np.float64), out["v"].astype(np.float64))
    assert np.max(np.abs(d)) < 1e-3     # FP32 storage of a projected field


def test_batch_members_match_their_own_runs(fs, O):
    # four cavities of three sizes in one batch, in the caller's order, against the
    # same cavities stepped alone (each run meets the same FP64 gate)
    cavs = [dict(n=64, lid=1.0, nu=1e-2, disk=dict(x=0.5, y=0.5, r=0.1)),
            dict(n=256, lid=0.7, nu=2e-4, disk=None),
            dict(n=128, lid=1.5, nu=1e-3, disk=dict(x=0.3, y=0.6, r=0.08)),
            dict(n=64, lid=2.0, nu=5e-3, disk=None)]
    steps = 6
    with fs.CavityBatch(cavs, tolerance=1e-8) as b:
        for _ in range(steps):
            b.stepCFL(0.5)
            assert b.stats()["last_residual"] <= 1e-8
        together = [b.readVelocity(k) for k in range(len(cavs))]
        st = b.stats()
    assert st["cells"] == sum(c["n"] ** 2 for c in cavs) and st["steps"] == steps
    for k, c in enumerate(cavs):
        with fs.CavityBatch([c], tolerance=1e-8) as s:
            for _ in range(steps):
                s.stepCFL(0.5)
            alone = s.readVelocity(0)
        assert max(rel(together[k]["u"], alone["u"]), rel(together[k]["v"], alone["v"])) < 1e-5, k


def test_random_cavity_draws(fs):
    cavs = fs.random_cavities(256, seed=1)
    assert len(cavs) == 256 and {c["n"] for c in cavs} == {64, 128, 256, 512}
    for c in cavs:
        h = 1.0 / c["n"]
        d = c["disk"]
        assert 0.5 <= c["lid"] <= 2 and 100 <= c["lid"] / c["nu"] <= 5000
        assert 0.05 <= d["r"] <= 0.2
        assert d["r"] + 2 * h <= d["x"] <= 1 - d["r"] - 2 * h and d["r"] + 2 * h <= d["y"] <= 1 - d["r"] - 2 * h


@pytest.mark.parametrize("pre,post", [(2, 2), (1, 3), (3, 1)])
def test_fused_cavity_multigrid_gives_the_same_bits(fs, pre, post):
    # the LDS-tiled level kernels against the one-pass-per-operator kernels
    cavs = [dict(n=256, lid=1.0, nu=1e-3, disk=dict(x=0.4, y=0.5, r=0.1)),
            dict(n=128, lid=1.5, nu=2e-3, disk=None),
            dict(n=512, lid=0.8, nu=4e-4, disk=dict(x=0.7, y=0.3, r=0.15))]
    outs = []
    for fused in (True, False):
        with fs.CavityBatch(cavs, tolerance=1e-7, mgPre=pre, mgPost=post, mgFused=fused) as b:
            for _ in range(4):
                b.stepCFL(0.5)
            outs.append([b.readVelocity(k) for k in range(len(cavs))])
            cyc = b.stats()["total_vcycles"]
    for a, c in zip(*outs):
        assert np.array_equal(a["u"], c["u"]) and np.array_equal(a["v"], c["v"])
    assert cyc > 0


@pytest.mark.parametrize("guess", ["zero", "previous"])
def test_graph_replay_gives_the_same_bits(fs, guess):
    # each step's head replayed as a captured hipGraph against the same launches one
    # by one: every step compared, with reads (synchronisations) between the replays
    cavs = [dict(n=256, lid=1.0, nu=1e-3, disk=dict(x=0.4, y=0.5, r=0.1)),
            dict(n=64, lid=2.0, nu=4e-4, disk=dict(x=0.5, y=0.5, r=0.2)),
            dict(n=512, lid=0.8, nu=4e-4, disk=None),
            dict(n=128, lid=1.5, nu=2e-3, disk=None)]
    steps = 8
    with fs.CavityBatch(cavs, tolerance=1e-7, guess=guess, graph=True) as g, \
            fs.CavityBatch(cavs, tolerance=1e-7, guess=guess, graph=False) as e:
        for s in range(steps):
            g.stepCFL(0.5)
            e.stepCFL(0.5)
            for k in range(len(cavs)):
                a, c = g.readVelocity(k), e.readVelocity(k)
                assert np.array_equal(a["u"], c["u"]) and np.array_equal(a["v"], c["v"]), (s, k)
            assert g.stats()["last_vcycles"] == e.stats()["last_vcycles"]
        sg, se = g.stats(), e.stats()
    assert sg["graph_replays"] == steps and se["graph_replays"] == 0
    assert 1 <= sg["graph_captures"] < steps          # heads are reused across steps


@pytest.mark.parametrize("tile", ["64x32", "32x32", "16x16", "8x8", "64x64", "32x32/1024", "16x16/256"])
def test_every_cavity_tile_shape_gives_the_same_bits(fs, tile):
    # what makes the tile autotuner safe: each shape against the one-pass kernels
    cavs = [dict(n=512, lid=1.0, nu=1e-3, disk=dict(x=0.4, y=0.5, r=0.1)),
            dict(n=128, lid=1.5, nu=2e-3, disk=None),
            dict(n=256, lid=0.8, nu=4e-4, disk=dict(x=0.7, y=0.3, r=0.15))]
    outs = []
    for fused in (True, False):
        with fs.CavityBatch(cavs, tolerance=1e-7, mgFused=fused, tile=tile if fused else None) as b:
            for _ in range(3):
                b.stepCFL(0.5)
            outs.append([b.readVelocity(k) for k in range(len(cavs))])
            if fused:
                assert b.tuneReport().count(tile) >= 2 * 3  # down and up of classes 512, 256, 128
    for a, c in zip(*outs):
        assert np.array_equal(a["u"], c["u"]) and np.array_equal(a["v"], c["v"])


def test_cavity_autotuner_picks_tiles_and_sweeps(fs):
    cavs = fs.random_cavities(32, seed=2)
    with fs.CavityBatch(cavs, tolerance=1e-6, autotune=2) as b:
        rep = b.tuneReport()
        st = b.stats()
        assert st["steps"] == 0 and st["total_vcycles"] == 0          # handed back at rest
        u = b.readVelocity(0)
        assert not np.any(u["u"]) and not np.any(u["v"])
        for _ in range(3):
            b.stepCFL(0.5)
            assert b.stats()["last_residual"] <= 1e-6
    print(rep)
    assert rep.count("trial pre") == 9
    for lg in (512, 256, 128):
        assert f"class {lg}^2" in rep and "us," in rep
