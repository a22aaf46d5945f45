"""Batched 2-D shots: one graph-captured launch per step over every shot.

Each shot's traces and final field must equal, bit for bit, those of its own
single-grid run (and the oracle's), for every tile shape of the autotuner,
for odd and even nt, and on repeated graph launches."""
import numpy as np
import pytest

from wavefield_imaging_amd import awefd, synth

pytestmark = pytest.mark.gpu


def _shots(n, nt, seed=5, fdorder=4, **kw):
    rng = np.random.default_rng(seed)
    wav = synth.ricker(nt, synth.DT, freq=4e6)
    out = []
    for s in range(n):
        nz, nx = int(rng.integers(30, 90)), int(rng.integers(40, 150))
        vel = synth.layered_2d(nz, nx, seed=int(rng.integers(1 << 20)), air=max(nz // 5, 2),
                               base_mean=int(nz * 0.7), amp=3.0, corr=6.0)
        col = int(rng.integers(nx // 4, 3 * nx // 4))
        sou = np.array([[col * 5.0 + 1.25, 7 * 5.0]], np.float32)
        rec = np.stack([np.arange(nx) * 5.0, np.full(nx, 6 * 5.0)], 1).astype(np.float32)
        par = dict(n=[nz, nx], o=[0.0, 0.0], d=[5.0, 5.0], nt=nt, dt=synth.DT, nb=int(rng.integers(6, 20)),
                   dabc=True, fdorder=fdorder, expl=False)
        par.update(kw)
        out.append((par, vel, sou, rec, wav))
    return out


def _models(shots):
    ms = []
    for par, vel, sou, rec, wav in shots:
        m = awefd.Model("cart2d", par)
        m.set_model(vel)
        m.set_sources(sou)
        m.set_receivers(rec)
        m.set_wavelet(wav)
        ms.append(m)
    return ms


def _single(shots):
    out = []
    for sh in shots:
        tr, _, field = awefd.run("cart2d", *sh)
        out.append((tr, field))
    return out


@pytest.mark.parametrize("nt", [61, 60])
def test_batch_every_tile_matches_single_runs(orc, nt):
    shots = _shots(7, nt)
    print("batch of", len(shots), "shots nt", nt, flush=True)
    ref = _single(shots)
    ms = _models(shots)
    b = awefd.Batch(ms)
    for rx, grp in ((2, 0), (4, 1), (8, 0), (16, 2)):
        b.set_tile(rx, grp)          # 1-2 MiB groups: several groups of these small shots
        b.run()
        for s, m in enumerate(ms):
            np.testing.assert_array_equal(m.traces(), ref[s][0], err_msg="shot %d rx %d" % (s, rx))
            np.testing.assert_array_equal(m.field(), ref[s][1], err_msg="shot %d rx %d" % (s, rx))
    # the single runs agree with the oracle
    for s in (0, len(shots) - 1):
        tr, fin, _ = orc.run("cart2d", *shots[s])
        np.testing.assert_array_equal(ref[s][0], tr)
        np.testing.assert_array_equal(ref[s][1], fin)
    b.close()
    for m in ms:
        m.close()


def test_batch_tune_and_replays(orc):
    shots = _shots(5, 50, seed=9, fdorder=8)
    ref = _single(shots)
    ms = _models(shots)
    b = awefd.Batch(ms)
    rx, grp, times = b.tune(steps=10)
    print("tuned rx", rx, "group", grp, times, flush=True)
    assert rx in (2, 4, 8, 16) and grp in b.GROUP_MB and all(t > 0 for t in times.values())
    for rep in range(3):
        b.run()
        # eager work on the same device between graph launches
        _ = awefd.run("cart2d", *shots[rep])
        for s, m in enumerate(ms):
            np.testing.assert_array_equal(m.traces(), ref[s][0], err_msg="replay %d shot %d" % (rep, s))
            np.testing.assert_array_equal(m.field(), ref[s][1], err_msg="replay %d shot %d" % (rep, s))
    b.close()
    for m in ms:
        m.close()


def test_batch_scalar_order_and_free_surface(orc):
    shots = _shots(4, 40, seed=3, sse=False, free=True)
    ref = _single(shots)
    ms = _models(shots)
    b = awefd.Batch(ms)
    b.run()
    for s, m in enumerate(ms):
        np.testing.assert_array_equal(m.traces(), ref[s][0])
    tr, _, _ = orc.run("cart2d", *shots[1])
    np.testing.assert_array_equal(ref[1][0], tr)
    b.close()
    for m in ms:
        m.close()


@pytest.mark.parametrize("nt,fdorder,kw", [
    (61, 4, {}), (60, 4, {}), (59, 8, {}), (62, 12, {}), (57, 4, dict(sse=False)),
    (58, 6, dict(free=True)), (55, 4, dict(optfd=True)),
])
def test_two_steps_per_pass_match_single_runs(orc, nt, fdorder, kw):
    """Depth 2 and 3 (two or three time steps per launch on overlapping
    tiles, each step's field through LDS; with three, the middle trace row
    extracted in-block): every shot's traces and final field equal its own
    single-step run bit for bit, for nt of every residue, every stencil
    order, both operation orders, one and several shot groups, and on graph
    replays."""
    shots = _shots(6, nt, seed=11, fdorder=fdorder, **kw)
    print("depth-2 batch nt", nt, "fdorder", fdorder, kw, flush=True)
    ref = _single(shots)
    ms = _models(shots)
    b = awefd.Batch(ms)
    for rx, grp, dep in ((4, 0, 2), (8, 1, 2), (4, 1, 2), (4, 0, 3), (8, 1, 3), (4, 1, 3)):
        b.set_tile(rx, grp)
        b.set_depth(dep)
        for rep in range(2):
            b.run()
            for s, m in enumerate(ms):
                msg = "shot %d rx %d depth %d rep %d" % (s, rx, dep, rep)
                np.testing.assert_array_equal(m.traces(), ref[s][0], err_msg=msg)
                np.testing.assert_array_equal(m.field(), ref[s][1], err_msg=msg)
    tr, fin, _ = orc.run("cart2d", *shots[0])
    np.testing.assert_array_equal(ref[0][0], tr)
    b.close()
    for m in ms:
        m.close()


def test_two_steps_sources_on_tile_edges_and_tuning(orc):
    """Several sources per shot, some on the overlap of neighbouring tiles
    (each tile injects into its whole first-step region), expl=y; then the
    depth-aware autotuner, whose choice must give the same results."""
    shots = _shots(4, 40, seed=21, expl=True)
    out = []
    for par, vel, sou, rec, wav in shots:
        nz, nx = par["n"]
        xs = np.array([11, 12, 23, 24, 35, 36, 47]) % nx
        sou = np.stack([xs * 5.0 + 0.5, np.full(len(xs), (nz // 2) * 5.0 + 2.0)], 1).astype(np.float32)
        out.append((par, vel, sou, rec, wav))
    ref = _single(out)
    ms = _models(out)
    b = awefd.Batch(ms)
    for dep in (2, 3):
        b.set_tile(4, 0)
        b.set_depth(dep)
        b.run()
        for s, m in enumerate(ms):
            np.testing.assert_array_equal(m.traces(), ref[s][0], err_msg="shot %d depth %d" % (s, dep))
    rx, grp, depth, times = b.tune_depth(steps=10)
    print("tuned rx", rx, "group", grp, "depth", depth, flush=True)
    assert depth in (1, 2, 3) and all(t > 0 for t in times.values())
    b.run()
    for s, m in enumerate(ms):
        np.testing.assert_array_equal(m.traces(), ref[s][0], err_msg="tuned shot %d" % s)
        np.testing.assert_array_equal(m.field(), ref[s][1], err_msg="tuned shot %d" % s)
    b.close()
    for m in ms:
        m.close()
