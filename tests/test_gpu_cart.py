"""Cartesian HIP kernels vs the CPU oracle (bit-exact FP32), through the C ABI."""
import numpy as np
import pytest

from wavefield_imaging_amd import awefd, synth

pytestmark = pytest.mark.gpu


def _case2d(nz=40, nx=56, nt=120, nb=8, **kw):
    c = synth.cart2d(scale=0.1, nt=nt)
    vel = synth.layered_2d(nz, nx, seed=11, air=6, base_mean=25, amp=3.0, corr=5.0)
    sou = np.array([[nx * 5 * 0.45, 9 * 5.0]], np.float32)
    rec = np.stack([np.arange(nx) * 5.0, np.full(nx, 7 * 5.0)], 1).astype(np.float32)
    par = dict(n=[nz, nx], o=[0.0, 0.0], d=[5.0, 5.0], nt=nt, dt=synth.DT, nb=nb,
               dabc=True, cden=True, fdorder=4, expl=False)
    par.update(kw)
    return par, vel, sou, rec, c["wav"][:nt]


def _case3d(nz=24, nx=28, ny=30, nt=60, nb=6, **kw):
    vel = synth.layered_3d(nz, nx, ny, seed=5, air=5, base_mean=15, amp=2.0, corr=4.0)
    sou = np.array([[nx * 2.5 + 1.3, ny * 2.5 - 2.1, 6 * 5.0 + 0.7]], np.float32)
    ix, iy = np.meshgrid(np.arange(0, nx, 3), np.arange(0, ny, 3), indexing="ij")
    rec = np.stack([ix.ravel() * 5.0, iy.ravel() * 5.0, np.full(ix.size, 4 * 5.0)], 1).astype(np.float32)
    par = dict(n=[nz, nx, ny], o=[0.0] * 3, d=[5.0] * 3, nt=nt, dt=synth.DT, nb=nb,
               dabc=True, cden=True, fdorder=4, expl=False)
    par.update(kw)
    return par, vel, sou, rec, synth.ricker(nt, synth.DT)


def _check(kind, par, vel, sou, rec, wav, orc, den=None):
    print("case", kind, par["n"], {k: par[k] for k in par if k not in ("n", "o", "d")},
          "den" if den is not None else "", flush=True)
    tr, _, field = awefd.run(kind, par, vel, sou, rec, wav, den=den)
    ref, ufin, _ = orc.run(kind, par, vel, sou, rec, wav, den=den)
    assert np.any(ref != 0)
    np.testing.assert_array_equal(tr, ref)
    np.testing.assert_array_equal(field, ufin)


@pytest.mark.parametrize("kw", [
    dict(), dict(dabc=False), dict(fdorder=2), dict(fdorder=6), dict(fdorder=8),
    dict(sse=False), dict(optfd=True), dict(free=True), dict(nb=None),
])
def test_cart2d_bit_exact(orc, kw):
    _check("cart2d", *_case2d(**kw), orc)


@pytest.mark.parametrize("narrow", ["0", "1"])
@pytest.mark.parametrize("kw", [dict(), dict(fdorder=8), dict(free=True, sse=False), dict(dabc=False, nb=3)])
def test_cart2d_both_tiles(orc, narrow, kw, monkeypatch, capfd):
    """Both 2-D step tiles, forced: 4 waves x 4 x rows (the default once a
    grid gives 3 such blocks per CU) and one wave x 3 rows (smaller grids).
    151 padded x rows: neither 16 nor 3 divides the width."""
    monkeypatch.setenv("AWEFD_VERBOSE", "1")
    monkeypatch.setenv("AWEFD_NARROW", narrow)
    par, vel, sou, rec, wav = _case2d(nz=70, nx=151 - 2 * kw.get("nb", 8), nt=90, **kw)
    sou = np.array([[40.0, 45.0], [(par["n"][1] - 2) * 5.0, 300.0], [371.5, 12.0]], np.float32)
    _check("cart2d", par, vel, sou, rec, np.stack([wav, -wav, 0.5 * wav], 1).astype(np.float32), orc)
    err = capfd.readouterr().err
    assert ("one-wave tiles" in err) == (narrow == "1"), err


def test_cart2d_many_sources_expl(orc):
    par, vel, sou, rec, wav = _case2d(expl=True)
    sou = np.stack([np.arange(3, 50, 2) * 5.0 + 0.5, np.full(24, 12 * 5.0 + 1.2)], 1).astype(np.float32)
    _check("cart2d", par, vel, sou, rec, wav, orc)


def test_cart2d_many_sources_separate_wavelets(orc):
    par, vel, sou, rec, wav = _case2d()
    sou = np.array([[100.0, 60.0], [100.0, 60.0], [103.0, 61.0], [-50.0, 10.0]], np.float32)
    w4 = np.stack([wav, -0.5 * wav, 2 * wav, wav], 1).astype(np.float32)
    _check("cart2d", par, vel, sou, rec, w4, orc)


@pytest.mark.parametrize("kw", [dict(), dict(dabc=False), dict(free=True), dict(fdorder=6),
                                dict(sse=False), dict(nb=3)])
def test_cart3d_bit_exact(orc, kw):
    _check("cart3d", *_case3d(**kw), orc)


def _density(vel, seed):
    """Density loosely tied to the velocity layers plus a smooth perturbation."""
    rng = np.random.default_rng(seed)
    d = 900.0 + vel.astype(np.float64) / 2.5e5
    d *= 1.0 + 0.1 * np.sin(np.cumsum(rng.standard_normal(vel.shape), axis=-1) / 8.0)
    return d.astype(np.float32)


@pytest.mark.parametrize("kind,kw", [
    ("cart2d", dict()), ("cart2d", dict(sse=False)), ("cart2d", dict(fdorder=8, dabc=False)),
    ("cart3d", dict()), ("cart3d", dict(sse=False)), ("cart3d", dict(fdorder=6, free=True)),
])
def test_variable_density_bit_exact(orc, kind, kw):
    """cden=n with den= (ts_kernel.c:101-135, 228-280)."""
    case = _case2d if kind == "cart2d" else _case3d
    par, vel, sou, rec, wav = case(cden=False, **kw)
    _check(kind, par, vel, sou, rec, wav, orc, den=_density(vel, 8))


def test_constant_density_flag_ignores_den(orc):
    """cden=y runs constant density even when den= is given."""
    par, vel, sou, rec, wav = _case2d(cden=True)
    den = _density(vel, 9)
    tr, _, _ = awefd.run("cart2d", par, vel, sou, rec, wav, den=den)
    ref, _, _ = orc.run("cart2d", par, vel, sou, rec, wav)
    np.testing.assert_array_equal(tr, ref)


def test_cart2d_graph_replays_setters_and_eager_path(orc):
    """The first run of a 2-D loop launches step by step, a repeated run
    captures the loop as one graph and later runs replay it: the direct run,
    the capture, replays with other work in between, a setter (new sources),
    a caller trace buffer and the profiled per-launch path all give the
    oracle's traces."""
    import ctypes as C

    par, vel, sou, rec, wav = _case2d(nt=97)
    ref, ufin, _ = orc.run("cart2d", par, vel, sou, rec, wav)
    sou2 = np.array([[61.0, 52.0]], np.float32)
    ref2, ufin2, _ = orc.run("cart2d", par, vel, sou2, rec, wav)
    m = awefd.Model("cart2d", par)
    m.set_model(vel)
    m.set_sources(sou)
    m.set_receivers(rec)
    m.set_wavelet(wav)
    for rep in range(3):
        tr, _ = m.run()
        np.testing.assert_array_equal(tr, ref, err_msg="replay %d" % rep)
        np.testing.assert_array_equal(m.field(), ufin)
        _ = awefd.run("cart2d", par, vel, sou2, rec, wav)      # other work between replays
    m.set_sources(sou2)
    tr, _ = m.run()
    np.testing.assert_array_equal(tr, ref2)
    np.testing.assert_array_equal(m.field(), ufin2)
    m.profile(True)                                           # eager path
    tr, _ = m.run()
    m.profile(False)
    np.testing.assert_array_equal(tr, ref2)
    np.testing.assert_array_equal(m.field(), ufin2)
    import torch

    buf = torch.zeros((97, len(rec)), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    m.run_dev(buf.data_ptr())
    np.testing.assert_array_equal(buf.cpu().numpy(), ref2)
    m.close()


@pytest.mark.parametrize("nz,kw", [
    (60, dict()),                                   # padded z 72 = 64 + 8
    (68, dict(fdorder=8)),                          # 80 = 64 + 16, nop 4
    (60, dict(free=True, sse=False)),
    (60, dict(dabc=False)),
    (60, dict(fdorder=12)),                         # nop 6: 72 = 64 + 8, 2 nop halo lanes of 16
    (53, dict()),                                   # 65 = 64 + 1: z-high ABC in its own pass
    (4, dict()),                                    # 16: the whole z axis in 16-lane rows
    (60, dict(cden=False)),                         # variable density
])
def test_cart3d_last_z_tile_in_rows(orc, nz, kw, monkeypatch, capfd):
    """Grids whose last z tile holds at most 16 nodes run it as four 16-lane
    rows of 4 x tiles per wave (z neighbours by row shifts, each row's
    addresses offset by its x tile).  82 x rows: 41 x tiles, so some XCDs
    own a group of fewer than 4; sources and receivers in the last z tile."""
    monkeypatch.setenv("AWEFD_VERBOSE", "1")
    nx, ny, nb = 70, 18, 6
    den = kw.pop("cden", True)
    par, vel, sou, rec, wav = _case3d(nz=nz, nx=nx, ny=ny, nt=50, nb=nb, cden=den, **kw)
    zdeep = (nz - 1.6) * 5.0
    sou = np.array([[nx * 2.5 + 1.3, ny * 2.5 - 2.1, min(6 * 5.0 + 0.7, zdeep)],
                    [3 * 5.0 + 0.4, 4 * 5.0 + 2.2, zdeep], [(nx - 2) * 5.0, (ny - 3) * 5.0, zdeep]], np.float32)
    ix, iy = np.meshgrid(np.arange(0, nx, 4), np.arange(0, ny, 4), indexing="ij")
    rec = np.concatenate([
        np.stack([ix.ravel() * 5.0, iy.ravel() * 5.0, np.full(ix.size, min(4 * 5.0, zdeep))], 1),
        np.stack([ix.ravel() * 5.0 + 1.1, iy.ravel() * 5.0 + 0.6, np.full(ix.size, (nz - 1.3) * 5.0)], 1),
    ]).astype(np.float32)
    w3 = np.stack([wav, -0.5 * wav, 2 * wav], 1).astype(np.float32)   # one wavelet per source
    _check("cart3d", par, vel, sou, rec, w3, orc, den=None if den else _density(vel, 3))
    err = capfd.readouterr().err
    assert "last z tile in 16-lane rows" in err, err
