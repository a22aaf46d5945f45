"""hybridbc=y and adj=y of the cartesian programs: HIP kernels vs the CPU
oracle (bit-exact FP32), through the C ABI."""
import numpy as np
import pytest

from wavefield_imaging_amd import awefd
from test_gpu_cart import _case2d, _case3d, _density

pytestmark = pytest.mark.gpu


def _run_both(kind, par, vel, sou, rec, wav, orc, den=None, jsnap=0):
    print("case", kind, par["n"], {k: par[k] for k in par if k not in ("n", "o", "d")}, flush=True)
    with awefd.Model(kind, par) as m:
        m.set_model(vel, den)
        m.set_sources(sou)
        m.set_receivers(rec)
        m.set_wavelet(wav)
        tr, sn = m.run(jsnap=jsnap, snapshots=jsnap > 0)
        field = m.field()
        # repeat plain runs: the 2-D loop is captured as a graph on the second
        tr2, _ = m.run()
        f2 = m.field()
        tr3, _ = m.run()
        f3 = m.field()
    ref, ufin, rsn = orc.run(kind, dict(par, snap=jsnap > 0, jsnap=jsnap or par["nt"]), vel, sou, rec, wav,
                             den=den, snaps=jsnap > 0)
    assert np.any(ufin != 0)
    np.testing.assert_array_equal(tr, ref)
    np.testing.assert_array_equal(field, ufin)
    for t, f in ((tr2, f2), (tr3, f3)):
        np.testing.assert_array_equal(t, ref)
        np.testing.assert_array_equal(f, ufin)
    assert [it for it, _ in sn] == [it for it, _ in rsn]
    for (_, a), (_, b) in zip(sn, rsn):
        np.testing.assert_array_equal(a, b)
    return ref, ufin


@pytest.mark.parametrize("kw", [
    dict(), dict(free=True), dict(nb=1), dict(nb=None, fdorder=8), dict(fdorder=6, nb=9),
    dict(sse=False, free=True), dict(optfd=True), dict(dabc=False, free=True),
])
def test_cart2d_hybrid_bit_exact(orc, kw):
    par, vel, sou, rec, wav = _case2d(nt=160, **dict(dict(hybridbc=True), **kw))
    sou = np.array([[30.0, 20.0]], np.float32)      # near a corner: the wave crosses every strip
    _run_both("cart2d", par, vel, sou, rec, wav, orc, jsnap=37)


@pytest.mark.parametrize("kw", [dict(), dict(free=True), dict(nb=2), dict(fdorder=6, nb=8), dict(sse=False)])
def test_cart3d_hybrid_bit_exact(orc, kw):
    par, vel, sou, rec, wav = _case3d(nt=70, **dict(dict(hybridbc=True), **kw))
    _run_both("cart3d", par, vel, sou, rec, wav, orc, jsnap=23)


def test_hybrid_variable_density(orc):
    par, vel, sou, rec, wav = _case2d(nt=140, hybridbc=True, cden=False, free=True)
    _run_both("cart2d", par, vel, sou, rec, wav, orc, den=_density(vel, 3))


@pytest.mark.parametrize("expl,kt,kw", [
    (True, 1, dict()), (True, 45, dict()), (False, 30, dict()), (False, 1, dict(hybridbc=True)),
    (True, 100, dict(free=True)), (False, 7, dict(fdorder=8, sse=False)),
])
def test_cart2d_adj_image_bit_exact(orc, expl, kt, kw):
    par, vel, sou, rec, wav = _case2d(nt=120, expl=expl, adj=True, kt=kt, **kw)
    rec = rec[::3]
    k = 1 if expl else len(rec)
    rng = np.random.default_rng(kt)
    wav = (wav[:, None] * rng.standard_normal(k)[None, :]).astype(np.float32)
    ref, img = _run_both("cart2d", par, vel, sou, rec, wav, orc, jsnap=11)
    assert orc.imaged() and np.all(ref == 0)


@pytest.mark.parametrize("expl,kw", [(True, dict()), (False, dict(hybridbc=True)), (False, dict(free=True))])
def test_cart3d_adj_bit_exact(orc, expl, kw):
    par, vel, sou, rec, wav = _case3d(nt=60, expl=expl, adj=True, **kw)
    sou = np.concatenate([sou, sou + np.float32([[20.0, -10.0, 5.0]])])
    w = np.stack([wav, -0.7 * wav[::-1]], 1).astype(np.float32)
    _run_both("cart3d", par, vel, sou, rec, wav if expl else w, orc)


def test_adj_and_hybrid_errors(orc):
    par, vel, sou, rec, wav = _case2d(nt=50, adj=True, kt=0)
    with pytest.raises(awefd.L.AwefdError):
        awefd.Model("cart2d", par)
    with pytest.raises(awefd.L.AwefdError):
        awefd.Model("cart2d", dict(par, kt=51))
    # without expl the 2-D adjoint loop reads nr samples per step
    with awefd.Model("cart2d", dict(par, kt=5)) as m:
        m.set_model(vel)
        m.set_sources(sou)
        m.set_receivers(rec)
        m.set_wavelet(wav)                      # nt samples, not nt * nr
        with pytest.raises(awefd.L.AwefdError):
            m.run()
    # batches take neither mode
    ms = []
    for kw in (dict(adj=True, kt=3, expl=True), dict(hybridbc=True)):
        p = _case2d(nt=50, **kw)[0]
        m = awefd.Model("cart2d", p)
        m.set_model(vel)
        m.set_sources(sou)
        m.set_receivers(rec)
        m.set_wavelet(wav)
        ms.append(m)
        with pytest.raises(awefd.L.AwefdError):
            awefd.Batch([m])
    for m in ms:
        m.close()


@pytest.mark.parametrize("kw", [dict(), dict(expl=True), dict(fdorder=8, sse=False), dict(hybridbc=True, free=True),
                                dict(adj=True, kt=9), dict(adj=True, kt=30, expl=True), dict(dabc=False)])
def test_cart2d_sinc_bit_exact(orc, kw):
    """sinc=y: 9 x 9 taps per point; overlapping taps of nearby sources add
    in source order; on-node and off-node points."""
    par, vel, sou, rec, wav = _case2d(nt=110, nb=10, sinc=True, **kw)
    sou = np.float32([[101.3, 62.2], [104.0, 60.0], [140.7, 33.9]])
    rec = np.float32([[x, 41.3 + 0.37 * i] for i, x in enumerate(np.arange(45.0, 240.0, 13.1))])
    adj2 = kw.get("adj", False)
    k = 1 if kw.get("expl") else (len(rec) if adj2 else len(sou))
    rng = np.random.default_rng(11)
    w = (wav[:, None] * rng.standard_normal(k)[None, :]).astype(np.float32)
    ref, _ = _run_both("cart2d", par, vel, sou, rec, w, orc, jsnap=17)
    if not adj2:
        assert np.abs(ref).max() > 0


@pytest.mark.parametrize("kw", [dict(), dict(expl=True, free=True), dict(hybridbc=True), dict(adj=True)])
def test_cart3d_sinc_bit_exact(orc, kw):
    par, vel, sou, rec, wav = _case3d(nt=50, nb=7, sinc=True, **kw)
    sou = np.float32([[71.3, 72.9, 33.4], [74.0, 70.0, 35.0]])
    rec = rec[::4] + np.float32([[0.7, -1.3, 2.1]])
    k = 1 if kw.get("expl") else len(sou)
    w = (wav[:, None] * np.float32([1.0, -0.6])[None, :k]).astype(np.float32)
    _run_both("cart3d", par, vel, sou, rec, w, orc, jsnap=21)


def test_sinc_taps_must_stay_on_the_grid(orc):
    par, vel, sou, rec, wav = _case2d(nt=20, nb=2, sinc=True)
    with awefd.Model("cart2d", par) as m:
        m.set_model(vel)
        with pytest.raises(awefd.L.AwefdError):
            m.set_sources(np.float32([[3.0, 90.0]]))      # 4 taps toward x < 0 leave the grid
        m.set_sources(np.float32([[0.0, 90.0]]))          # on a node: one tap


@pytest.mark.parametrize("kw", [dict(hybridbc=True), dict(hybridbc=True, free=True), dict(adj=True, expl=True),
                                dict(sinc=True, nb=7), dict(sinc=True, nb=7, adj=True)])
def test_cart3d_modes_with_last_z_tile_in_rows(orc, kw):
    """hybridbc, adj and sinc over grids whose last z tile runs as 16-lane rows
    of 4 x tiles (padded z 72 = 64 + 8, or 74 with nb = 7), with a source
    whose taps reach into that tile and snapshots."""
    nz, nx, ny = 60, 34, 20
    par, vel, sou, rec, wav = _case3d(nz=nz, nx=nx, ny=ny, nt=50, **kw)
    sou = np.float32([[nx * 2.5 + 1.3, ny * 2.5 - 2.1, 6 * 5.0 + 0.7], [71.3, 52.9, (nz - 6) * 5.0 + 0.4]])
    rec = np.concatenate([rec[::3], rec[::5] + np.float32([[0.7, -1.3, (nz - 9) * 5.0]])]).astype(np.float32)
    k = 1 if kw.get("expl") else len(sou)
    w = (wav[:, None] * np.float32([1.0, -0.6])[None, :k]).astype(np.float32)
    _run_both("cart3d", par, vel, sou, rec, w if k > 1 else wav, orc, jsnap=19)
