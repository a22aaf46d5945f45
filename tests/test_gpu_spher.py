"""Polar and spherical HIP kernels vs the CPU oracle (bit-exact FP32), through
the C ABI.  The reference programs are MAWEFDspher2Dcpu.c / MAWEFDspher3Dcpu.c."""
import os

import numpy as np
import pytest

from wavefield_imaging_amd import awefd, synth

pytestmark = pytest.mark.gpu

R0 = np.float32(3387.5e3)


def _grid(n, nb):
    """r from 3387.5 km (dr = 5 m); d(theta) = d(phi) = 5 m at 3388.5 km,
    theta centred on pi/4 (the configuration's choice)."""
    dth = np.float32(5.0 / 3388.5e3)
    o = [R0, np.float32(np.pi / 4 - n[1] / 2 * dth), np.float32(0.4)][: len(n)]
    d = [np.float32(5.0), dth, dth][: len(n)]
    return o, d


def _vel(n, seed, top=6):
    """Layered model, array (phi, theta, r) / (theta, r): a rough base in r
    under a 3e8 top layer (the outer radial nodes)."""
    rng = np.random.default_rng(seed)
    v = np.full(n[::-1], synth.V_ICE, np.float32)
    base = n[0] // 2 + rng.integers(-2, 3, size=n[1:][::-1])
    r = np.arange(n[0])
    v[r[(None,) * (len(n) - 1)] < base[..., None]] = synth.V_BASE
    v[..., n[0] - top:] = synth.V_AIR
    return v


def _coords(o, d, pts):
    """(r, theta[, phi]) in cells -> the coordinate files' order
    (spher2d: (r, theta); spher3d: (r, phi, theta))."""
    pts = np.asarray(pts, np.float64)
    out = []
    for p in pts:
        if len(o) == 2:
            out.append([o[0] + p[0] * d[0], o[1] + p[1] * d[1]])
        else:
            out.append([o[0] + p[0] * d[0], o[2] + p[2] * d[2], o[1] + p[1] * d[1]])
    return np.array(out, np.float32)


def _case(ndim, n=None, nb=6, nt=80, src=None, **kw):
    if n is None:
        n = [30, 28, 26][:ndim]
    o, d = _grid(n, nb)
    vel = _vel(n, seed=3)
    ph = n[2] / 2 if ndim == 3 else 0.0
    if src is None:
        # just under the fast top layer, so the reflectivity (expl=y) is not 0
        src = [[n[0] - 6.5, n[1] / 2 + 0.4, ph - 0.7][:ndim]]
    sou = _coords(o, d, src)
    rr = np.linspace(1.0, n[1] - 2.0, 9)
    rec = _coords(o, d, [[n[0] - 8.6, t, ph + 0.3][:ndim] for t in rr]
                  + [[2.2, n[1] / 2, 1.5][:ndim], [n[0] - 1.5, 1.25, n[-1] - 1.75][:ndim]])
    par = dict(n=n, o=o, d=d, nt=nt, dt=synth.DT, nb=nb, expl=False)
    par.update(kw)
    return par, vel, sou, rec, synth.ricker(nt, synth.DT, freq=8e6)


def _check(kind, par, vel, sou, rec, wav, orc):
    print("case", kind, par["n"], {k: par[k] for k in par if k not in ("n", "o", "d")}, flush=True)
    tr, _, field = awefd.run(kind, par, vel, sou, rec, wav)
    ref, ufin, _ = orc.run(kind, par, vel, sou, rec, wav)
    if not par.get("free") or kind == "spher2d":
        assert np.any(ref != 0)
    np.testing.assert_array_equal(tr, ref)
    np.testing.assert_array_equal(field, ufin)


@pytest.mark.parametrize("kw", [
    dict(), dict(expl=True), dict(free=True), dict(nb=None), dict(nb=12),
    dict(src=[[2.5, 2.25]]), dict(src=[[27.5, 26.75]]), dict(n=[150, 90]),
])
def test_spher2d_bit_exact(orc, kw):
    _check("spher2d", *_case(2, **kw), orc)


def test_spher2d_many_sources(orc):
    par, vel, sou, rec, wav = _case(2, expl=True)
    o, d = par["o"], par["d"]
    src = [[23.5, 10.0], [23.5, 10.0], [23.25, 10.25], [5.75, 3.5], [15.0, 14.5]]
    _check("spher2d", par, vel, _coords(o, d, src), rec, wav, orc)


@pytest.mark.parametrize("kw", [
    dict(), dict(expl=True), dict(nb=None), dict(nb=9),
    dict(src=[[1.5, 1.25, 2.5]]), dict(src=[[27.5, 26.5, 24.25]]),
    dict(n=[24, 18, 30]),                    # n_theta < n_phi - 3: phi-high table layer
])
def test_spher3d_bit_exact(orc, kw):
    _check("spher3d", *_case(3, **kw), orc)


def test_spher3d_free_surface_zeroes(orc):
    par, vel, sou, rec, wav = _case(3, free=True, nt=20)
    tr, _, field = awefd.run("spher3d", par, vel, sou, rec, wav)
    assert not tr.any() and not field.any()


@pytest.mark.parametrize("blocks", ["1", "64", "100000"])
def test_spher3d_tilings(orc, blocks, monkeypatch):
    """Several r tiles, theta tiles and phi chunks (down to one plane)."""
    monkeypatch.setenv("AWEFD_SPHER_BLOCKS", blocks)
    _check("spher3d", *_case(3, n=[140, 40, 36], nt=40,
                             src=[[120.5, 20.25, 18.75], [3.0, 3.0, 3.0]]), orc)


def test_spher3d_many_sources_same_node(orc):
    par, vel, sou, rec, wav = _case(3, expl=True, nt=50)
    o, d = par["o"], par["d"]
    # three next to the fast top layer (two on one node); one in a homogeneous
    # region, whose reflectivity is 0
    src = [[23.5, 13.0, 12.0], [23.5, 13.0, 12.0], [23.25, 13.5, 12.25], [6.0, 6.0, 20.0]]
    _check("spher3d", par, vel, _coords(o, d, src), rec, wav, orc)


@pytest.mark.parametrize("which", ["interior", "shell_receivers", "receiver_on_source", "receivers_on_sources"])
def test_spher3d_receivers_in_shell_and_on_sources(orc, which):
    """Receivers whose corners lie in the shell (r, theta and phi faces: their
    values come from the shell pass) or on injection nodes (read after the
    step, before the next injection), with nb = 4 so that the shell reaches
    into the model."""
    n, nb = [30, 28, 26], 4
    par, vel, sou, rec, wav = _case(3, n=n, nb=nb, nt=60)
    o, d = par["o"], par["d"]
    src = [[n[0] - 6.5, n[1] / 2 + 0.4, n[2] / 2 - 0.7]]
    pts = [[n[0] - 8.6, t, 13.3] for t in np.linspace(3.0, n[1] - 4.0, 5)]   # corners off the shell
    if which == "shell_receivers":
        # padded r 4.5 and 32.5 of 38, theta 4.25 of 36, phi 28.5 of 34
        pts += [[0.5, 10.0, 10.0], [n[0] - 1.5, 12.0, 9.0], [10.0, 0.25, 9.0], [11.0, 13.0, n[2] - 1.5]]
    if which == "receiver_on_source":
        pts += [[n[0] - 6.5, n[1] / 2 + 0.4, n[2] / 2 - 0.7]]
    if which == "receivers_on_sources":
        # two sources; three receivers share corners with them (two on the
        # sources, one elsewhere in the first source's cell)
        src += [[10.2, 7.3, 9.9]]
        pts += [[n[0] - 6.5, n[1] / 2 + 0.4, n[2] / 2 - 0.7], [10.2, 7.3, 9.9],
                [n[0] - 6.5, n[1] / 2 + 0.9, n[2] / 2 - 0.7]]
    par = dict(par, expl=True)
    _check("spher3d", par, vel, _coords(o, d, src), _coords(o, d, pts), wav, orc)


def test_spher3d_source_in_shell(orc):
    par, vel, sou, rec, wav = _case(3, nb=4, nt=50, src=[[0.5, 14.0, 13.0]])
    _check("spher3d", par, vel, sou, rec, wav, orc)


@pytest.mark.parametrize("case", ["reciprocal_products", "divisor_rejected", "ieee_forced"])
def test_spher3d_derivative_divisions(orc, case, monkeypatch, capfd):
    """The five derivative divisions by launch constants run as products
    with the divisors' double reciprocals when the host has shown them to
    round like the IEEE quotient for every float (dr = 5 m: 2 dr = 10 has
    subnormal midpoint quotients, all rounded alike); dr = 49 m gives
    2 dr = 98, for which some subnormal quotient rounds the other way, so
    the kernel divides; AWEFD_SPHER_IEEEDIV=1 asks for the division.  Every
    case is bit-exact against the oracle."""
    monkeypatch.setenv("AWEFD_VERBOSE", "1")
    if case == "ieee_forced":
        monkeypatch.setenv("AWEFD_SPHER_IEEEDIV", "1")
    par, vel, sou, rec, wav = _case(3, nt=60)
    if case == "divisor_rejected":
        d = list(par["d"])
        d[0] = np.float32(49.0)
        par["d"] = d
        sou = _coords(par["o"], d, [[30 - 6.5, 14.4, 12.3]])
        rec = _coords(par["o"], d, [[30 - 8.6, t, 13.3] for t in np.linspace(1.0, 26.0, 9)])
    _check("spher3d", par, vel, sou, rec, wav, orc)
    err = capfd.readouterr().err
    want = "as double reciprocal products" if case == "reciprocal_products" else "by IEEE float division"
    assert want in err, err
