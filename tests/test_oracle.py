"""Pins the CPU oracle (CPU only; no GPU).

The reference ships no tests or golden outputs for this path, so the oracle is
pinned by what the mathematics fixes, each chosen so that a plausible slip
(a dropped term, a wrong sign or index, a transposed operand) fails:

* FD coefficients: the closed form and the Taylor moment conditions.
* Sponge tables: the closed forms and the outermost factors quoted for the
  reference defaults.
* Index map: hand-computed on-node, mid-cell, fractional and out-of-grid
  cases (tests/golden/known_answers.json).
* Stencils and time ordering: brute force against a textbook numpy leapfrog
  (float64, written from the PDE, not from the oracle) on tiny grids before the
  wave reaches the absorbing strips, for all four geometries, including the
  spherical metric terms and the reference's polar d/dr sign.
* Whole loop: linearity, superposition, source/receiver reciprocity,
  absorption against a larger domain, and the FP64 error.
"""
import json
import os
from fractions import Fraction
from math import factorial

import numpy as np
import pytest

from wavefield_imaging_amd import synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _known():
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        return json.load(f)


# --------------------------------------------------------------------------
# FD coefficients (MAWEFDcart2Dcpu.c:423-551)

def _exact_d2(nop):
    """Central second-derivative weights of order 2*nop (exact rationals)."""
    return [Fraction(2 * (-1) ** (k + 1) * factorial(nop) ** 2,
                     k * k * factorial(nop - k) * factorial(nop + k)) for k in range(1, nop + 1)]


@pytest.mark.parametrize("nop", [1, 2, 3, 4, 5, 6])
def test_fd_second_derivative_closed_form_and_moments(orc, nop):
    d = [5.0, 4.0, 2.5]
    c = orc.fdcoef(nop, d)
    ex = _exact_d2(nop)
    # Taylor: sum c_k k^2 = 1 and sum c_k k^(2m) = 0 for m = 2..nop
    assert sum(e * k * k for k, e in enumerate(ex, 1)) == 1
    for m in range(2, nop + 1):
        assert sum(e * k ** (2 * m) for k, e in enumerate(ex, 1)) == 0
    for a, h in enumerate(d):
        for k in range(1, nop + 1):
            np.testing.assert_allclose(c[a * nop + k], float(ex[k - 1]) / h ** 2, rtol=2e-7)
    # c0 = -2 * sum of all axis weights (the centre weight of the Laplacian)
    c0 = -2.0 * sum(float(ex[k]) / h ** 2 for h in d for k in range(nop))
    np.testing.assert_allclose(c[0], c0, rtol=1e-6)
    if nop == 2:
        np.testing.assert_allclose(c[1] * 25.0, 4.0 / 3.0, rtol=1e-7)
        np.testing.assert_allclose(c[2] * 25.0, -1.0 / 12.0, rtol=1e-7)


@pytest.mark.parametrize("nop", [1, 2, 3, 4, 5, 6])
def test_fd_first_derivative_moments(orc, nop):
    b = orc.fdcoef(nop, [2.0, 2.0], order=1)
    ex = [e * k / 2 for k, e in enumerate(_exact_d2(nop), 1)]
    # sum b_k 2k = 1, sum b_k k^(2m+1) = 0 for m = 1..nop-1
    assert sum(e * 2 * k for k, e in enumerate(ex, 1)) == 1
    for m in range(1, nop):
        assert sum(e * k ** (2 * m + 1) for k, e in enumerate(ex, 1)) == 0
    np.testing.assert_allclose(b[1:nop + 1] * 2.0, [float(e) for e in ex], rtol=2e-7)
    assert b[0] == 0


def _band(resp, target, tol):
    """Upper end (fraction of pi) of the wavenumber band [0, k] in which the
    spectral response stays within ``tol`` relative error."""
    th = np.linspace(0.01, np.pi, 4000)
    err = np.abs(resp(th) / target(th) - 1)
    return th[np.argmax(err > tol)] / np.pi if (err > tol).any() else 1.0


@pytest.mark.parametrize("nop", [2, 3, 4, 5, 6])
def test_fd_optimized_tables_widen_the_accurate_band(orc, nop):
    """optfd tables (Liu 2013; Zhang & Yao 2013) trade the Taylor moments for
    a wider wavenumber band: at 1 % (and 0.3 %) error their band must exceed
    the Taylor band of the same length, for both derivatives."""
    c = orc.fdcoef(nop, [1.0, 1.0], optfd=True)[1:nop + 1].astype(np.float64)
    t = _taylor_d2(nop)
    d2 = lambda w: (lambda th: sum(2 * w[k - 1] * (1 - np.cos(k * th)) for k in range(1, nop + 1)))
    for tol in (1e-2, 3e-3):
        assert _band(d2(c), np.square, tol) > _band(d2(t), np.square, tol) + 0.03
    b = orc.fdcoef(nop, [1.0, 1.0], optfd=True, order=1)[1:nop + 1].astype(np.float64)
    tb = t * np.arange(1, nop + 1) / 2
    d1 = lambda w: (lambda th: sum(2 * w[k - 1] * np.sin(k * th) for k in range(1, nop + 1)))
    ident = lambda th: th
    assert _band(d1(b), ident, 1e-2) > _band(d1(tb), ident, 1e-2) + 0.03
    assert abs(sum(c[k - 1] * k * k for k in range(1, nop + 1)) - 1) < 0.02


# --------------------------------------------------------------------------
# Sponge tables (fdutil.c:2181-2203; spher_utils.c:665-667)

@pytest.mark.parametrize("nb", [1, 2, 7, 40, 50, 300])
def test_cartesian_sponge_table(orc, nb):
    w = orc.sponge_table(nb)
    ib = np.arange(nb)
    closed = np.exp(-(ib / (np.sqrt(2.0) * 4.0 * nb)) ** 2)
    np.testing.assert_allclose(w, closed, rtol=1e-6)
    assert w[0] == 1.0 and np.all(np.diff(w) <= 0)
    if nb == 50:  # outermost factor of the 2-D default (applied as w[nb-1])
        assert abs(w[nb - 1] - 0.97044) < 1e-4


def test_spherical_damp(orc):
    for nb in (4, 30):
        for i in range(nb + 2):
            closed = np.exp(-((i - 1) ** 2) * abs(np.log(float(np.float32(0.9)))) / nb ** 2)
            assert orc.spher_damp(i, nb) == pytest.approx(closed, rel=1e-14)
        assert orc.spher_damp(1, nb) == 1.0
        assert orc.spher_damp(nb + 1, nb) == pytest.approx(0.9, rel=1e-7)
    assert orc.spher_damp(30, 30) == pytest.approx(0.9062, abs=1e-4)


# --------------------------------------------------------------------------
# Index map (fdutil.c:1363-1411, 1474-1539)

def test_lint2d_known_answers(orc):
    for case in _known()["lint2d"]:
        jz, jx, w = orc.lint2d(np.float32(case["xz"]), *case["grid"])
        assert (int(jz[0]), int(jx[0])) == tuple(case["j"]), case
        np.testing.assert_allclose(w[0], case["w"], atol=1e-6, err_msg=str(case))


def test_lint3d_known_answers(orc):
    for case in _known()["lint3d"]:
        jz, jx, jy, w = orc.lint3d(np.float32(case["xyz"]), *case["grid"])
        assert (int(jz[0]), int(jx[0]), int(jy[0])) == tuple(case["j"]), case
        np.testing.assert_allclose(w[0], case["w"], atol=1e-6, err_msg=str(case))


def test_lint_weights_partition_unity(orc):
    rng = np.random.default_rng(3)
    pts = rng.uniform(-40, 140, size=(500, 3)).astype(np.float32)
    _, _, _, w = orc.lint3d(pts, 30, 30, 30, -40.0, -40.0, -40.0, 5.0, 5.0, 5.0)
    np.testing.assert_allclose(w.sum(1), 1.0, atol=1e-6)
    assert np.all(w >= -1e-7)


# --------------------------------------------------------------------------
# Textbook brute force (float64), written from the PDE

def _taylor_d2(nop):
    return np.array([float(e) for e in _exact_d2(nop)])


def _lap_cart(u, h, nop, ndim):
    """Sum over axes of the order-2*nop second difference, interior nodes
    only (distance >= nop from every face); array axes (.., x, z)."""
    c = _taylor_d2(nop)
    lap = np.zeros_like(u)
    inner = tuple(slice(nop, n - nop) for n in u.shape)
    for ax in range(ndim):
        acc = -2.0 * c.sum() * u
        for k in range(1, nop + 1):
            acc = acc + c[k - 1] * (np.roll(u, k, ax) + np.roll(u, -k, ax))
        lap[inner] += (acc / h[ndim - 1 - ax] ** 2)[inner]
    return lap


def _points(orc, kind, n, o, d, nb, pts):
    """Padded-array nodes and weights of each point, from the index map that
    the known-answer tests pin: [[(array index, weight), ...], ...].
    Array axes are slowest first: cartesian (y, x, z), spherical (phi, theta, r)."""
    f32 = np.float32
    pad = lambda a: f32(o[a]) - f32(nb) * f32(d[a])
    npd = [k + 2 * nb for k in n]
    pts = np.asarray(pts, np.float32)
    out = []
    if kind == "cart2d":
        jz, jx, w = orc.lint2d(pts, npd[0], npd[1], pad(0), pad(1), d[0], d[1])
        for i in range(len(pts)):
            out.append([((jx[i] + (q >> 1), jz[i] + (q & 1)), w[i, q]) for q in range(4)])
    elif kind == "cart3d":
        jz, jx, jy, w = orc.lint3d(pts, npd[0], npd[1], npd[2], pad(0), pad(1), pad(2),
                                   d[0], d[1], d[2])
        for i in range(len(pts)):
            out.append([((jy[i] + (q >> 2), jx[i] + ((q >> 1) & 1), jz[i] + (q & 1)), w[i, q])
                        for q in range(8)])
    elif kind == "spher2d":       # points (r, theta): z = theta, x = r
        jz, jx, w = orc.lint2d(pts, npd[1], npd[0], pad(1), pad(0), d[1], d[0])
        for i in range(len(pts)):
            out.append([((jz[i] + (q & 1), jx[i] + (q >> 1)), w[i, q]) for q in range(4)])
    else:                         # points (r, phi, theta): z = theta, x = r, y = phi
        jz, jx, jy, w = orc.lint3d(pts, npd[1], npd[0], npd[2], pad(1), pad(0), pad(2),
                                   d[1], d[0], d[2])
        for i in range(len(pts)):
            out.append([((jy[i] + (q >> 2), jz[i] + (q & 1), jx[i] + ((q >> 1) & 1)), w[i, q])
                        for q in range(8)])
    return out


def _gather(u, pt):
    return sum(float(w) * u[idx] for idx, w in pt)


def _gauss_taper(n, nb):
    """Per-axis sponge profile: layer ib from either face is scaled by
    exp(-((nb-1-ib) / (4 sqrt(2) nb))^2) (the outermost layer the most)."""
    f = np.ones(n)
    for ib in range(nb):
        g = np.exp(-((nb - 1 - ib) / (np.sqrt(2.0) * 4.0 * nb)) ** 2)
        f[ib] *= g
        f[n - 1 - ib] *= g
    return f


def _sponge(u, nb):
    for ax in range(u.ndim):
        shape = [1] * u.ndim
        shape[ax] = u.shape[ax]
        u *= _gauss_taper(u.shape[ax], nb).reshape(shape)


def _one_way_faces(un, u, velp, dt, h, nop, nb, free):
    """First-order one-way (Higdon) condition on the outer nop layers of
    every face of a 3-D grid, swept outward from layer nop:
        un[b] = u[b+1] + (u[b] - un[b+1]) (1-c)/(1+c),  c = v dt / h
    with b+1 the inward neighbour and v the raw velocity nb layers in.
    Faces in the order z (low face skipped under a free surface), x, y.
    Array axes (y, x, z); h = (dz, dx, dy)."""
    for ax, hh in ((2, h[0]), (1, h[1]), (0, h[2])):
        n = un.shape[ax]
        take = lambda a, i: np.take(a, i, axis=ax)

        def put(i, val):
            idx = [slice(None)] * 3
            idx[ax] = i
            un[tuple(idx)] = val
        c_lo = take(velp, nb) * dt / hh
        c_hi = take(velp, n - nb - 1) * dt / hh
        k_lo, k_hi = (1 - c_lo) / (1 + c_lo), (1 - c_hi) / (1 + c_hi)
        for k in range(nop):
            if not (free and ax == 2):
                i = nop - k
                put(i, take(u, i + 1) + (take(u, i) - take(un, i + 1)) * k_lo)
            i = n - nop + k - 1
            put(i, take(u, i - 1) + (take(u, i) - take(un, i - 1)) * k_hi)


def _grad_dot(u, rho, h, nop, ndim):
    """sum over axes of D1(u) D1(rho) with the order-2*nop central first
    difference, interior nodes only; array axes (.., x, z)."""
    b = _taylor_d2(nop) * np.arange(1, nop + 1) / 2
    out = np.zeros_like(u)
    inner = tuple(slice(nop, n - nop) for n in u.shape)
    for ax in range(ndim):
        du = sum(b[k - 1] * (np.roll(u, -k, ax) - np.roll(u, k, ax)) for k in range(1, nop + 1))
        dr = sum(b[k - 1] * (np.roll(rho, -k, ax) - np.roll(rho, k, ax)) for k in range(1, nop + 1))
        out[inner] += (du * dr)[inner] / h[ndim - 1 - ax] ** 2
    return out


def _leapfrog_cart(velp, dt, h, nop, src, recs, wav, nt, nb=0, dabc=False, free=False, rho=None):
    """u[n+1] = 2u[n] - u[n-1] + (v dt)^2 lap u[n]; the source adds
    w[n] W (v dt)^2 at each of its nodes; with dabc the 3-D one-way faces
    follow, then the sponge scales u[n+1] and u[n]; trace row n = u[n]
    (u[0] = 0).  A free surface zeroes (v dt)^2 on the top nb+1 layers.
    Returns the traces and the last recorded field u[nt-1]."""
    ndim = velp.ndim
    V = (velp.astype(np.float64) * dt) ** 2
    if free:
        V[..., :nb + 1] = 0
    um = np.zeros_like(V)
    u = np.zeros_like(V)
    tr = np.zeros((nt, len(recs)))
    for it in range(nt):
        lap = _lap_cart(u, h, nop, ndim)
        if rho is not None:   # variable density: lap - grad(rho).grad(u) / rho
            lap = lap - _grad_dot(u, rho, h, nop, ndim) / rho
        un = 2 * u - um + V * lap
        for idx, w in src:
            un[idx] += float(wav[it]) * float(w) * V[idx]
        if dabc:
            if ndim == 3:
                _one_way_faces(un, u, velp.astype(np.float64), dt, h, nop, nb, free)
            _sponge(un, nb)
            _sponge(u, nb)
        tr[it] = [_gather(u, r) for r in recs]
        um, u = u, un
    return tr, um


@pytest.mark.parametrize("ndim,fdorder,nt", [(2, 4, 9), (2, 8, 5), (3, 4, 7), (3, 6, 5)])
def test_cartesian_against_textbook_leapfrog(orc, ndim, fdorder, nt):
    nop = fdorder // 2
    nb = 6
    n = [30, 34, 28][:ndim]
    rng = np.random.default_rng(7)
    vel = rng.uniform(1.2e8, 3.0e8, size=n[::-1]).astype(np.float32)
    h = [5.0, 4.0, 6.0][:ndim]
    # (z, x, y) positions in cells: an on-node source, on-node and off-node receivers
    si = [15, 17, 14][:ndim]
    rz = [si, [si[0] + 1] + si[1:], [si[0]] + [si[1] + 2] + si[2:],
          [si[0] - 3] + [si[1] + 1] + si[2:], [si[0] + 0.5, si[1] - 0.25] + [s + 0.75 for s in si[2:]]]
    co = lambda z: ([z[1] * h[1], z[0] * h[0]] if ndim == 2 else
                    [z[1] * h[1], z[2] * h[2], z[0] * h[0]])
    sou = np.array([co(si)], np.float32)
    rec = np.array([co(r) for r in rz], np.float32)
    wav = synth.ricker(nt, synth.DT, freq=2e7, delay=0.0)
    kind = "cart%dd" % ndim
    par = dict(n=n, o=[0.0] * ndim, d=h, nt=nt, dt=synth.DT, nb=nb, dabc=True, fdorder=fdorder)
    ref32, _, _ = orc.run(kind, par, vel, sou, rec, wav)
    ref64, _, _ = orc.run(kind, par, vel, sou, rec, wav, precision="f64")
    pt = lambda q: _points(orc, kind, n, par["o"], h, nb, q)
    s_pts, r_pts = pt(sou), pt(rec)
    assert max(w for _, w in s_pts[0]) == 1.0                  # on-node source
    tb, _ = _leapfrog_cart(np.pad(vel, nb, mode="edge"), np.float64(np.float32(synth.DT)), h, nop,
                           s_pts[0], r_pts, wav.astype(np.float64), nt)
    # time ordering: row 0 is u[0] = 0; row 1 holds only the injected node
    assert np.all(ref32[0] == 0) and ref32[1, 0] != 0 and np.all(ref32[1, 1:4] == 0)
    assert np.abs(tb[-1]).min() > 0
    np.testing.assert_allclose(ref64, tb, rtol=1e-9, atol=1e-12 * np.abs(tb).max())
    np.testing.assert_allclose(ref32, tb, rtol=2e-5, atol=2e-6 * np.abs(tb).max())


@pytest.mark.parametrize("ndim,fdorder,sse", [(2, 4, True), (2, 8, False), (3, 4, True), (3, 6, False)])
def test_cartesian_density_against_textbook(orc, ndim, fdorder, sse):
    """cden=n: the operator is lap(u) - grad(rho).grad(u) / rho (ts_kernel.c:101-135,
    228-280), with the first differences of the same order."""
    nop, nb, nt = fdorder // 2, 6, 7
    n = [30, 34, 28][:ndim]
    rng = np.random.default_rng(17)
    vel = rng.uniform(1.2e8, 3.0e8, size=n[::-1]).astype(np.float32)
    den = rng.uniform(900.0, 2600.0, size=n[::-1]).astype(np.float32)
    h = [5.0, 4.0, 6.0][:ndim]
    si = [15, 17, 14][:ndim]
    co = lambda z: ([z[1] * h[1], z[0] * h[0]] if ndim == 2 else
                    [z[1] * h[1], z[2] * h[2], z[0] * h[0]])
    rz = [si, [si[0] + 2] + si[1:], [si[0] - 1, si[1] + 1.5] + [s + 0.5 for s in si[2:]]]
    sou = np.array([co(si)], np.float32)
    rec = np.array([co(r) for r in rz], np.float32)
    wav = synth.ricker(nt, synth.DT, freq=2e7, delay=0.0)
    kind = "cart%dd" % ndim
    par = dict(n=n, o=[0.0] * ndim, d=h, nt=nt, dt=synth.DT, nb=nb, dabc=True, fdorder=fdorder,
               cden=False, sse=sse)
    ref32, _, _ = orc.run(kind, par, vel, sou, rec, wav, den=den)
    ref64, _, _ = orc.run(kind, par, vel, sou, rec, wav, den=den, precision="f64")
    pt = lambda q: _points(orc, kind, n, par["o"], h, nb, q)
    rhop = np.pad(den, nb, mode="edge").astype(np.float64)
    tb, _ = _leapfrog_cart(np.pad(vel, nb, mode="edge"), np.float64(np.float32(synth.DT)), h, nop,
                           pt(sou)[0], pt(rec), wav.astype(np.float64), nt, rho=rhop)
    tb0, _ = _leapfrog_cart(np.pad(vel, nb, mode="edge"), np.float64(np.float32(synth.DT)), h, nop,
                            pt(sou)[0], pt(rec), wav.astype(np.float64), nt)
    assert np.abs(tb - tb0).max() > 1e-3 * np.abs(tb).max()      # the term matters
    np.testing.assert_allclose(ref64, tb, rtol=1e-9, atol=1e-12 * np.abs(tb).max())
    np.testing.assert_allclose(ref32, tb, rtol=2e-5, atol=2e-6 * np.abs(tb).max())


@pytest.mark.parametrize("ndim,free", [(2, False), (2, True), (3, False), (3, True)])
def test_cartesian_boundaries_against_textbook(orc, ndim, free):
    """A source next to a grid corner, so the wave crosses the sponge strips
    (and in 3-D the one-way faces) within a few steps: whole final field and
    traces against the textbook sweep."""
    nb, nt, fdorder = 5, 24, 4
    n = [10, 12, 11][:ndim]
    rng = np.random.default_rng(12)
    vel = rng.uniform(1.2e8, 3.0e8, size=n[::-1]).astype(np.float32)
    h = [5.0, 4.0, 6.0][:ndim]
    si = [1.5, 1.25, 9.5][:ndim]
    rz = [[0.0, 0.0, 0.0][:ndim], [9.0, 11.0, 10.0][:ndim], [1.0, 6.0, 9.75][:ndim]]
    co = lambda z: ([z[1] * h[1], z[0] * h[0]] if ndim == 2 else
                    [z[1] * h[1], z[2] * h[2], z[0] * h[0]])
    sou = np.array([co(si)], np.float32)
    rec = np.array([co(r) for r in rz], np.float32)
    wav = synth.ricker(nt, synth.DT, freq=2e7, delay=0.0)
    kind = "cart%dd" % ndim
    par = dict(n=n, o=[0.0] * ndim, d=h, nt=nt, dt=synth.DT, nb=nb, dabc=True,
               fdorder=fdorder, free=free)
    ref32, f32, _ = orc.run(kind, par, vel, sou, rec, wav)
    ref64, f64, _ = orc.run(kind, par, vel, sou, rec, wav, precision="f64")
    pt = lambda q: _points(orc, kind, n, par["o"], h, nb, q)
    tb, ftb = _leapfrog_cart(np.pad(vel, nb, mode="edge"), np.float64(np.float32(synth.DT)), h,
                             fdorder // 2, pt(sou)[0], pt(rec), wav.astype(np.float64), nt,
                             nb=nb, dabc=True, free=free)
    assert np.abs(ftb[..., :nb, :]).max() > 1e-3 * np.abs(ftb).max()   # x-low strip reached
    np.testing.assert_allclose(f64, ftb, rtol=0, atol=1e-10 * np.abs(ftb).max())
    np.testing.assert_allclose(ref64, tb, rtol=0, atol=1e-10 * np.abs(tb).max())
    np.testing.assert_allclose(f32, ftb, rtol=0, atol=1e-5 * np.abs(ftb).max())


def _spher_grid(ndim, n, nb, dr=5.0):
    """r from 3387.5 km (dr = 5 m), d(theta) = d(phi) = 5 m at 3388.5 km,
    theta centred on pi/4."""
    r0 = np.float32(3387.5e3)
    dth = np.float32(5.0 / 3388.5e3)
    o = [r0, np.float32(np.pi / 4 - n[1] / 2 * dth), np.float32(0.3)][:ndim]
    d = [np.float32(dr), dth, dth][:ndim]
    return o, d


def _axis_values(o, d, n, nb):
    """Node coordinates as the reference computes them: float32 padded origin
    o - nb*d, then d*i + o."""
    op = np.float32(o) - np.float32(nb) * np.float32(d)
    return np.float64(np.float32(d)) * np.arange(n + 2 * nb) + np.float64(op)


def _leapfrog_spher(velp, dt, r, th, dr, dth, dph, src, recs, wav, nt, polar_sign=-1.0):
    """Textbook spherical Laplacian with second-order central differences:
    p_rr + (2/r) p_r + p_thth / r^2 + cot(th) p_th / r^2 + p_phph / (r sin th)^2,
    on nodes at least 4 from every face.  Polar (2-D): p_rr + polar_sign (1/r) p_r
    + p_thth / r^2 (the reference program has polar_sign = -1).  The source adds
    w[n] W to p[n] before the update; trace row n = p[n+1]."""
    ndim = velp.ndim
    V = (velp.astype(np.float64) * dt) ** 2
    pm = np.zeros_like(V)
    p = np.zeros_like(V)
    R = r[None, :] if ndim == 2 else r[None, None, :]
    inner = tuple(slice(4, k - 4) for k in V.shape)
    ax_r, ax_t = ndim - 1, ndim - 2
    tr = np.zeros((nt, len(recs)))
    for it in range(nt):
        for idx, w in src:
            p[idx] += float(wav[it]) * float(w)
        d1r = (np.roll(p, -1, ax_r) - np.roll(p, 1, ax_r)) / (2 * dr)
        d2r = (np.roll(p, -1, ax_r) - 2 * p + np.roll(p, 1, ax_r)) / dr ** 2
        d2t = (np.roll(p, -1, ax_t) - 2 * p + np.roll(p, 1, ax_t)) / dth ** 2
        if ndim == 2:
            lap = d2r + polar_sign * d1r / R + d2t / R ** 2
        else:
            TH = th[None, :, None]
            d1t = (np.roll(p, -1, ax_t) - np.roll(p, 1, ax_t)) / (2 * dth)
            d2p = (np.roll(p, -1, 0) - 2 * p + np.roll(p, 1, 0)) / dph ** 2
            lap = (d2r + 2 * d1r / R + d2t / R ** 2 + np.cos(TH) / np.sin(TH) * d1t / R ** 2
                   + d2p / (R * np.sin(TH)) ** 2)
        lap_in = np.zeros_like(lap)
        lap_in[inner] = lap[inner]
        pm, p = p, V * lap_in + 2 * p - pm
        tr[it] = [_gather(p, q) for q in recs]
    return tr


@pytest.mark.parametrize("ndim,nt", [(2, 12), (3, 9)])
def test_spherical_against_textbook_leapfrog(orc, ndim, nt):
    nb = 6
    n = [26, 24, 22][:ndim]
    o, d = _spher_grid(ndim, n, nb)
    rng = np.random.default_rng(9)
    vel = rng.uniform(1.2e8, 3.0e8, size=n[::-1]).astype(np.float32)
    # (r, theta, phi) positions in cells
    si = [13.25, 12.5, 11.75][:ndim]
    rq = [si, [si[0] + 1] + si[1:], [si[0] - 1.5] + si[1:], [si[0], si[1] + 1] + si[2:],
          [si[0], si[1] - 2.25] + si[2:]]
    if ndim == 3:
        rq.append([si[0], si[1], si[2] + 1.5])
    # coordinate files: spher2d (r, theta); spher3d (r, phi, theta)
    co = lambda q: ([o[0] + q[0] * d[0], o[1] + q[1] * d[1]] if ndim == 2 else
                    [o[0] + q[0] * d[0], o[2] + q[2] * d[2], o[1] + q[1] * d[1]])
    sou = np.array([co(si)], np.float32)
    rec = np.array([co(q) for q in rq], np.float32)
    wav = synth.ricker(nt, synth.DT, freq=2e7, delay=0.0)
    kind = "spher%dd" % ndim
    par = dict(n=n, o=o, d=d, nt=nt, dt=synth.DT, nb=nb, expl=False)
    ref32, _, _ = orc.run(kind, par, vel, sou, rec, wav)
    ref64, _, _ = orc.run(kind, par, vel, sou, rec, wav, precision="f64")
    pt = lambda q: _points(orc, kind, n, o, d, nb, q)
    s_pts, r_pts = pt(sou), pt(rec)
    r = _axis_values(o[0], d[0], n[0], nb)
    th = _axis_values(o[1], d[1], n[1], nb)
    args = (np.pad(vel, nb, mode="edge"), np.float64(np.float32(synth.DT)), r, th,
            float(d[0]), float(d[1]), float(d[2]) if ndim == 3 else 1.0, s_pts[0], r_pts,
            wav.astype(np.float64), nt)
    tb = _leapfrog_spher(*args)
    assert ref32[0, 0] != 0                         # injection precedes the solve
    np.testing.assert_allclose(ref64, tb, rtol=1e-9, atol=1e-12 * np.abs(tb).max())
    np.testing.assert_allclose(ref32, tb, rtol=1e-4, atol=1e-5 * np.abs(tb).max())
    if ndim == 2:
        # the reference's polar operator is d_rr - (1/r) d_r; the textbook sign
        # (+) changes the radial neighbours by far more than the oracle's error
        tb_plus = _leapfrog_spher(*args, polar_sign=+1.0)
        assert np.abs(tb_plus - ref64).max() > 1e3 * np.abs(tb - ref64).max()


def _polar_last_rule(ir, ith, nr, nth):
    """Last one-way rule applied at a polar node (spher_utils.c:518-566):
    iop from the first of r, theta within NOP = 4 of a face; then theta-low (1),
    theta-high (2), r-low (3), r-high (4) on exact layer matches."""
    if ir < 4:
        iop = ir
    elif ir > nr - 4:
        iop = nr - ir
    elif ith < 4:
        iop = ith
    elif ith > nth - 4:
        iop = nth - ith
    else:
        iop = 0
    rule = 0
    if ith == 4 - iop:
        rule = 1
    if ith == nth - 4 + iop - 1:
        rule = 2
    if ir == 4 - iop:
        rule = 3
    if ir == nr - 4 + iop - 1:
        rule = 4
    return rule


def _k32(v, dt, h):
    """(1-c)/(1+c) with c = v (dt/h), evaluated in float32."""
    c = np.float32(v) * (np.float32(dt) / np.float32(h))
    return np.float32((np.float32(1) - c) / (np.float32(1) + c))


def _spher_boundaries(orc, velp, dt, d, nb):
    """Serial one-way update list and per-node sponge factors.

    One-way: p_new[a] = p_old[n] + (p_old[a] - p_new[n]) k with n the inward
    neighbour of the rule and k = (1-c)/(1+c), c = v dt / d(axis) from the raw
    velocity at layer NOP = 4 in from the face the table is named for.  The
    r-low rule takes the table built at the r-high layer and the r-high rule the
    one built at the r-low layer (the program passes the r tables swapped); the
    phi-high table is built at phi = n_theta - 5 (the program indexes that layer
    with the theta size).  Nodes are visited r-outermost, then theta, then phi.
    Sponge: one factor per node, exp(-(i-1)^2 |ln 0.9f| / nb^2), i the depth
    into the first strip found in the order r, phi, theta (low side first)."""
    ndim = velp.ndim
    upd = []
    if ndim == 2:
        nth, nr = velp.shape
        for ir in range(nr):
            for ith in range(nth):
                rule = _polar_last_rule(ir, ith, nr, nth)
                if rule == 1:
                    upd.append(((ith, ir), (ith + 1, ir), _k32(velp[4, ir], dt, d[1])))
                elif rule == 2:
                    upd.append(((ith, ir), (ith - 1, ir), _k32(velp[nth - 5, ir], dt, d[1])))
                elif rule == 3:
                    upd.append(((ith, ir), (ith, ir + 1), _k32(velp[ith, 4], dt, d[0])))
                elif rule == 4:
                    upd.append(((ith, ir), (ith, ir - 1), _k32(velp[ith, nr - 5], dt, d[0])))
    else:
        nph, nth, nr = velp.shape
        for ir in range(nr):
            for ith in range(nth):
                for iph in range(nph):
                    rule = orc.spher3d_rule(ir, ith, iph, nr, nth, nph)
                    a = (iph, ith, ir)
                    if rule == 1:
                        upd.append((a, (iph, ith + 1, ir), _k32(velp[iph, 4, ir], dt, d[1])))
                    elif rule == 2:
                        upd.append((a, (iph, ith - 1, ir), _k32(velp[iph, nth - 5, ir], dt, d[1])))
                    elif rule == 3:
                        upd.append((a, (iph, ith, ir + 1), _k32(velp[iph, ith, nr - 5], dt, d[0])))
                    elif rule == 4:
                        upd.append((a, (iph, ith, ir - 1), _k32(velp[iph, ith, 4], dt, d[0])))
                    elif rule == 5:
                        upd.append((a, (iph + 1, ith, ir), _k32(velp[4, ith, ir], dt, d[2])))
                    elif rule == 6:
                        upd.append((a, (iph - 1, ith, ir), _k32(velp[nth - 5, ith, ir], dt, d[2])))
    fac = np.ones(velp.shape)
    damp = lambda i: np.exp(-((i - 1) ** 2) * abs(np.log(float(np.float32(0.9)))) / nb ** 2)
    for idx in np.ndindex(*velp.shape):
        if ndim == 2:
            ith, ir = idx
            axes = [(ir, velp.shape[1]), (ith, velp.shape[0])]
        else:
            iph, ith, ir = idx
            axes = [(ir, velp.shape[2]), (iph, velp.shape[0]), (ith, velp.shape[1])]
        lo = [nb - i for i, n in axes if i < nb]
        hi = [i - (n - nb) for i, n in axes if i > n - nb]
        if lo:
            fac[idx] = damp(lo[0])
        elif hi:
            fac[idx] = damp(hi[0])
    return upd, fac


@pytest.mark.parametrize("ndim,corner,free", [(2, "low", False), (2, "high", False),
                                              (2, "high", True), (3, "low", False),
                                              (3, "high", False)])
def test_spherical_boundaries_against_textbook(orc, ndim, corner, free):
    """A source next to a corner of a small spherical grid: the wave crosses
    the one-way layers and the sponge within the run; traces and the final
    field against the textbook sweep.  The polar free surface zeroes the
    outermost nb-1 radial layers after the sponge."""
    nb, nt = 7, 20
    # n_theta < n_phi - 3 keeps the phi-high table's layer (n_theta + 2 here)
    # apart from the edge-replicated padding, so that layer choice is visible
    n = [10, 9] if ndim == 2 else [10, 8, 12]
    o, d = _spher_grid(ndim, n, nb)
    rng = np.random.default_rng(21)
    vel = rng.uniform(1.2e8, 3.0e8, size=n[::-1]).astype(np.float32)
    si = ([1.5, 1.25, 1.5] if corner == "low" else [8.5, 6.75, 10.5])[:ndim]
    rq = [[0.0, 0.0, 0.0][:ndim], [9.0, 8.0, 7.0][:ndim], [1.0, 6.0, 6.75][:ndim]]
    co = lambda q: ([o[0] + q[0] * d[0], o[1] + q[1] * d[1]] if ndim == 2 else
                    [o[0] + q[0] * d[0], o[2] + q[2] * d[2], o[1] + q[1] * d[1]])
    sou = np.array([co(si)], np.float32)
    rec = np.array([co(q) for q in rq], np.float32)
    wav = synth.ricker(nt, synth.DT, freq=2e7, delay=0.0)
    kind = "spher%dd" % ndim
    par = dict(n=n, o=o, d=d, nt=nt, dt=synth.DT, nb=nb, expl=False, free=free)
    ref64, f64, _ = orc.run(kind, par, vel, sou, rec, wav, precision="f64")
    ref32, f32, _ = orc.run(kind, par, vel, sou, rec, wav)
    velp = np.pad(vel, nb, mode="edge")
    dt = np.float64(np.float32(synth.DT))
    upd, fac = _spher_boundaries(orc, velp, np.float32(synth.DT), d, nb)
    r = _axis_values(o[0], d[0], n[0], nb)
    th = _axis_values(o[1], d[1], n[1], nb)
    src = _points(orc, kind, n, o, d, nb, sou)[0]
    recs = _points(orc, kind, n, o, d, nb, rec)
    V = (velp.astype(np.float64) * dt) ** 2
    R = r[None, :] if ndim == 2 else r[None, None, :]
    inner = tuple(slice(4, k - 4) for k in V.shape)
    pm = np.zeros_like(V)
    p = np.zeros_like(V)
    tr = np.zeros((nt, len(recs)))
    ax_r, ax_t = ndim - 1, ndim - 2
    for it in range(nt):
        for idx, w in src:
            p[idx] += float(wav[it]) * float(w)
        d1r = (np.roll(p, -1, ax_r) - np.roll(p, 1, ax_r)) / (2 * float(d[0]))
        d2r = (np.roll(p, -1, ax_r) - 2 * p + np.roll(p, 1, ax_r)) / float(d[0]) ** 2
        d2t = (np.roll(p, -1, ax_t) - 2 * p + np.roll(p, 1, ax_t)) / float(d[1]) ** 2
        if ndim == 2:
            lap = d2r - d1r / R + d2t / R ** 2
        else:
            TH = th[None, :, None]
            d1t = (np.roll(p, -1, ax_t) - np.roll(p, 1, ax_t)) / (2 * float(d[1]))
            d2p = (np.roll(p, -1, 0) - 2 * p + np.roll(p, 1, 0)) / float(d[2]) ** 2
            lap = (d2r + 2 * d1r / R + d2t / R ** 2 + np.cos(TH) / np.sin(TH) * d1t / R ** 2
                   + d2p / (R * np.sin(TH)) ** 2)
        lap_in = np.zeros_like(lap)
        lap_in[inner] = lap[inner]
        pm, p = p, V * lap_in + 2 * p - pm
        for a, nbr, k in upd:
            p[a] = pm[nbr] + (pm[a] - p[nbr]) * float(k)
        p *= fac
        pm *= fac
        if free:
            p[..., p.shape[-1] - nb + 1:] = 0
        tr[it] = [_gather(p, q) for q in recs]
    # the damped layers and the one-way rows next to the chosen corner are reached
    edge = p[..., 4:6] if corner == "low" else p[..., -7:-4]
    assert np.abs(edge).max() > 1e-3 * np.abs(p).max()
    if ndim == 3:
        ph = p[:5] if corner == "low" else p[-7:]
        assert np.abs(ph).max() > 1e-3 * np.abs(p).max()
    np.testing.assert_allclose(f64, p, rtol=0, atol=1e-9 * np.abs(p).max())
    np.testing.assert_allclose(ref64, tr, rtol=0, atol=1e-9 * np.abs(tr).max())
    np.testing.assert_allclose(f32, p, rtol=0, atol=1e-4 * np.abs(p).max())


# --------------------------------------------------------------------------
# Whole-loop invariants

def _small2d(nz=60, nx=70, nt=260, nb=20, **kw):
    vel = synth.layered_2d(nz, nx, seed=4, air=20, base_mean=40, amp=3.0, corr=6.0)
    sou = np.array([[nx / 2 * 5.0, 25 * 5.0]], np.float32)
    rec = np.stack([np.arange(0, nx, 3) * 5.0, np.full(len(range(0, nx, 3)), 22 * 5.0)],
                   1).astype(np.float32)
    par = dict(n=[nz, nx], o=[0.0, 0.0], d=[5.0, 5.0], nt=nt, dt=synth.DT, nb=nb, dabc=True)
    par.update(kw)
    return par, vel, sou, rec, synth.ricker(nt, synth.DT)


def test_linearity(orc):
    """Doubling the wavelet doubles every output.  Exact in FP64; in FP32 the
    gradual underflow at the wave's leading edge (the Ricker tail is ~1e-8 at
    t = 0 and the first values far from the source are subnormal) breaks
    bit-exact homogeneity at the 1e-6 level of the peak."""
    par, vel, sou, rec, wav = _small2d(nt=400)
    a, fa, _ = orc.run("cart2d", par, vel, sou, rec, wav, precision="f64")
    b, fb, _ = orc.run("cart2d", par, vel, sou, rec, 2 * wav, precision="f64")
    np.testing.assert_array_equal(b, 2 * a)
    np.testing.assert_array_equal(fb, 2 * fa)
    a, fa, _ = orc.run("cart2d", par, vel, sou, rec, wav)
    b, fb, _ = orc.run("cart2d", par, vel, sou, rec, 2 * wav)
    np.testing.assert_allclose(b, 2 * a, rtol=0, atol=1e-5 * np.abs(a).max())
    np.testing.assert_allclose(fb, 2 * fa, rtol=0, atol=1e-5 * np.abs(fa).max())


def test_superposition(orc):
    par, vel, sou, rec, wav = _small2d(nt=150)
    s2 = np.array([[100.0, 90.0], [230.0, 140.0]], np.float32)
    w2 = np.stack([wav, -0.7 * wav], 1).astype(np.float32)
    both, _, _ = orc.run("cart2d", par, vel, s2, rec, w2)
    a, _, _ = orc.run("cart2d", par, vel, s2[:1], rec, wav)
    b, _, _ = orc.run("cart2d", par, vel, s2[1:], rec, (-0.7 * wav).astype(np.float32))
    np.testing.assert_allclose(both, a + b, atol=1e-6 * np.abs(both).max())


@pytest.mark.parametrize("kind", ["cart2d", "cart3d"])
def test_reciprocity(orc, kind):
    """Without absorbing strips the discrete propagator P(VL)V is symmetric,
    so swapping an on-node source and receiver leaves the trace unchanged."""
    if kind == "cart2d":
        n = [40, 44]
        vel = synth.layered_2d(40, 44, seed=2, air=10, base_mean=25, amp=3.0, corr=5.0)
        a, b = [60.0, 70.0], [155.0, 120.0]
    else:
        n = [24, 26, 22]
        vel = synth.layered_3d(24, 26, 22, seed=2, air=6, base_mean=15, amp=2.0, corr=4.0)
        a, b = [40.0, 50.0, 35.0], [85.0, 60.0, 75.0]
    nt = 150 if kind == "cart2d" else 110
    par = dict(n=n, o=[0.0] * len(n), d=[5.0] * len(n), nt=nt, dt=synth.DT, nb=4, dabc=False)
    wav = synth.ricker(nt, synth.DT, freq=6e6)
    ab, _, _ = orc.run(kind, par, vel, np.float32([a]), np.float32([b]), wav)
    ba, _, _ = orc.run(kind, par, vel, np.float32([b]), np.float32([a]), wav)
    assert np.abs(ab).max() > 0
    np.testing.assert_allclose(ab, ba, atol=2e-5 * np.abs(ab).max())


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_cart2d_sponge_absorbs(orc):
    """Traces of a small domain against the same region inside a domain large
    enough that no boundary echo returns in time: the sponge removes most of
    the echo that the bare (dabc=n) boundary sends back."""
    par, vel, sou, rec, wav = _small2d(nt=500)
    big = dict(par, n=[par["n"][0] + 240, par["n"][1] + 240], o=[-600.0, -600.0])
    velb = np.pad(vel, 120, mode="edge")
    ref, _, _ = orc.run("cart2d", big, velb, sou, rec, wav)
    sp, _, _ = orc.run("cart2d", par, vel, sou, rec, wav)
    bare, _, _ = orc.run("cart2d", dict(par, dabc=False), vel, sou, rec, wav)
    e_sp, e_bare = _rel(sp, ref), _rel(bare, ref)
    assert e_sp < 0.6 * e_bare and e_sp < 0.2, (e_sp, e_bare)


def test_cart3d_abc_absorbs(orc):
    nz, nx, ny, nb, nt = 26, 26, 26, 8, 200
    vel = np.full((ny, nx, nz), 3.0e8, np.float32)
    sou = np.float32([[65.0, 65.0, 65.0]])
    rec = np.float32([[65.0, 65.0, 40.0], [30.0, 65.0, 65.0], [65.0, 100.0, 65.0]])
    wav = synth.ricker(nt, synth.DT, freq=6e6)
    par = dict(n=[nz, nx, ny], o=[0.0] * 3, d=[5.0] * 3, nt=nt, dt=synth.DT, nb=nb, dabc=True)
    big = dict(par, n=[nz + 80, nx + 80, ny + 80], o=[-200.0] * 3)
    ref, _, _ = orc.run("cart3d", big, np.pad(vel, 40, mode="edge"), sou, rec, wav)
    abc, _, _ = orc.run("cart3d", par, vel, sou, rec, wav)
    bare, _, _ = orc.run("cart3d", dict(par, dabc=False), vel, sou, rec, wav)
    e_abc, e_bare = _rel(abc, ref), _rel(bare, ref)
    assert e_abc < 0.25 * e_bare, (e_abc, e_bare)


def test_spher2d_boundaries_absorb(orc):
    nr, nth, nb, nt = 50, 50, 20, 500
    o, d = _spher_grid(2, [nr, nth], nb)
    vel = np.full((nth, nr), 1.68e8, np.float32)
    sou = np.float32([[o[0] + 25 * d[0], o[1] + 25 * d[1]]])
    rec = np.float32([[o[0] + 20 * d[0], o[1] + 25 * d[1]], [o[0] + 25 * d[0], o[1] + 30 * d[1]]])
    wav = synth.ricker(nt, synth.DT)
    par = dict(n=[nr, nth], o=o, d=d, nt=nt, dt=synth.DT, nb=nb, expl=False)
    pad = 150
    big = dict(par, n=[nr + 2 * pad, nth + 2 * pad],
               o=[np.float32(o[0] - pad * d[0]), np.float32(o[1] - pad * d[1])])
    ref, _, _ = orc.run("spher2d", big, np.pad(vel, pad, mode="edge"), sou, rec, wav)
    sm, _, _ = orc.run("spher2d", par, vel, sou, rec, wav)
    assert np.abs(ref).max() > 0
    assert _rel(sm, ref) < 0.2, _rel(sm, ref)


def test_spher3d_free_surface_zeroes_everything(orc):
    """freeBC_3D (spher_utils.c:765-785) zeroes the whole field each step."""
    n, nb = [14, 12, 12], 4
    o, d = _spher_grid(3, n, nb)
    vel = np.full(n[::-1], 1.68e8, np.float32)
    sou = np.float32([[o[0] + 7 * d[0], o[2] + 6 * d[2], o[1] + 6 * d[1]]])
    par = dict(n=n, o=o, d=d, nt=12, dt=synth.DT, nb=nb, expl=False, free=True)
    tr, fin, _ = orc.run("spher3d", par, vel, sou, sou, synth.ricker(12, synth.DT))
    assert not tr.any() and not fin.any()


def test_spher_reflectivity_injection(orc):
    """expl=y scales the source by |v0-vk|/(v0+vk) of the first differing
    neighbour: nothing is injected inside a homogeneous neighbourhood."""
    n, nb = [30, 28], 6
    o, d = _spher_grid(2, n, nb)
    vel = np.full((n[1], n[0]), 1.68e8, np.float32)
    sou = np.float32([[o[0] + 15 * d[0], o[1] + 14.5 * d[1]]])
    par = dict(n=n, o=o, d=d, nt=30, dt=synth.DT, nb=nb, expl=True)
    wav = synth.ricker(30, synth.DT)
    tr, _, _ = orc.run("spher2d", par, vel, sou, sou, wav)
    assert not tr.any()
    vel2 = vel.copy()
    vel2[14 + 1:, :] = 1.2e8               # theta + 1 neighbour differs
    tr2, _, _ = orc.run("spher2d", par, vel2, sou, sou, wav)
    r = (1.68e8 - 1.2e8) / (1.68e8 + 1.2e8)
    tr_const, _, _ = orc.run("spher2d", dict(par, expl=False), vel2, sou, sou, wav)
    np.testing.assert_allclose(tr2, r * tr_const, rtol=1e-5, atol=1e-7 * np.abs(tr_const).max())


def _refl_grid(case):
    """Raw padded velocity of a reflectivity known-answer case, array axes
    slowest first ([phi,] theta, r), and the source node (r, theta[, phi])."""
    g, node = case["grid"], case["node"]
    vel = np.full(g[::-1], np.float32(case["v0"]), np.float32)
    for e in case["set"]:
        off, v = e[:-1], e[-1]
        idx = tuple(node[a] + off[a] for a in range(len(node)))[::-1]
        vel[idx] = np.float32(v)
    return vel, node


@pytest.mark.parametrize("key", ["refl2d", "refl3d"])
def test_spher_reflectivity_search_order(orc, key):
    """expl=y injection scale (spher_utils.c:153-173, 250-310): one case per
    position of the reference's neighbour search order, in which that
    position is the first whose velocity differs; the expected value names
    the position.  Covers the 3-D order's repeated (phi+1, r-1) entry and the
    homogeneous neighbourhood (r = 0)."""
    cases = _known()[key]
    assert len(cases) == (8 if key == "refl2d" else 28)
    for case in cases:
        vel, node = _refl_grid(case)
        got = orc.refl_at(vel, node)
        assert got == np.float32(case["r"]), (key, case["case"], case["position"], got, case["r"])


def test_spher3d_rule_known_answers(orc):
    """Last serial one-way rule per node (spher_utils.c:583-631), hand-derived."""
    for case in _known()["spher3d_rule"]:
        assert orc.spher3d_rule(*case["node"], *case["grid"]) == case["rule"], case


@pytest.mark.parametrize("kind", ["cart2d", "cart3d", "spher2d", "spher3d"])
def test_fp64_error_small(orc, kind):
    if kind == "cart2d":
        par, vel, sou, rec, wav = _small2d(nt=300)
    elif kind == "cart3d":
        n = [20, 22, 24]
        vel = synth.layered_3d(*n, seed=1, air=5, base_mean=13, amp=2.0, corr=4.0)
        sou = np.float32([[55.0, 60.0, 32.0]])
        rec = np.float32([[x, 60.0, 30.0] for x in range(10, 100, 10)])
        wav = synth.ricker(150, synth.DT)
        par = dict(n=n, o=[0.0] * 3, d=[5.0] * 3, nt=150, dt=synth.DT, nb=6, dabc=True)
    else:
        ndim = 2 if kind == "spher2d" else 3
        n = [30, 28, 26][:ndim]
        o, d = _spher_grid(ndim, n, 6)
        vel = np.full(n[::-1], 1.68e8, np.float32)
        vel[..., :8] = 3.0e8
        c = ([o[0] + 12.3 * d[0], o[1] + 14.2 * d[1]] if ndim == 2 else
             [o[0] + 12.3 * d[0], o[2] + 13.1 * d[2], o[1] + 14.2 * d[1]])
        sou = np.float32([c])
        rec = np.float32([c])
        wav = synth.ricker(150, synth.DT)
        par = dict(n=n, o=o, d=d, nt=150, dt=synth.DT, nb=6, expl=False)
    a, fa, _ = orc.run(kind, par, vel, sou, rec, wav)
    b, fb, _ = orc.run(kind, par, vel, sou, rec, wav, precision="f64")
    assert np.abs(b).max() > 0
    assert _rel(a, b) < 1e-4 and _rel(fa, fb) < 1e-4, (_rel(a, b), _rel(fa, fb))


@pytest.mark.parametrize("kind,n,nb,kw", [
    ("cart2d", [37, 52], 6, {}),                       # (nz - 2 nop) % 4 != 0: scalar tail
    ("cart2d", [40, 44], 8, dict(fdorder=8, free=True)),
    ("cart3d", [21, 18, 16], 4, {}),
    ("cart3d", [20, 17, 15], 6, dict(fdorder=6, optfd=True)),
    ("cart3d", [19, 16, 16], 4, dict(sse=False)),     # the scalar order keeps the checker's step
])
def test_reference_form_build_is_bit_identical(orc, kind, n, nb, kw):
    """liborc_ref.so (-O3, the 4-wide SSE step of ts_kernel.c with
    schedule(static,1), used only to time the CPU baseline) gives the
    checker's traces and final field bit for bit."""
    ndim = len(n)
    rng = np.random.default_rng(sum(n))
    vel = rng.uniform(1.2e8, 3.0e8, n[::-1]).astype(np.float32)
    h = 5.0
    ext = [(k - 1) * h for k in n]
    sou = np.float32([[ext[1] * 0.5, ext[0] * 0.4] if ndim == 2 else [ext[1] * 0.5, ext[2] * 0.4, ext[0] * 0.3]])
    rec = np.float32([[ext[1] * f, ext[0] * 0.2] if ndim == 2 else [ext[1] * f, ext[2] * 0.6, ext[0] * 0.2]
                      for f in np.linspace(0.05, 0.95, 7)])
    nt = 60
    par = dict(n=n, o=[0.0] * ndim, d=[h] * ndim, nt=nt, dt=synth.DT, nb=nb, dabc=True, **kw)
    wav = synth.ricker(nt, synth.DT, freq=8e6)
    a = orc.run(kind, par, vel, sou, rec, wav)
    b = orc.run(kind, par, vel, sou, rec, wav, ref_form=True)
    assert np.any(a[0])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("ndim,expl", [(2, True), (3, False), (3, True)])
def test_reference_form_spherical_nest_is_bit_identical(orc, ndim, expl):
    """The checker's spherical solve visits nodes r fastest with the trig per
    theta; liborc_ref.so keeps the reference's r-outer nest with the trig per
    node (spher_utils.c:390-398).  Same values, same bits."""
    n, nb = [22, 18, 16][:ndim], 5
    o, d = _spher_grid(ndim, n, nb)
    rng = np.random.default_rng(3 + ndim)
    vel = rng.uniform(1.2e8, 3.0e8, n[::-1]).astype(np.float32)
    c = [11.3, 9.6, 7.2][:ndim]
    sou = np.float32([[o[0] + c[0] * d[0], o[1] + c[1] * d[1]] if ndim == 2 else
                      [o[0] + c[0] * d[0], o[2] + c[2] * d[2], o[1] + c[1] * d[1]]])
    nt = 40
    par = dict(n=n, o=o, d=d, nt=nt, dt=synth.DT, nb=nb, expl=expl)
    wav = synth.ricker(nt, synth.DT, freq=8e6)
    kind = "spher%dd" % ndim
    a = orc.run(kind, par, vel, sou, sou, wav)
    b = orc.run(kind, par, vel, sou, sou, wav, ref_form=True)
    assert np.any(a[1])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
