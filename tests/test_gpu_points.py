"""The host index map and the injection reflectivity of the library, read back
through the C ABI (awefd_debug_points), bit for bit.

- Index map (lint2d_make / lint3d_make, fdutil.c:1363-1411, 1474-1539):
  the known-answer cases of tests/golden/known_answers.json, and the source
  and receiver geometries of the measured workloads (cart2d 400x800 with 800
  receivers, cart3d 256x512x512 with 128x128, spher3d 400x512x512 with
  256x256), plus points on the last padded node, just inside it and beyond
  the grid.  Integer indices and the float bit patterns of the weights must
  equal the oracle's lint2d/lint3d.
- Reflectivity (expl=y, spher_utils.c:153-173, 250-310): the known-answer
  cases, one per position of the neighbour search order, rebuilt around the
  source node of a polar and a spherical model.

These calls create handles, which allocate device memory: GPU tests."""
import json
import os

import numpy as np
import pytest

from wavefield_imaging_amd import awefd, synth

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "known_answers.json")
f32 = np.float32


def _known():
    with open(GOLDEN) as f:
        return json.load(f)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _pad_origin(o, nb, d):
    return f32(o) - f32(nb) * f32(d)


def _oracle_map(orc, kind, par, nb, pts):
    """(idx fastest first, weights[8]) of the oracle's index map for `kind`."""
    n, o, d = par["n"], par["o"], par["d"]
    npd = [k + 2 * nb for k in n]
    pad = [_pad_origin(o[a], nb, d[a]) for a in range(len(n))]
    pts = np.asarray(pts, np.float32)
    w8 = np.zeros((len(pts), 8), np.float32)
    if kind == "cart2d":
        jz, jx, w = orc.lint2d(pts, npd[0], npd[1], pad[0], pad[1], d[0], d[1])
        idx = np.stack([jz, jx, np.zeros_like(jz)], 1)
        w8[:, :4] = w
    elif kind == "cart3d":
        jz, jx, jy, w = orc.lint3d(pts, npd[0], npd[1], npd[2], pad[0], pad[1], pad[2], d[0], d[1], d[2])
        idx = np.stack([jz, jx, jy], 1)
        w8[:] = w
    else:  # spher3d points (r, phi, theta): z = theta, x = r, y = phi
        jz, jx, jy, w = orc.lint3d(pts, npd[1], npd[0], npd[2], pad[1], pad[0], pad[2], d[1], d[0], d[2])
        idx = np.stack([jx, jz, jy], 1)       # (r, theta, phi)
        w8[:] = w
    return idx, w8


def _edge_points(kind, par, nb):
    """Points on the padded grid's last node (outside the map's half-open
    range), a hair inside it, at the padded origin and far outside."""
    n, o, d = par["n"], par["o"], par["d"]
    ndim = len(n)
    lo = [_pad_origin(o[a], nb, d[a]) for a in range(ndim)]
    last = [f32(lo[a] + f32(n[a] + 2 * nb - 1) * f32(d[a])) for a in range(ndim)]
    inside = [np.nextafter(last[a], f32(-np.inf)) for a in range(ndim)]
    mid = [f32(o[a] + f32(n[a] // 2) * f32(d[a])) for a in range(ndim)]
    cases = []
    for a in range(ndim):
        for v in (last[a], inside[a], lo[a], f32(last[a] + 1000 * d[a]), f32(lo[a] - d[a])):
            p = list(mid)
            p[a] = v
            cases.append(p)

    def order(p):  # model axes -> coordinate-file order
        if kind == "cart2d":
            return [p[1], p[0]]
        if kind == "cart3d":
            return [p[1], p[2], p[0]]
        return [p[0], p[2], p[1]]   # spher3d (r, phi, theta)

    return np.array([order(p) for p in cases], np.float32)


@pytest.mark.parametrize("kind", ["cart2d", "cart3d", "spher3d"])
def test_index_map_workload_geometry(orc, kind):
    case = getattr(synth, kind)(nt=4)
    par = case["par"]
    nb = par["nb"]
    extra = _edge_points(kind, par, nb)
    with awefd.Model(kind, par) as m:
        for which, pts in ((0, np.concatenate([case["sou"], extra])), (1, np.concatenate([case["rec"], extra]))):
            if which == 0:
                m.set_sources(pts)
            else:
                m.set_receivers(pts)
            print("index map %s which=%d: %d points" % (kind, which, len(pts)), flush=True)
            idx, w = m.points(which)
            ridx, rw = _oracle_map(orc, kind, par, nb, pts)
            assert np.array_equal(idx, ridx), kind
            assert np.array_equal(_bits(w), _bits(rw)), kind
            # the edge cases: the last node and beyond map to node 0 with the
            # whole weight on the z+1 corner; just inside it to n-2
            assert np.any(idx[-len(extra):] == 0)


def test_index_map_known_answers():
    """The hand-derived cases, through grids whose padded origin and sampling
    are the cases' (nb = 4)."""
    k = _known()
    for case in k["lint2d"]:
        nzp, nxp, ozp, oxp, dz, dx = case["grid"]
        par = dict(n=[nzp - 8, nxp - 8], o=[ozp + 4 * dz, oxp + 4 * dx], d=[dz, dx], nt=4, dt=synth.DT, nb=4,
                   dabc=False)
        with awefd.Model("cart2d", par) as m:
            m.set_sources(np.float32([case["xz"]]))
            idx, w = m.points(0)
        assert list(idx[0, :2]) == case["j"], case["name"]
        assert np.array_equal(_bits(w[0, :4]), _bits(np.float32(case["w"]))), case["name"]
    for case in k["lint3d"]:
        nzp, nxp, nyp, ozp, oxp, oyp, dz, dx, dy = case["grid"]
        par = dict(n=[nzp - 8, nxp - 8, nyp - 8], o=[ozp + 4 * dz, oxp + 4 * dx, oyp + 4 * dy], d=[dz, dx, dy],
                   nt=4, dt=synth.DT, nb=4, dabc=False)
        with awefd.Model("cart3d", par) as m:
            m.set_sources(np.float32([case["xyz"]]))
            idx, w = m.points(0)
        assert list(idx[0]) == case["j"], case["name"]
        assert np.array_equal(_bits(w[0]), _bits(np.float32(case["w"]))), case["name"]


@pytest.mark.parametrize("key", ["refl2d", "refl3d"])
def test_reflectivity_known_answers(orc, key):
    """The library's expl=y injection factor at each search position equals
    the hand-derived answer (and the oracle's on the same padded model)."""
    ndim = 2 if key == "refl2d" else 3
    kind = "spher2d" if ndim == 2 else "spher3d"
    n, nb = [16, 14, 12][:ndim], 4
    dth = f32(5.0 / 3388.5e3)
    o = [f32(3387.5e3), f32(np.pi / 4 - n[1] / 2 * dth), f32(0.4)][:ndim]
    d = [f32(5.0), dth, dth][:ndim]
    par = dict(n=n, o=o, d=d, nt=4, dt=synth.DT, nb=nb, expl=True)
    # a source on a cell corner in the middle of the model
    c = [6.25, 6.5, 5.75][:ndim]
    src = [o[a] + c[a] * d[a] for a in range(ndim)]
    sou = np.float32([[src[0], src[1]] if ndim == 2 else [src[0], src[2], src[1]]])
    with awefd.Model(kind, par) as m:
        m.set_sources(sou)
        m.set_receivers(sou)
        m.set_wavelet(np.zeros(4, np.float32))
        node = [int(v) for v in m.points(0)[0][0][:ndim]]        # padded (r, theta[, phi])
        for case in _known()[key]:
            vel = np.full(n[::-1], f32(case["v0"]), np.float32)   # ([phi,] theta, r)
            for e in case["set"]:
                off, v = e[:-1], e[-1]
                u = [node[a] + off[a] - nb for a in range(ndim)]
                assert all(0 < u[a] < n[a] - 1 for a in range(ndim))
                vel[tuple(u[::-1])] = f32(v)
            m.set_model(vel)
            got = m.points(2)[1][0, 0]
            assert got == f32(case["r"]), (key, case["case"], case["position"], got, case["r"])
            padded = np.pad(vel, nb, mode="edge")
            assert orc.refl_at(padded, node) == got
