"""The allocation-free run (awefd_run_noalloc) against awefd_run, through the
C ABI: the same traces and the same final field bit for bit, on two calls in
a row on one GPU, the second with a smaller input; trace blocks beyond the
fixed scratch are refused with an error before anything runs."""
import ctypes as C

import numpy as np
import pytest

from wavefield_imaging_amd import _lib as L
from wavefield_imaging_amd import awefd, synth

pytestmark = pytest.mark.gpu


def _setup(kind, scale, nt, **kw):
    case = getattr(synth, kind)(scale=scale, nt=nt)
    par = dict(case["par"], **kw)
    m = awefd.Model(kind, par)
    m.set_model(case["vel"])
    m.set_sources(case["sou"])
    m.set_receivers(case["rec"])
    m.set_wavelet(case["wav"])
    return m


def _noalloc(m):
    tr = np.full((m.nt, m.nr), np.nan, np.float32)
    L.check(m.lib.awefd_run_noalloc(m.h, L.fptr(tr)))
    return tr, m.field()


def _reference(m):
    tr, _ = m.run()
    return tr, m.field()


CASES = [
    # (kind, [(scale, nt, params) of the first call, of the smaller second call])
    ("cart2d", [(0.2, 300, {}), (0.1, 200, {})]),
    ("cart2d", [(0.15, 260, dict(adj=True, expl=True, kt=60)), (0.1, 120, dict(adj=True, expl=True, kt=30))]),
    ("cart3d", [(0.12, 60, {}), (0.07, 40, dict(fdorder=8, nb=8))]),
    ("spher2d", [(0.12, 120, dict(expl=True)), (0.08, 80, {})]),
    ("spher3d", [(0.07, 40, {}), (0.05, 30, dict(expl=True))]),
]


@pytest.mark.parametrize("kind,calls", CASES, ids=[c[0] + str(i) for i, c in enumerate(CASES)])
def test_noalloc_matches_run_twice(kind, calls):
    models = [_setup(kind, s, nt, **kw) for s, nt, kw in calls]
    got = []
    for i, m in enumerate(models):                  # two calls in a row, the second smaller
        print("noalloc %s call %d: padded %s nt %d nr %d" % (kind, i, m.padded, m.nt, m.nr), flush=True)
        got.append(_noalloc(m))
    assert models[1].nt * models[1].nr < models[0].nt * models[0].nr
    for i, m in enumerate(models):
        tr, fld = _reference(m)
        if m.params.adj and kind == "cart2d":
            assert np.array_equal(got[i][1], fld), (kind, i)     # the image
        else:
            assert np.any(tr), (kind, i)
            assert np.array_equal(got[i][0], tr), (kind, i)
            assert np.array_equal(got[i][1], fld), (kind, i)
    # and once more on the first model, after the smaller call left its traces
    tr, fld = _noalloc(models[0])
    ref, rfld = _reference(models[0])
    if not (models[0].params.adj and kind == "cart2d"):
        assert np.array_equal(tr, ref)
    assert np.array_equal(fld, rfld)
    for m in models:
        m.close()


def test_noalloc_refuses_large_trace_blocks():
    """nt * nr * 4 bytes above AWEFD_NOALLOC_TRACE_BYTES (32 MiB): an error,
    nothing launched; the model still runs through awefd_run."""
    case = synth.cart2d(scale=0.1, nt=400)
    m = awefd.Model("cart2d", case["par"])
    m.set_model(case["vel"])
    m.set_sources(case["sou"])
    nx = case["par"]["n"][1]
    nr = (32 << 20) // (4 * 400) + 1
    rec = np.stack([np.linspace(0, (nx - 1) * 5.0, nr), np.full(nr, 20.0)], 1).astype(np.float32)
    m.set_receivers(rec)
    m.set_wavelet(case["wav"])
    with pytest.raises(L.AwefdError, match="scratch"):
        L.check(m.lib.awefd_run_noalloc(m.h, None))
    m.profile(True)
    with pytest.raises(L.AwefdError, match="profiling"):
        L.check(m.lib.awefd_run_noalloc(m.h, None))
    m.close()
