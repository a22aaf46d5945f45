"""Drop-in programs sfAWEFDcart2D/3D and sfAWEFDspher2D/3D (RSF in, RSF out).

CPU tests: the RSF reader/writer and the programs' argument errors (which
stop before any GPU call).  GPU tests: each program run on RSF files, its
trace and snapshot files compared bit for bit with the oracle (traces,
snapshot windows cut as the reference's cut2d/cut3d do)."""
import os
import subprocess

import numpy as np
import pytest

from wavefield_imaging_amd import rsf, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "wavefield_imaging_amd", "bin")


def _prog(name):
    path = os.path.join(BIN, name)
    assert os.path.exists(path), "build the programs first (make -C wavefield_imaging_amd/cli)"
    return path


def test_rsf_round_trip(tmp_path):
    a = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    p = str(tmp_path / "a.rsf")
    rsf.write(p, a, [(1.5, 0.25, "z"), (-2.0, 4.0, "x"), (0.0, 1e-9, "t")])
    b, h = rsf.read(p)
    np.testing.assert_array_equal(a, b)
    assert h["n1"] == "4" and float(h["d3"]) == 1e-9 and h["label1"] == "z"


def test_rsf_separate_xdr_data(tmp_path):
    a = np.linspace(-1, 1, 12, dtype=np.float32).reshape(3, 4)
    a.astype(">f4").tofile(str(tmp_path / "a.rsf@"))
    with open(str(tmp_path / "a.rsf"), "w") as f:
        f.write('sfspike n1=9\n\tn1=4\n\tn2=3\n\tin="%s"\n\tdata_format="xdr_float"\n\tesize=4\n'
                % str(tmp_path / "a.rsf@"))
    b, _ = rsf.read(str(tmp_path / "a.rsf"))
    np.testing.assert_array_equal(a, b)


def _run(prog, args, stdin_path, stdout_path, cwd):
    with open(stdin_path, "rb") as fi, open(stdout_path, "wb") as fo:
        return subprocess.run([_prog(prog)] + args, stdin=fi, stdout=fo, stderr=subprocess.PIPE,
                              cwd=cwd, timeout=120)


def test_program_argument_errors(tmp_path):
    w = str(tmp_path / "w.rsf")
    rsf.write(w, np.zeros((10, 1), np.float32), [(0, 1), (0, 1e-8)])
    r = _run("sfAWEFDcart2D", [], w, str(tmp_path / "o.rsf"), str(tmp_path))
    assert r.returncode == 1 and b"vel=" in r.stderr
    r = _run("sfAWEFDspher3D", ["vel=nothere.rsf"], w, str(tmp_path / "o.rsf"), str(tmp_path))
    assert r.returncode == 1 and b"nothere" in r.stderr


def test_adj_needs_an_image_step(tmp_path):
    """2-D adj=y writes the field at step nt-kt; a kt outside 1..nt (0 when
    absent) would write no image, so the program stops before the GPU."""
    case = synth.cart2d(scale=0.05, nt=40)
    _write_case(tmp_path, "cart2d", case["par"], case["vel"], case["sou"], case["rec"], case["wav"])
    for extra in ([], ["kt=0"], ["kt=41"]):
        r = _run("sfAWEFDcart2D", ["vel=vel.rsf", "sou=sou.rsf", "rec=rec.rsf", "adj=y", "expl=y"] + extra,
                 str(tmp_path / "wav.rsf"), str(tmp_path / "o.rsf"), str(tmp_path))
        assert r.returncode == 1 and b"kt" in r.stderr, r.stderr


# ---------------------------------------------------------------- GPU

def _write_case(tmp, kind, par, vel, sou, rec, wav, den=None):
    ndim = 3 if kind.endswith("3d") else 2
    o, d = list(par["o"]), list(par["d"])
    rsf.write(str(tmp / "vel.rsf"), vel, [(o[a], d[a]) for a in range(ndim)])
    if den is not None:
        rsf.write(str(tmp / "den.rsf"), den, [(o[a], d[a]) for a in range(ndim)])
    rsf.write(str(tmp / "sou.rsf"), np.asarray(sou, np.float32).reshape(-1, ndim), [(0, 1), (0, 1)])
    rsf.write(str(tmp / "rec.rsf"), np.asarray(rec, np.float32).reshape(-1, ndim), [(0, 1), (0, 1)])
    w = np.asarray(wav, np.float32)
    w = w.reshape(par["nt"], -1)
    rsf.write(str(tmp / "wav.rsf"), w, [(0, 1), (0.0, par["dt"], "t")])


def _snaps_window(snaps, sl):
    return np.stack([f[sl] for _, f in snaps])


@pytest.mark.gpu
def test_cart2d_program(orc, tmp_path):
    case = synth.cart2d(scale=0.08, nt=160)
    par = dict(case["par"], nb=8, snap=True, jsnap=40)
    _write_case(tmp_path, "cart2d", par, case["vel"], case["sou"], case["rec"], case["wav"])
    args = ["vel=vel.rsf", "sou=sou.rsf", "rec=rec.rsf", "dabc=y", "nb=8", "snap=y", "jsnap=40",
            "wfl=wfl.rsf"]
    r = _run("sfAWEFDcart2D", args, str(tmp_path / "wav.rsf"), str(tmp_path / "dat.rsf"), str(tmp_path))
    assert r.returncode == 0, r.stderr.decode()
    dat, h = rsf.read(str(tmp_path / "dat.rsf"))
    wfl, hw = rsf.read(str(tmp_path / "wfl.rsf"))
    ref, _, snaps = orc.run("cart2d", par, case["vel"], case["sou"], case["rec"], case["wav"], snaps=True)
    np.testing.assert_array_equal(dat, ref)
    assert int(h["n1"]) == ref.shape[1] and int(h["n2"]) == par["nt"]
    nb, (nz, nx) = 8, par["n"]
    assert wfl.shape == (len(snaps), nx, nz) and [it for it, _ in snaps] == [0, 40, 80, 120]
    np.testing.assert_array_equal(wfl, _snaps_window(snaps, (slice(nb, nb + nx), slice(nb, nb + nz))))


@pytest.mark.gpu
def test_cart3d_program_window_and_jdata(orc, tmp_path):
    case = synth.cart3d(scale=0.06, nt=90)
    par = dict(case["par"], nb=6, snap=True, jsnap=30)
    den = (1000.0 + case["vel"] / 3e5).astype(np.float32)
    par_den = dict(par, cden=False)
    _write_case(tmp_path, "cart3d", par, case["vel"], case["sou"], case["rec"], case["wav"], den=den)
    nz, nx, ny = par["n"]
    # window: every 2nd z from z=10 m, x from 5 m, y from 0, 4 nodes each axis
    args = ["vel=vel.rsf", "sou=sou.rsf", "rec=rec.rsf", "den=den.rsf", "dabc=y", "nb=6", "snap=y",
            "jsnap=30", "wfl=wfl.rsf", "jdata=2", "nqz=4", "oqz=10", "dqz=10", "nqx=4", "oqx=5",
            "nqy=4", "oqy=0"]
    r = _run("sfAWEFDcart3D", args, str(tmp_path / "wav.rsf"), str(tmp_path / "dat.rsf"), str(tmp_path))
    assert r.returncode == 0, r.stderr.decode()
    dat, h = rsf.read(str(tmp_path / "dat.rsf"))
    wfl, _ = rsf.read(str(tmp_path / "wfl.rsf"))
    ref, _, snaps = orc.run("cart3d", par_den, case["vel"], case["sou"], case["rec"], case["wav"], den=den,
                            snaps=True)
    assert int(h["n2"]) == (par["nt"] - 1) // 2 + 1 and float(h["d2"]) == pytest.approx(2 * par["dt"])
    np.testing.assert_array_equal(dat, ref[::2])
    nb = 6
    sl = (slice(nb, nb + 4), slice(nb + 1, nb + 5), slice(nb + 2, nb + 10, 2))   # (y, x, z)
    np.testing.assert_array_equal(wfl, _snaps_window(snaps, sl))


@pytest.mark.gpu
def test_spher2d_program_padded_snapshots(orc, tmp_path):
    case = synth.spher2d(scale=0.08, nt=120)
    par = dict(case["par"], nb=6, expl=False, snap=True, jsnap=50)
    _write_case(tmp_path, "spher2d", par, case["vel"], case["sou"], case["rec"], case["wav"])
    args = ["vel=vel.rsf", "sou=sou.rsf", "rec=rec.rsf", "nb=6", "expl=n", "jsnap=50", "wfl=wfl.rsf"]
    r = _run("sfAWEFDspher2D", args, str(tmp_path / "wav.rsf"), str(tmp_path / "dat.rsf"), str(tmp_path))
    assert r.returncode == 0, r.stderr.decode()
    dat, _ = rsf.read(str(tmp_path / "dat.rsf"))
    wfl, hw = rsf.read(str(tmp_path / "wfl.rsf"))
    ref, _, snaps = orc.run("spher2d", par, case["vel"], case["sou"], case["rec"], case["wav"], snaps=True)
    np.testing.assert_array_equal(dat, ref)
    # snap=y bnds=y by default: padded field, theta fastest
    np.testing.assert_array_equal(wfl, np.stack([f.T for _, f in snaps]))
    assert hw["label1"] == "th"


@pytest.mark.gpu
def test_spher3d_program(orc, tmp_path):
    case = synth.spher3d(scale=0.06, nt=60)
    par = dict(case["par"], nb=5, snap=True, jsnap=25)
    _write_case(tmp_path, "spher3d", par, case["vel"], case["sou"], case["rec"], case["wav"])
    args = ["vel=vel.rsf", "sou=sou.rsf", "rec=rec.rsf", "nb=5", "expl=n", "snap=y", "jsnap=25",
            "wfl=wfl.rsf"]
    r = _run("sfAWEFDspher3D", args, str(tmp_path / "wav.rsf"), str(tmp_path / "dat.rsf"), str(tmp_path))
    assert r.returncode == 0, r.stderr.decode()
    dat, _ = rsf.read(str(tmp_path / "dat.rsf"))
    wfl, _ = rsf.read(str(tmp_path / "wfl.rsf"))
    ref, _, snaps = orc.run("spher3d", par, case["vel"], case["sou"], case["rec"], case["wav"], snaps=True)
    np.testing.assert_array_equal(dat, ref)
    nr, nth, nph = par["n"]
    sl = (slice(5, 5 + nph), slice(5, 5 + nth), slice(5, 5 + nr))
    np.testing.assert_array_equal(wfl, _snaps_window(snaps, sl))


@pytest.mark.gpu
def test_cart2d_program_adj_image(orc, tmp_path):
    """adj=y kt=: the receiver data injected back to front at the receivers,
    the image (the interior at step nt-kt) written instead of traces, the
    snapshots up to that step; hybridbc=y on the same run."""
    case = synth.cart2d(scale=0.08, nt=150)
    nt, kt = 150, 20
    rec = case["rec"][::4]
    rng = np.random.default_rng(2)
    data = (case["wav"][:, None] * rng.standard_normal(len(rec))[None, :]).astype(np.float32)
    par = dict(case["par"], nb=3, dabc=True, hybridbc=True, expl=False, adj=True, kt=kt, snap=True,
               jsnap=40)
    _write_case(tmp_path, "cart2d", par, case["vel"], case["sou"], rec, data)
    args = ["vel=vel.rsf", "sou=sou.rsf", "rec=rec.rsf", "dabc=y", "hybridbc=y", "nb=3", "adj=y",
            "kt=%d" % kt, "snap=y", "jsnap=40", "wfl=wfl.rsf"]
    r = _run("sfAWEFDcart2D", args, str(tmp_path / "wav.rsf"), str(tmp_path / "img.rsf"), str(tmp_path))
    assert r.returncode == 0, r.stderr.decode()
    img, h = rsf.read(str(tmp_path / "img.rsf"))
    wfl, hw = rsf.read(str(tmp_path / "wfl.rsf"))
    _, fin, snaps = orc.run("cart2d", par, case["vel"], case["sou"], rec, data, snaps=True)
    assert orc.imaged()
    nb, (nz, nx) = 3, par["n"]
    assert img.shape == (nx, nz) and int(h["n1"]) == nz and int(h["n2"]) == nx
    assert float(h["o1"]) == par["o"][0] and float(h["d2"]) == par["d"][1]
    np.testing.assert_array_equal(img, fin[nb:nb + nx, nb:nb + nz])
    assert np.abs(img).max() > 0
    assert [it for it, _ in snaps] == [0, 40, 80, 120] and int(hw["n3"]) == 4
    np.testing.assert_array_equal(wfl, _snaps_window(snaps, (slice(nb, nb + nx), slice(nb, nb + nz))))


@pytest.mark.gpu
def test_cart3d_program_adj_hybrid(orc, tmp_path):
    """3-D adj=y reads the wavelet back to front and still writes traces."""
    case = synth.cart3d(scale=0.06, nt=70)
    par = dict(case["par"], nb=5, dabc=True, hybridbc=True, free=True, adj=True)
    _write_case(tmp_path, "cart3d", par, case["vel"], case["sou"], case["rec"], case["wav"])
    args = ["vel=vel.rsf", "sou=sou.rsf", "rec=rec.rsf", "dabc=y", "hybridbc=y", "free=y", "nb=5", "adj=y"]
    r = _run("sfAWEFDcart3D", args, str(tmp_path / "wav.rsf"), str(tmp_path / "dat.rsf"), str(tmp_path))
    assert r.returncode == 0, r.stderr.decode()
    dat, _ = rsf.read(str(tmp_path / "dat.rsf"))
    ref, _, _ = orc.run("cart3d", par, case["vel"], case["sou"], case["rec"], case["wav"])
    assert np.abs(ref).max() > 0
    np.testing.assert_array_equal(dat, ref)


@pytest.mark.gpu
def test_cart2d_program_sinc(orc, tmp_path):
    case = synth.cart2d(scale=0.08, nt=120)
    sou = np.float32([[101.3, 62.2]])
    rec = case["rec"] + np.float32([[0.0, 1.7]])
    par = dict(case["par"], nb=8, sinc=True, optfd=True)
    _write_case(tmp_path, "cart2d", par, case["vel"], sou, rec, case["wav"])
    args = ["vel=vel.rsf", "sou=sou.rsf", "rec=rec.rsf", "dabc=y", "nb=8", "sinc=y", "optfd=y"]
    r = _run("sfAWEFDcart2D", args, str(tmp_path / "wav.rsf"), str(tmp_path / "dat.rsf"), str(tmp_path))
    assert r.returncode == 0, r.stderr.decode()
    dat, _ = rsf.read(str(tmp_path / "dat.rsf"))
    ref, _, _ = orc.run("cart2d", par, case["vel"], sou, rec, case["wav"])
    assert np.abs(ref).max() > 0
    np.testing.assert_array_equal(dat, ref)
