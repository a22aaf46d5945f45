"""ctypes front end of the CPU oracle (test infrastructure only).

Only the test suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module.  It loads ``oracle/liborc.so`` (built by
``oracle/Makefile``), an independent C restatement of the reference CPU
programs ``CODE/MAWEFD{cart,spher}{2,3}Dcpu.c`` of Nacnud04/wavefield_imaging.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_LIB_REF = None

KINDS = ("cart2d", "cart3d", "spher2d", "spher3d")


class OrcPar(C.Structure):
    _fields_ = [
        ("n", C.c_int * 3), ("o", C.c_float * 3), ("d", C.c_float * 3),
        ("nt", C.c_int), ("dt", C.c_float), ("nb", C.c_int),
        ("fdorder", C.c_int), ("optfd", C.c_int), ("dabc", C.c_int),
        ("fsrf", C.c_int), ("expl", C.c_int), ("cden", C.c_int), ("sse", C.c_int),
        ("ns", C.c_int), ("nr", C.c_int), ("snap", C.c_int), ("jsnap", C.c_int),
        ("adj", C.c_int), ("kt", C.c_int), ("hybrid", C.c_int), ("sinc", C.c_int),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(ref_form=False):
    """The checker build (liborc.so), or with ref_form=True the
    reference-form timing build of the same code (liborc_ref.so: -O3, the
    4-wide SSE step with schedule(static,1)), used for CPU baselines."""
    global _LIB, _LIB_REF
    if ref_form:
        if _LIB_REF is None:
            path = os.path.join(_HERE, "liborc_ref.so")
            if not os.path.exists(path):
                build()
            _LIB_REF = C.CDLL(path)
        return _LIB_REF
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
    return _LIB


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def fdcoef(nop, d, optfd=False, order=2):
    d = np.ascontiguousarray(d, dtype=np.float32)
    out = np.zeros(len(d) * nop + 1, np.float32)
    rc = lib().orc_fdcoef(nop, len(d), _fp(d), int(optfd), order, _fp(out))
    if rc:
        raise ValueError("unsupported fd order")
    return out


def sponge_table(nb):
    w = np.zeros(nb, np.float32)
    lib().orc_sponge_table(nb, _fp(w))
    return w


def spher_damp(i, nb):
    f = lib().orc_spher_damp
    f.restype = C.c_double
    return f(int(i), int(nb))


def spher3d_rule(ir, ith, iph, nr, nth, nph):
    return lib().orc_spher3d_rule(ir, ith, iph, nr, nth, nph)


def refl_at(vel, j):
    """expl=y reflectivity at padded node j = (jr, jth[, jph]) of the raw
    padded velocity vel (array axes slowest first: [phi,] theta, r)."""
    vel = np.ascontiguousarray(vel, np.float32)
    f = lib().orc_refl_at
    f.restype = C.c_float
    ndim = vel.ndim
    jph = int(j[2]) if ndim == 3 else 0
    return np.float32(f(ndim, _fp(vel), int(j[0]), int(j[1]), jph, vel.shape[-1], vel.shape[-2]))


def lint2d(xz, nzpad, nxpad, ozpad, oxpad, dz, dx):
    xz = np.ascontiguousarray(xz, np.float32).reshape(-1)
    na = xz.size // 2
    jz = np.zeros(na, np.int32)
    jx = np.zeros(na, np.int32)
    w = np.zeros(4 * na, np.float32)
    lib().orc_lint2d(na, _fp(xz), nzpad, nxpad, C.c_float(ozpad), C.c_float(oxpad),
                     C.c_float(dz), C.c_float(dx), _ip(jz), _ip(jx), _fp(w))
    return jz, jx, w.reshape(na, 4)


def lint3d(xyz, nzpad, nxpad, nypad, ozpad, oxpad, oypad, dz, dx, dy):
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1)
    na = xyz.size // 3
    jz = np.zeros(na, np.int32)
    jx = np.zeros(na, np.int32)
    jy = np.zeros(na, np.int32)
    w = np.zeros(8 * na, np.float32)
    lib().orc_lint3d(na, _fp(xyz), nzpad, nxpad, nypad, C.c_float(ozpad), C.c_float(oxpad),
                     C.c_float(oypad), C.c_float(dz), C.c_float(dx), C.c_float(dy),
                     _ip(jz), _ip(jx), _ip(jy), _fp(w))
    return jz, jx, jy, w.reshape(na, 8)


def padded_shape(kind, n, nb=None, fdorder=4, hybrid=False):
    """Padded grid (fastest axis first) the oracle uses for ``kind``;
    hybrid: dabc=y hybridbc=y, which widens nb <= nop to 2 nop."""
    if kind.startswith("cart"):
        nop = fdorder // 2
        b = nop if nb is None or nb < nop else nb
        if hybrid and b <= nop:
            b = 2 * nop
    else:
        b = 4 if nb is None or nb < 4 else nb
    dims = 3 if kind.endswith("3d") else 2
    return tuple(n[i] + 2 * b for i in range(dims)), b


def run(kind, par, vel, sou, rec, wav, den=None, precision="f32", final=True, snaps=False, ref_form=False):
    """Run one oracle propagation.

    par keys: n (list), o, d, nt, dt, nb, fdorder, optfd, dabc, fsrf, expl, cden,
    sse, snap, jsnap, adj, kt, hybridbc.  Returns (traces[nt, nr], final_field or
    None, snapshots list); for a cartesian 2-D adj=y run the final field is the
    one the image is cut from (``imaged()`` tells whether step nt-kt was
    reached) and the traces stay zero.
    """
    if kind not in KINDS:
        raise ValueError(kind)
    real = np.float32 if precision == "f32" else np.float64
    p = OrcPar()
    n = list(par["n"]) + [1] * (3 - len(par["n"]))
    o = list(par.get("o", [0.0] * 3)) + [0.0] * 3
    d = list(par["d"]) + [1.0] * 3
    for i in range(3):
        p.n[i] = int(n[i])
        p.o[i] = float(o[i])
        p.d[i] = float(d[i])
    p.nt = int(par["nt"])
    p.dt = float(par["dt"])
    p.nb = int(par.get("nb", -1) if par.get("nb") is not None else -1)
    p.fdorder = int(par.get("fdorder", 4))
    p.optfd = int(_yn(par.get("optfd", False)))
    p.dabc = int(_yn(par.get("dabc", False)))
    p.fsrf = int(_yn(par.get("free", par.get("fsrf", False))))
    p.expl = int(_yn(par.get("expl", kind.startswith("spher"))))
    p.cden = int(_yn(par.get("cden", True)))
    p.sse = int(_yn(par.get("sse", True)))
    p.snap = int(_yn(par.get("snap", False)))
    p.jsnap = int(par.get("jsnap", p.nt))
    p.adj = int(_yn(par.get("adj", False)))
    p.kt = int(par.get("kt", 0))
    p.hybrid = int(_yn(par.get("hybridbc", False)))
    p.sinc = int(_yn(par.get("sinc", False))) if kind.startswith("cart") else 0
    ndim = 3 if kind.endswith("3d") else 2
    sou = np.ascontiguousarray(sou, np.float32).reshape(-1)
    rec = np.ascontiguousarray(rec, np.float32).reshape(-1)
    p.ns = sou.size // ndim
    p.nr = rec.size // ndim
    vel = np.ascontiguousarray(vel, np.float32)
    wav = np.ascontiguousarray(wav, np.float32).reshape(-1)
    dat = np.zeros((p.nt, p.nr), real)
    if kind.startswith("cart"):
        shp, _ = padded_shape(kind, n, p.nb, p.fdorder, hybrid=bool(p.dabc and p.hybrid))
    else:
        shp, _ = padded_shape(kind, n, p.nb)
    ufin = np.zeros(int(np.prod(shp)), real) if final else None
    rp = _fp if real is np.float32 else _dp
    CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int,
                     C.POINTER(C.c_float if real is np.float32 else C.c_double))
    shots = []
    ncell = int(np.prod(shp))

    def _cb(_ud, it, u):
        shots.append((it, np.ctypeslib.as_array(u, shape=(ncell,)).copy().reshape(shp[::-1])))

    cb = CB(_cb) if snaps else CB()
    fn = getattr(lib(ref_form), "orc_%s_%s" % (kind, precision))
    if kind.startswith("cart"):
        dn = None if den is None else np.ascontiguousarray(den, np.float32)
        rc = fn(C.byref(p), _fp(vel), _fp(dn) if dn is not None else None, _fp(sou), _fp(rec),
                _fp(wav), rp(dat), rp(ufin) if final else None, cb, None)
    else:
        rc = fn(C.byref(p), _fp(vel), _fp(sou), _fp(rec), _fp(wav), rp(dat),
                rp(ufin) if final else None, cb, None)
    if rc:
        raise ValueError("oracle rejected the configuration (rc=%d)" % rc)
    return dat, (ufin.reshape(shp[::-1]) if final else None), shots


def imaged():
    """Whether the last cartesian 2-D adj=y run reached its image step."""
    return bool(lib().orc_last_imaged())


def _yn(v):
    if isinstance(v, str):
        return v.lower().startswith("y")
    return bool(v)


def _par2d(par, sou, rec):
    p = OrcPar()
    for i in range(3):
        p.n[i] = int((list(par["n"]) + [1])[i])
        p.o[i] = float((list(par.get("o", [0.0, 0.0])) + [0.0])[i])
        p.d[i] = float((list(par["d"]) + [1.0])[i])
    p.nt = int(par["nt"])
    p.dt = float(par["dt"])
    p.nb = int(par.get("nb", -1) if par.get("nb") is not None else -1)
    p.fdorder = int(par.get("fdorder", 4))
    p.optfd = int(_yn(par.get("optfd", False)))
    p.dabc = int(_yn(par.get("dabc", False)))
    p.fsrf = int(_yn(par.get("free", par.get("fsrf", False))))
    p.expl = int(_yn(par.get("expl", False)))
    p.ns = np.asarray(sou).size // 2
    p.nr = np.asarray(rec).size // 2
    return p


def _f(a):
    return np.ascontiguousarray(a, np.float32).reshape(-1)


def grad_f64(par, vel, sou, rec, wav, dobs=None, res=None, slow=None):
    """FP64 L2-misfit gradient w.r.t. squared slowness (model grid): returns
    (grad, misfit, traces).  res, when given, replaces d - dobs (the adjoint
    source); slow (float64, model grid) replaces vel as the model."""
    p = _par2d(par, sou, rec)
    vel = np.ascontiguousarray(vel, np.float32)
    g = np.zeros(vel.shape, np.float64)
    J = C.c_double()
    dat = np.zeros((p.nt, p.nr), np.float64)
    do = None if dobs is None else _f(dobs)
    rs = None if res is None else np.ascontiguousarray(res, np.float64)
    sl = None if slow is None else np.ascontiguousarray(slow, np.float64)
    rc = lib().orc_cart2d_grad_f64(C.byref(p), _fp(vel), _dp(sl) if sl is not None else None, _fp(_f(sou)),
                                   _fp(_f(rec)), _fp(_f(wav)), _fp(do) if do is not None else None,
                                   _dp(rs) if rs is not None else None, _dp(g), C.byref(J), _dp(dat))
    if rc:
        raise ValueError("oracle gradient rejected the configuration (rc=%d)" % rc)
    return g, J.value, dat


def tlm_f64(par, vel, sou, rec, wav, ds, slow=None):
    """FP64 traces of the squared-slowness perturbation ds (model grid)."""
    p = _par2d(par, sou, rec)
    vel = np.ascontiguousarray(vel, np.float32)
    ds = np.ascontiguousarray(ds, np.float64)
    sl = None if slow is None else np.ascontiguousarray(slow, np.float64)
    dd = np.zeros((p.nt, p.nr), np.float64)
    rc = lib().orc_cart2d_tlm_f64(C.byref(p), _fp(vel), _dp(sl) if sl is not None else None, _fp(_f(sou)),
                                  _fp(_f(rec)), _fp(_f(wav)), _dp(ds), _dp(dd))
    if rc:
        raise ValueError("oracle TLM rejected the configuration (rc=%d)" % rc)
    return dd


def fwd_f64(par, vel, sou, rec, wav, slow=None):
    """Traces of the FP64 scheme the adjoint oracle differentiates."""
    p = _par2d(par, sou, rec)
    sl = None if slow is None else np.ascontiguousarray(slow, np.float64)
    dat = np.zeros((p.nt, p.nr), np.float64)
    rc = lib().orc_cart2d_fwd_f64(C.byref(p), _fp(np.ascontiguousarray(vel, np.float32)),
                                  _dp(sl) if sl is not None else None, _fp(_f(sou)), _fp(_f(rec)),
                                  _fp(_f(wav)), _dp(dat))
    if rc:
        raise ValueError("oracle forward rejected the configuration (rc=%d)" % rc)
    return dat
