"""Python mirror of the four Madagascar programs over libawefd.

``run(kind, par, vel, sou, rec, wav, den=None)`` takes what the program
sfAWEFD<kind> reads (RSF arrays in axis-1-fastest order, getpar keys in
``par``) and returns the traces and, on request, the wavefield snapshots,
computed by the HIP kernels through the C ABI.  ``Model`` is the handle-level
interface used by the benchmark and the batched/adjoint layers.

Parameter keys and defaults follow the reference mains
(CODE/MAWEFDcart2Dcpu.c:116-211, CODE/MAWEFDcart3Dcpu.c:116-210,
CODE/MAWEFDspher2Dcpu.c:56-116, CODE/MAWEFDspher3Dcpu.c:57-133).
"""
import ctypes as C

import numpy as np

from . import _lib as L

_BOOL_KEYS = ("optfd", "dabc", "expl", "cden")


def _yn(v):
    if isinstance(v, str):
        return v.lower().startswith("y")
    return bool(v)


def make_params(kind, par):
    """Build an awefd_params from program parameters.

    par: n (list), o, d, nt, dt and any getpar key (nb, fdorder, optfd, dabc,
    free/fsrf, expl, cden, sse, jdata, gpu, adj, kt, hybridbc, sinc).
    """
    k = L.KIND_NAMES[kind] if isinstance(kind, str) else int(kind)
    lib = L.load()
    p = L.Params()
    L.check(lib.awefd_params_default(k, C.byref(p)))
    n = list(par["n"]) + [1, 1]
    o = list(par.get("o", [0.0, 0.0, 0.0])) + [0.0, 0.0]
    d = list(par["d"]) + [1.0, 1.0]
    for i in range(3):
        p.n[i] = int(n[i])
        p.o[i] = float(o[i])
        p.d[i] = float(d[i])
    p.nt = int(par["nt"])
    p.dt = float(par["dt"])
    if par.get("nb") is not None:
        p.nb = int(par["nb"])
    if "fdorder" in par:
        p.fdorder = int(par["fdorder"])
    for key in _BOOL_KEYS:
        if key in par:
            setattr(p, key, int(_yn(par[key])))
    if "free" in par or "fsrf" in par:
        p.fsrf = int(_yn(par.get("free", par.get("fsrf"))))
    if "sse" in par:
        p.sse = int(_yn(par["sse"]))
    if "jdata" in par:
        p.jdata = int(par["jdata"])
    if "adj" in par:
        p.adj = int(_yn(par["adj"]))
    if "kt" in par:
        p.kt = int(par["kt"])
    if "hybridbc" in par:
        p.hybrid = int(_yn(par["hybridbc"]))
    if "sinc" in par:
        p.sinc = int(_yn(par["sinc"]))
    p.device = int(par.get("gpu", 0))
    return p


class _DevArray:
    """A float32 device buffer described by __cuda_array_interface__."""

    def __init__(self, ptr, shape):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": "<f4", "data": (int(ptr), False),
                                         "version": 3, "strides": None}


class Model:
    """One grid resident on one GPU (an awefd_model handle)."""

    def __init__(self, kind, par):
        self.kind = kind
        self.lib = L.load()
        self.params = make_params(kind, par)
        self.ndim = 3 if str(kind).endswith("3d") or kind in (L.CART3D, L.SPHER3D) else 2
        h = C.c_void_p()
        L.check(self.lib.awefd_create(C.byref(self.params), C.byref(h)))
        self.h = h
        dims = (C.c_int * 3)()
        nb = C.c_int()
        L.check(self.lib.awefd_padded_dims(self.h, dims, C.byref(nb)))
        self.padded = tuple(dims[i] for i in range(3))
        self.nb = nb.value
        self.nt = self.params.nt
        self.ns = self.nr = 0

    # -- inputs -------------------------------------------------------
    def set_model(self, vel, den=None):
        vel = L.f32(vel)
        den = None if den is None else L.f32(den)
        L.check(self.lib.awefd_set_model(self.h, L.fptr(vel), L.fptr(den) if den is not None else None))

    def set_model_dev(self, vel):
        """Velocity from a torch tensor on the model's GPU (RSF order,
        unpadded, float32, contiguous): awefd_set_model_dev."""
        import torch

        if vel.dtype != torch.float32 or not vel.is_cuda or not vel.is_contiguous():
            raise ValueError("set_model_dev: a contiguous float32 CUDA tensor")
        torch.cuda.current_stream(vel.device).synchronize()
        L.check(self.lib.awefd_set_model_dev(self.h, vel.data_ptr()))

    def set_sources(self, coords):
        c = L.f32(coords).reshape(-1)
        self.ns = c.size // self.ndim
        L.check(self.lib.awefd_set_sources(self.h, self.ns, L.fptr(c)))

    def set_receivers(self, coords):
        c = L.f32(coords).reshape(-1)
        self.nr = c.size // self.ndim
        L.check(self.lib.awefd_set_receivers(self.h, self.nr, L.fptr(c)))

    def set_wavelet(self, wav):
        w = L.f32(wav).reshape(-1)
        L.check(self.lib.awefd_set_wavelet(self.h, w.size, L.fptr(w)))

    # -- runs ---------------------------------------------------------
    def run(self, jsnap=0, snapshots=False):
        """Full time loop; returns (traces[nt, nr], [(it, field), ...])."""
        tr = np.zeros((self.nt, max(self.nr, 0)), np.float32)
        shots = []
        shape = self.padded[: self.ndim][::-1]

        def _cb(_u, it, field, dims):
            n = int(np.prod(shape))
            shots.append((it, np.ctypeslib.as_array(field, shape=(n,)).copy().reshape(shape)))

        cb = L.SNAP_FN(_cb) if snapshots else L.SNAP_FN()
        L.check(self.lib.awefd_run(self.h, L.fptr(tr), int(jsnap) if snapshots else 0, cb, None))
        return tr, shots

    def run_dev(self, d_traces_ptr=None):
        """Full time loop with device-resident output (pointer or None)."""
        L.check(self.lib.awefd_run_dev(self.h, C.c_void_p(d_traces_ptr) if d_traces_ptr else None))

    def traces(self):
        """Traces of the last run, read back from device memory."""
        tr = np.zeros((self.nt, max(self.nr, 0)), np.float32)
        L.check(self.lib.awefd_get_traces(self.h, L.fptr(tr)))
        return tr

    def traces_dev(self):
        """Device address of the handle's trace buffer (nt x nr float32)."""
        p = C.c_void_p()
        L.check(self.lib.awefd_traces_dev(self.h, C.byref(p)))
        return p.value

    def traces_tensor(self):
        """The handle's trace buffer as a torch tensor sharing its memory
        (no copy); its contents change with the next run."""
        import torch

        return torch.as_tensor(_DevArray(self.traces_dev(), (self.nt, self.nr)), device="cuda")

    def field(self):
        """Padded field the reference would write after the last step."""
        out = np.zeros(self.padded[: self.ndim][::-1], np.float32)
        L.check(self.lib.awefd_get_field(self.h, L.fptr(out)))
        return out

    def points(self, which):
        """Index map of the sources (which=0) or receivers (1): (j1, j2, j3)
        fastest first and 8 corner weights per point; which=2: each source's
        injection factor in column 0 of the weights."""
        n = self.nr if which == 1 else self.ns
        idx = np.zeros((max(n, 1), 3), np.int32)
        w = np.zeros((max(n, 1), 8), np.float32)
        L.check(self.lib.awefd_debug_points(self.h, which, L.iptr(idx), L.fptr(w)))
        return idx[:n], w[:n]

    def profile(self, enable=True):
        L.check(self.lib.awefd_profile(self.h, int(enable)))

    def kernel_stats(self):
        n = C.c_long()
        ms = C.c_double()
        L.check(self.lib.awefd_kernel_stats(self.h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.awefd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Many 2-D cartesian shots advanced together (one launch per step over
    every shot, the time loop captured into one HIP graph)."""

    def __init__(self, models):
        self.lib = L.load()
        self.models = list(models)
        arr = (C.c_void_p * len(self.models))(*[m.h.value for m in self.models])
        h = C.c_void_p()
        L.check(self.lib.awefd_batch_create(len(self.models), arr, C.byref(h)))
        self.h = h

    RX = (2, 4, 8, 16)
    GROUP_MB = (0, 32, 64, 128, 192)

    def set_tile(self, rx, group_mb=0):
        L.check(self.lib.awefd_batch_set_tile(self.h, int(rx), int(group_mb)))

    def tune(self, steps=20):
        """Time every (tile, group) candidate on the batch and keep the fastest;
        returns (rx, group_mb, {(rx, group_mb): ms per step})."""
        rx, grp = C.c_int(), C.c_int()
        ms = (C.c_double * (len(self.RX) * len(self.GROUP_MB)))()
        L.check(self.lib.awefd_batch_tune(self.h, int(steps), C.byref(rx), C.byref(grp), ms))
        keys = [(r, g) for r in self.RX for g in self.GROUP_MB]
        return rx.value, grp.value, dict(zip(keys, list(ms)))

    def run(self):
        L.check(self.lib.awefd_batch_run(self.h))

    def set_depth(self, depth):
        """Steps per launch: 1, or 2 or 3 (tile rx 4 or 8)."""
        L.check(self.lib.awefd_batch_set_depth(self.h, int(depth)))

    def tune_depth(self, steps=20):
        """Time every (depth, tile, group) choice and keep the fastest;
        returns (rx, group_mb, depth, {(depth, rx, group_mb): ms per step})."""
        rx, grp, dep = C.c_int(), C.c_int(), C.c_int()
        n = len(self.RX) * len(self.GROUP_MB)
        ms = (C.c_double * (3 * n))()
        L.check(self.lib.awefd_batch_tune_depth(self.h, int(steps), C.byref(rx), C.byref(grp), C.byref(dep), ms))
        keys = [(d, r, g) for d in (1, 2, 3) for r in self.RX for g in self.GROUP_MB]
        return rx.value, grp.value, dep.value, {k: v for k, v in zip(keys, list(ms)) if v > 0}

    def run_ckpt(self, ckpt=0):
        """Every shot's forward loop keeping checkpoints every ``ckpt`` steps
        (0: ceil(sqrt(nt))) for adjoint_grad."""
        L.check(self.lib.awefd_batch_run_ckpt(self.h, int(ckpt)))

    def adjoint_grad(self, res_ptrs, grad_ptr):
        """Sum over the shots of dJ/ds (device pointer grad_ptr, model grid)
        for per-shot residuals at device pointers res_ptrs (nt x nr each)."""
        arr = (C.c_void_p * len(self.models))(*[int(p) for p in res_ptrs])
        L.check(self.lib.awefd_batch_adjoint_grad(self.h, arr, C.c_void_p(int(grad_ptr))))

    def close(self):
        if getattr(self, "h", None):
            self.lib.awefd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run(kind, par, vel, sou, rec, wav, den=None, snapshots=False):
    """sfAWEFD<kind> on the GPU: returns (traces[nt, nr], snapshots, final_field)."""
    with Model(kind, par) as m:
        m.set_model(vel, den)
        m.set_sources(sou)
        m.set_receivers(rec)
        m.set_wavelet(wav)
        tr, shots = m.run(jsnap=par.get("jsnap", par["nt"]), snapshots=snapshots)
        return tr, shots, m.field()
