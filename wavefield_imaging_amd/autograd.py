"""Traces of the cartesian 2-D loop as a differentiable function of squared
slowness, for torch.

``Traces(par, sou, rec, wav, ckpt=0)`` is a module: ``traces = m(slow)`` maps
a squared-slowness tensor (shape (n2, n1), n1 fastest, float32, on the
model's GPU) to the (nt, nr) trace tensor of the same loop as sfAWEFDcart2D
(MAWEFDcart2Dcpu.c:297-383).  The forward pass keeps a checkpoint of the
field pair every ``ckpt`` steps (0: ceil(sqrt(nt))); ``backward`` runs the
hand-written adjoint kernels from those checkpoints (awefd_adjoint_grad in
the C ABI), so ``loss.backward()`` gives dloss/dslow for any loss of the
traces.  The reference has no gradient; this layer has no counterpart there.

The checkpoints live in the model, one set at a time.  Each forward call is
stamped; when a backward belongs to an earlier forward than the one whose
checkpoints are live (two calls of one module before one backward, e.g.
f(m(s1)) + f(m(s2))), it first runs its own forward again from the saved
slowness, so every branch gets the gradient at its own model.

``BatchTraces(par, sources, rec, wav, ckpt=0)`` is the same for many shots
on one grid: ``m(slow)`` gives the (nshot, nt, nr) traces through the
batched graphs (awefd_batch_run_ckpt: three-step passes between
checkpoints) and ``backward`` sums the shots' gradients in shot order
(awefd_batch_adjoint_grad: three-step recompute, two-step adjoint passes),
bit for bit the sum of the single-shot gradients.

The model is set from device memory (awefd_set_model_dev), so an update of
the slowness never passes through the host.  The binding is ctypes over
libawefd; torch provides device memory and the autograd graph only.  The library works on its own HIP stream and returns
after synchronising it; torch's current stream is synchronised before each
call, so tensors torch produced are complete when the library reads them.
"""
import ctypes as C

import torch

from . import _lib as L
from .awefd import Batch, Model


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _forward_ckpt(owner, slow):
    """Set the model from squared slowness and run the checkpointed forward
    loop; returns the (nt, nr) traces."""
    m = owner.model
    # the library takes velocity; v = s^(-1/2), formed in float64 on the GPU
    # and handed over in device memory
    vel = slow.detach().double().rsqrt().float().contiguous()
    m.set_model_dev(vel)
    tr = torch.empty((m.nt, m.nr), dtype=torch.float32, device=owner.device)
    torch.cuda.current_stream(owner.device).synchronize()
    L.check(m.lib.awefd_run_ckpt(m.h, _ptr(tr), int(owner.ckpt)))
    return tr


class _TracesFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, slow, owner):
        if slow.dtype != torch.float32 or slow.device != owner.device:
            raise ValueError("slowness must be float32 on %s" % owner.device)
        if tuple(slow.shape) != owner.shape:
            raise ValueError("slowness shape %s, the model is %s" % (tuple(slow.shape), owner.shape))
        tr = _forward_ckpt(owner, slow)
        owner._calls += 1
        owner._live = owner._calls
        ctx.stamp = owner._live
        ctx.owner = owner
        ctx.save_for_backward(slow)
        return tr

    @staticmethod
    def backward(ctx, gtr):
        owner = ctx.owner
        m = owner.model
        if owner._live != ctx.stamp:
            # a later forward replaced the checkpoints: rebuild this call's
            (slow,) = ctx.saved_tensors
            _forward_ckpt(owner, slow)
            owner._live = ctx.stamp
        r = gtr.detach().to(torch.float32).contiguous()
        g = torch.empty(owner.shape, dtype=torch.float32, device=owner.device)
        torch.cuda.current_stream(owner.device).synchronize()
        L.check(m.lib.awefd_adjoint_grad(m.h, _ptr(r), _ptr(g)))
        return g, None


class Traces(torch.nn.Module):
    """Differentiable cartesian 2-D forward model (see the module docstring).

    par: the sfAWEFDcart2D parameters (n, o, d, nt, dt, nb, fdorder, optfd,
    dabc, free, expl, sse; cden must stay y).  sou, rec: (ns, 2) and (nr, 2)
    coordinates (x, z); wav: the wavelet as the program reads it."""

    def __init__(self, par, sou, rec, wav, ckpt=0):
        super().__init__()
        self.model = Model("cart2d", par)
        self.model.set_sources(sou)
        self.model.set_receivers(rec)
        self.model.set_wavelet(wav)
        self.ckpt = int(ckpt)
        self.shape = (int(par["n"][1]), int(par["n"][0]))
        self._calls = 0          # forward calls so far
        self._live = 0           # the call whose checkpoints the model holds
        dev = self.model.lib.awefd_device(self.model.h)
        self.device = torch.device("cuda", dev)

    def forward(self, slow):
        return _TracesFn.apply(slow, self)

    def gradient(self, residual):
        """dJ/ds for trace weights ``residual`` (nt, nr) at the model whose
        checkpoints the model holds (the last forward call, or the call the
        last backward rebuilt), without an autograd graph."""
        m = self.model
        r = residual.detach().to(device=self.device, dtype=torch.float32).contiguous()
        g = torch.empty(self.shape, dtype=torch.float32, device=self.device)
        torch.cuda.current_stream(self.device).synchronize()
        L.check(m.lib.awefd_adjoint_grad(m.h, _ptr(r), _ptr(g)))
        return g

    def close(self):
        if self.model is not None:
            self.model.close()
            self.model = None


def _forward_batch(owner, slow):
    """Set every shot's model from squared slowness and run the batched
    checkpointed forward loops; returns the (nshot, nt, nr) traces."""
    vel = slow.detach().double().rsqrt().float().contiguous()
    for m in owner.models:
        m.set_model_dev(vel)
    torch.cuda.current_stream(owner.device).synchronize()
    owner.batch.run_ckpt(owner.ckpt)
    # each shot's traces live in its model's buffer until the next run
    return torch.stack([m.traces_tensor() for m in owner.models])


class _BatchTracesFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, slow, owner):
        if slow.dtype != torch.float32 or slow.device != owner.device:
            raise ValueError("slowness must be float32 on %s" % owner.device)
        if tuple(slow.shape) != owner.shape:
            raise ValueError("slowness shape %s, the model is %s" % (tuple(slow.shape), owner.shape))
        tr = _forward_batch(owner, slow)
        owner._calls += 1
        owner._live = owner._calls
        ctx.stamp = owner._live
        ctx.owner = owner
        ctx.save_for_backward(slow)
        return tr

    @staticmethod
    def backward(ctx, gtr):
        owner = ctx.owner
        if owner._live != ctx.stamp:
            (slow,) = ctx.saved_tensors
            _forward_batch(owner, slow)
            owner._live = ctx.stamp
        return owner._gradient(gtr), None


class BatchTraces(torch.nn.Module):
    """Differentiable cartesian 2-D forward model of many shots on one grid
    (see the module docstring).

    par: the sfAWEFDcart2D parameters, shared by every shot; sources: one
    (ns, 2) coordinate array (x, z) per shot; rec, wav: one for all shots or
    one per shot.  ``m(slow)`` returns (nshot, nt, nr); the gradient of a
    loss is the sum over the shots."""

    def __init__(self, par, sources, rec, wav, ckpt=0):
        super().__init__()
        n = len(sources)
        if n < 1:
            raise ValueError("no shots")
        recs = rec if isinstance(rec, (list, tuple)) else [rec] * n
        wavs = wav if isinstance(wav, (list, tuple)) else [wav] * n
        if len(recs) != n or len(wavs) != n:
            raise ValueError("one receiver set and one wavelet for all shots, or one per shot")
        self.models = []
        self.batch = None
        try:
            for sou, r, w in zip(sources, recs, wavs):
                m = Model("cart2d", par)
                self.models.append(m)
                m.set_sources(sou)
                m.set_receivers(r)
                m.set_wavelet(w)
                # a model until the first forward call: the batch checks its shots
                m.set_model_dev(torch.ones((int(par["n"][1]), int(par["n"][0])), dtype=torch.float32,
                                           device=torch.device("cuda", m.lib.awefd_device(m.h))))
            self.batch = Batch(self.models)
        except Exception:
            self.close()
            raise
        if len({m.nr for m in self.models}) != 1:
            raise ValueError("every shot needs the same number of receivers (the traces are one tensor)")
        self.ckpt = int(ckpt)
        self.shape = (int(par["n"][1]), int(par["n"][0]))
        self._calls = 0
        self._live = 0
        self.device = torch.device("cuda", self.models[0].lib.awefd_device(self.models[0].h))

    def forward(self, slow):
        return _BatchTracesFn.apply(slow, self)

    def _gradient(self, residual):
        r = residual.detach().to(device=self.device, dtype=torch.float32).contiguous()
        if tuple(r.shape) != (len(self.models), self.models[0].nt, self.models[0].nr):
            raise ValueError("residual shape %s" % (tuple(r.shape),))
        g = torch.empty(self.shape, dtype=torch.float32, device=self.device)
        torch.cuda.current_stream(self.device).synchronize()
        self.batch.adjoint_grad([r[s].data_ptr() for s in range(len(self.models))], g.data_ptr())
        return g

    def gradient(self, residual):
        """Sum over the shots of dJ/ds for trace weights ``residual``
        (nshot, nt, nr) at the model of the live checkpoints, without an
        autograd graph."""
        return self._gradient(residual)

    def close(self):
        if self.batch is not None:
            self.batch.close()
            self.batch = None
        for m in self.models:
            m.close()
        self.models = []
