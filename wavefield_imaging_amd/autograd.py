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

The model is set from device memory (awefd_set_model_dev), so an update of
the slowness never passes through the host.  The binding is ctypes over
libawefd; torch provides device memory and the autograd graph only.  The library works on its own HIP stream and returns
after synchronising it; torch's current stream is synchronised before each
call, so tensors torch produced are complete when the library reads them.
"""
import ctypes as C

import torch

from . import _lib as L
from .awefd import Model


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
