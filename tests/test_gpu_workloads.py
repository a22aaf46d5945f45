"""The batched workloads at the sizes they are measured at.

- 256 independent 2-D shots of uneven size (nz ~ U[200,600], nx ~
  U[400,1200], nb = 50, nt = 2000) through the three-steps-per-pass batch
  (tile rx 8, one group: the tuned choice on MI355X): every shot's traces and
  final field equal its own single-grid run bit for bit, and two shots (the
  smallest and the largest grid) equal the oracle's FP32 loop.
- The misfit gradient on the full 1000 x 2000 grid (padded 1100 x 2100) with
  8 shots, nt = 400 (of the workload's 4000): the batched gradient equals the
  sum of the single-shot gradients bit for bit; one shot's gradient is within
  2e-3 (relative L2) of the FP64 oracle gradient; and the dot-product test
  against the FP64 tangent-linear model holds to 1e-3 on the full grid."""
import numpy as np
import pytest

from wavefield_imaging_amd import _lib as L
from wavefield_imaging_amd import awefd, synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_batch_256_uneven_shots_three_steps_per_pass(orc):
    bat = synth.batch2d(nshot=256, nt=2000)
    ms = []
    for sh in bat["shots"]:
        m = awefd.Model("cart2d", sh["par"])
        m.set_model(sh["vel"])
        m.set_sources(sh["sou"])
        m.set_receivers(sh["rec"])
        m.set_wavelet(bat["wav"])
        ms.append(m)
    b = awefd.Batch(ms)
    b.set_tile(8, 0)
    b.set_depth(3)
    print("batch of 256 shots, nt 2000, three steps per pass", flush=True)
    b.run()
    sizes = [int(np.prod(m.padded)) for m in ms]
    check = {int(np.argmin(sizes)), int(np.argmax(sizes))}
    for s, m in enumerate(ms):
        tb, fb = m.traces(), m.field()
        tr, _ = m.run()                  # the shot's own single-grid loop
        assert np.array_equal(tb, tr), "shot %d traces" % s
        assert np.array_equal(fb, m.field()), "shot %d field" % s
        if s in check:
            sh = bat["shots"][s]
            print("oracle check of shot %d, padded %s" % (s, m.padded), flush=True)
            ref, ufin, _ = orc.run("cart2d", sh["par"], sh["vel"], sh["sou"], sh["rec"], bat["wav"])
            assert np.any(ref)
            assert np.array_equal(tb, ref), "shot %d vs oracle" % s
            assert np.array_equal(fb, ufin), "shot %d field vs oracle" % s
    b.close()
    for m in ms:
        m.close()


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_gradient_full_grid_batched_8_shots(orc):
    nt, nshot, K = 400, 8, 20
    case = synth.adjoint2d(nt=nt, nshot=nshot)
    par, rec, wav = case["par"], case["rec"], case["wav"]
    dev = torch.device("cuda", 0)
    nr = len(rec)
    # observed data: the true model, every shot
    mt = awefd.Model("cart2d", par)
    mt.set_model(case["true"])
    mt.set_receivers(rec)
    mt.set_wavelet(wav)
    dobs = torch.empty((nshot, nt, nr), dtype=torch.float32, device=dev)
    for s in range(nshot):
        mt.set_sources(case["shots"][s])
        torch.cuda.synchronize()
        L.check(mt.lib.awefd_run_dev(mt.h, dobs[s].data_ptr()))
    mt.close()
    ms = []
    for s in range(nshot):
        m = awefd.Model("cart2d", par)
        m.set_model(case["start"])
        m.set_sources(case["shots"][s])
        m.set_receivers(rec)
        m.set_wavelet(wav)
        ms.append(m)
    # single shots: forward with checkpoints, residual, adjoint
    shape = (par["n"][1], par["n"][0])
    res, single = [], []
    for s, m in enumerate(ms):
        tr = torch.empty((nt, nr), dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        L.check(m.lib.awefd_run_ckpt(m.h, tr.data_ptr(), K))
        r = (tr - dobs[s]).contiguous()
        g = torch.empty(shape, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        L.check(m.lib.awefd_adjoint_grad(m.h, r.data_ptr(), g.data_ptr()))
        res.append(r)
        single.append(g.cpu().numpy())
    gsum = single[0].copy()
    for g in single[1:]:
        gsum = gsum + g
    print("single-shot gradients done", flush=True)
    b = awefd.Batch(ms)
    b.set_tile(8, 0)
    b.run_ckpt(K)
    gb = torch.empty(shape, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    b.adjoint_grad([r.data_ptr() for r in res], gb.data_ptr())
    assert np.array_equal(gb.cpu().numpy(), gsum)
    print("batched gradient equals the sum of the single shots", flush=True)
    b.close()
    # shot 0 against the FP64 oracle gradient of the same residual weights
    r0 = res[0].cpu().numpy()
    g64, _, _ = orc.grad_f64(par, case["start"], case["shots"][0], rec, wav, res=r0)
    err = _rel(single[0], g64)
    print("  full-grid gradient vs FP64 oracle: rel L2 %.3e" % err, flush=True)
    assert err < 2e-3
    # dot-product test against the FP64 tangent-linear model, full grid
    rng = np.random.default_rng(17)
    s = 1.0 / case["start"].astype(np.float64) ** 2
    ds = rng.standard_normal(s.shape) * s
    dd = orc.tlm_f64(par, case["start"], case["shots"][0], rec, wav, ds)
    lhs, rhs = float(np.sum(dd * r0)), float(np.sum(single[0].astype(np.float64) * ds))
    print("  dot test: <J ds, r> %.6e  <ds, J^T r> %.6e" % (lhs, rhs), flush=True)
    assert abs(lhs - rhs) <= 1e-3 * abs(lhs)
    for m in ms:
        m.close()
