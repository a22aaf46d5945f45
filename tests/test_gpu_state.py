"""State across calls: a batch whose models change between runs, and an
autograd module called twice before one backward.  Both must give what
fresh single-model runs give, bit for bit."""
import numpy as np
import pytest

from wavefield_imaging_amd import _lib as L
from wavefield_imaging_amd import awefd, synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _shot(nz, nx, nt, seed, xsrc, nb=8):
    vel = synth.layered_2d(nz, nx, seed=seed, air=max(nz // 5, 2), base_mean=int(nz * 0.65), amp=2.0, corr=5.0)
    sou = np.float32([[xsrc, 40.0]])
    xs = np.arange(1, nx - 1, 3)
    rec = np.stack([xs * 5.0 + 1.3, np.full(len(xs), 8.4 * 5.0)], 1).astype(np.float32)
    par = dict(n=[nz, nx], o=[0.0, 0.0], d=[5.0, 5.0], nt=nt, dt=synth.DT, nb=nb, dabc=True)
    return par, vel, sou, rec, synth.ricker(nt, synth.DT, freq=8e6)


def _model(par, vel, sou, rec, wav):
    m = awefd.Model("cart2d", par)
    m.set_model(vel)
    m.set_sources(sou)
    m.set_receivers(rec)
    m.set_wavelet(wav)
    return m


def _single(par, vel, sou, rec, wav, ckpt=0):
    """traces of a fresh model's checkpointed forward run"""
    m = _model(par, vel, sou, rec, wav)
    tr = torch.empty((m.nt, m.nr), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    L.check(m.lib.awefd_run_ckpt(m.h, tr.data_ptr(), ckpt))
    out = tr.cpu().numpy()
    m.close()
    return out


@pytest.mark.parametrize("depth", [1, 3])
def test_batch_follows_model_changes(depth):
    """Run a batch, then change one shot's velocity (dabc=y: its sponge tables
    are rebuilt), its sources and its receivers, and run again: every shot
    equals a fresh single run of its current set-up."""
    nt = 120
    shots = [_shot(40 + 8 * i, 60 + 10 * i, nt, seed=i, xsrc=60.0 + 20 * i) for i in range(4)]
    ms = [_model(*sh) for sh in shots]
    b = awefd.Batch(ms)
    b.set_tile(8, 0)
    if depth > 1:
        b.set_depth(depth)
    b.run()
    for s, sh in enumerate(shots):
        np.testing.assert_array_equal(ms[s].traces(), _single(*sh), err_msg="shot %d before" % s)
    # change shot 2: a faster velocity, another source, fewer receivers
    par, vel, sou, rec, wav = shots[2]
    vel2 = (vel * np.float32(1.1)).astype(np.float32)
    sou2 = np.float32([[150.0, 55.0]])
    rec2 = rec[::2].copy()
    ms[2].set_model(vel2)
    ms[2].set_sources(sou2)
    ms[2].set_receivers(rec2)
    shots[2] = (par, vel2, sou2, rec2, wav)
    for rep in range(2):
        print("batch after set_model, depth %d, run %d" % (depth, rep), flush=True)
        b.run()
        for s, sh in enumerate(shots):
            np.testing.assert_array_equal(ms[s].traces(), _single(*sh), err_msg="shot %d rep %d" % (s, rep))
    b.close()
    for m in ms:
        m.close()


def test_batch_checkpointed_forward_follows_model_changes():
    """batch_run_ckpt, set_model on a dabc=y shot, batch_run_ckpt again: the
    traces and the summed gradient equal fresh single-shot runs."""
    nt = 100
    shots = [_shot(48, 70, nt, seed=10 + i, xsrc=80.0 + 60 * i) for i in range(3)]
    ms = [_model(*sh) for sh in shots]
    b = awefd.Batch(ms)
    b.run_ckpt(10)
    par, vel, sou, rec, wav = shots[1]
    vel2 = vel.copy()
    vel2[:, 20:30] = np.float32(1.5e8)
    ms[1].set_model(vel2)
    shots[1] = (par, vel2, sou, rec, wav)
    with pytest.raises(L.AwefdError, match="checkpoints"):
        dres = [torch.zeros((nt, m.nr), dtype=torch.float32, device="cuda") for m in ms]
        g = torch.zeros((70, 48), dtype=torch.float32, device="cuda")
        b.adjoint_grad([t.data_ptr() for t in dres], g.data_ptr())
    print("checkpointed batch after set_model", flush=True)
    b.run_ckpt(10)
    rng = np.random.default_rng(1)
    res = [rng.standard_normal((nt, m.nr)).astype(np.float32) for m in ms]
    dres = [torch.from_numpy(r).cuda() for r in res]
    gb = torch.zeros((70, 48), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    b.adjoint_grad([t.data_ptr() for t in dres], gb.data_ptr())
    gsum = None
    for s, sh in enumerate(shots):
        np.testing.assert_array_equal(ms[s].traces(), _single(*sh, ckpt=10), err_msg="shot %d" % s)
        m = _model(*sh)
        tr = torch.empty((nt, m.nr), dtype=torch.float32, device="cuda")
        g = torch.empty((70, 48), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        L.check(m.lib.awefd_run_ckpt(m.h, tr.data_ptr(), 10))
        L.check(m.lib.awefd_adjoint_grad(m.h, dres[s].data_ptr(), g.data_ptr()))
        gs = g.cpu().numpy()
        gsum = gs if gsum is None else gsum + gs
        m.close()
    np.testing.assert_array_equal(gb.cpu().numpy(), gsum)
    b.close()
    for m in ms:
        m.close()


def test_autograd_two_forwards_one_backward(orc):
    """loss = f(op(s1)) + f(op(s2)) with one module: each slowness gets the
    gradient of its own forward (the module rebuilds the checkpoints of the
    earlier call), bit for bit equal to separate forward/backward pairs."""
    from wavefield_imaging_amd.autograd import Traces

    par, vel, sou, rec, wav = _shot(40, 56, 180, seed=5, xsrc=130.0)
    dobs, _, _ = orc.run("cart2d", par, vel, sou, rec, wav)
    dobs = torch.from_numpy(dobs).cuda()
    s1 = 1.0 / vel.astype(np.float64) ** 2
    s2 = s1.copy()
    s2[:, 20:26] *= 0.8
    op = Traces(par, sou, rec, wav, ckpt=12)

    def leaf(s):
        return torch.tensor(s, dtype=torch.float32, device="cuda", requires_grad=True)

    sep = []
    for s in (s1, s2):
        t = leaf(s)
        (0.5 * ((op(t) - dobs) ** 2).sum()).backward()
        sep.append(t.grad.cpu().numpy())
    a, c = leaf(s1), leaf(s2)
    loss = 0.5 * ((op(a) - dobs) ** 2).sum() + 0.5 * ((op(c) - dobs) ** 2).sum()
    loss.backward()
    np.testing.assert_array_equal(a.grad.cpu().numpy(), sep[0])
    np.testing.assert_array_equal(c.grad.cpu().numpy(), sep[1])
    assert not np.array_equal(sep[0], sep[1])
    op.close()
