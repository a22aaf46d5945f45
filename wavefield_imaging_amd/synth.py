"""Seeded synthetic inputs for the five measured workloads.

There are no datasets: every model, wavelet and geometry is generated here
from a recorded seed, in the shapes the reference programs read (RSF order,
axis 1 fastest; coordinates as the sou=/rec= files hold them).

Workloads (sizes are the full ones; ``scale`` shrinks the grids for tests):
  cart2d   400 x 800 (z, x), h = 5 m, nt = 3000, nb = 50, one source
  cart3d   256 x 512 x 512 (z, x, y), nt = 2000, nb = 40, 128 x 128 receivers
  spher3d  400 x 512 x 512 (r, theta, phi), r in [3387.5, 3389.5] km, nb = 30
  batch2d  256 shots, nz ~ U[200, 600], nx ~ U[400, 1200], nt = 2000, nb = 50
  adjoint2d 1000 x 2000, nt = 4000, 64 shots, 2000 receivers per shot
"""
import numpy as np

V_AIR, V_ICE, V_BASE = 3.0e8, 1.68e8, 1.2e8
H = 5.0
DT = 6.67e-9
FREQ = 2.0e6


def ricker(nt, dt, freq=FREQ, delay=None):
    """Ricker wavelet peaking at ``freq``, delayed 1.5/freq, max-abs 1."""
    t0 = 1.5 / freq if delay is None else delay
    t = np.arange(nt) * dt - t0
    a = (np.pi * freq * t) ** 2
    w = (1.0 - 2.0 * a) * np.exp(-a)
    return (w / np.abs(w).max()).astype(np.float32)


def _gauss_kernel(sigma):
    r = int(np.ceil(4 * sigma))
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return k / k.sum()


def _smooth(a, sigma, axes):
    """Separable Gaussian filter (wrap-around boundaries)."""
    out = np.asarray(a, np.float64)
    k = _gauss_kernel(sigma)
    r = len(k) // 2
    for ax in axes:
        n = out.shape[ax]
        acc = np.zeros_like(out)
        for i, c in enumerate(k):
            acc += c * np.roll(out, i - r, axis=ax) if n > 1 else c * out
        out = acc
    return out


def rough_base(shape, mean, amp, corr, seed):
    """Gaussian-filtered white-noise surface rescaled to max-abs ``amp``."""
    rng = np.random.default_rng(seed)
    s = _smooth(rng.standard_normal(shape), corr, range(len(shape)))
    s = s / np.abs(s).max()
    return mean + amp * s


def layered_2d(nz, nx, seed, air=100, base_mean=300, amp=10.0, corr=40.0):
    """Three-layer model, array (nx, nz) = RSF n1=z, n2=x."""
    zb = rough_base((nx,), base_mean, amp, corr, seed)
    z = np.arange(nz)[None, :]
    v = np.full((nx, nz), V_ICE, np.float32)
    v[:, :air] = V_AIR
    v[z >= zb[:, None]] = V_BASE
    v[:, :air] = V_AIR
    return v


def layered_3d(nz, nx, ny, seed, air=100, base_mean=180, amp=10.0, corr=40.0):
    """Three-layer model, array (ny, nx, nz) = RSF n1=z, n2=x, n3=y."""
    zb = rough_base((ny, nx), base_mean, amp, corr, seed)
    z = np.arange(nz)[None, None, :]
    v = np.full((ny, nx, nz), V_ICE, np.float32)
    v[z >= zb[:, :, None]] = V_BASE
    v[:, :, :air] = V_AIR
    return v


def cart2d(scale=1.0, nt=None, seed=1234):
    nz, nx = max(int(400 * scale), 24), max(int(800 * scale), 24)
    nb = max(int(50 * scale), 4)
    nt = nt or 3000
    air = max(int(100 * scale), 2)
    vel = layered_2d(nz, nx, seed, air=air, base_mean=int(300 * scale), amp=10 * scale,
                     corr=max(40 * scale, 1.0))
    src_row, src_col = max(int(20 * scale), 1), nx // 2
    sou = np.array([[src_col * H, src_row * H]], np.float32)        # (x, z)
    rec = np.stack([np.arange(nx) * H, np.full(nx, src_row * H)], 1).astype(np.float32)
    par = dict(n=[nz, nx], o=[0.0, 0.0], d=[H, H], nt=nt, dt=DT, nb=nb, dabc=True,
               free=False, cden=True, fdorder=4, expl=False)
    return dict(kind="cart2d", par=par, vel=vel, sou=sou, rec=rec, wav=ricker(nt, DT), seed=seed)


def cart3d(scale=1.0, nt=None, seed=4321):
    nz = max(int(256 * scale), 16)
    nx = ny = max(int(512 * scale), 16)
    nb = max(int(40 * scale), 4)
    nt = nt or 2000
    air = max(int(100 * scale), 2)
    vel = layered_3d(nz, nx, ny, seed, air=air, base_mean=int(180 * scale), amp=10 * scale,
                     corr=max(40 * scale, 1.0))
    d15 = max(int(15 * scale), 1)
    sou = np.array([[nx // 2 * H, ny // 2 * H, d15 * H]], np.float32)   # (x, y, z)
    nrl = max(int(128 * scale), 2)
    step = max(nx // nrl, 1)
    ix, iy = np.meshgrid(np.arange(nrl) * step, np.arange(nrl) * step, indexing="ij")
    rec = np.stack([ix.ravel() * H, iy.ravel() * H, np.full(ix.size, d15 * H)], 1).astype(np.float32)
    par = dict(n=[nz, nx, ny], o=[0.0] * 3, d=[H] * 3, nt=nt, dt=DT, nb=nb, dabc=True,
               free=False, cden=True, fdorder=4, expl=False)
    return dict(kind="cart3d", par=par, vel=vel, sou=sou, rec=rec, wav=ricker(nt, DT), seed=seed)


def spher3d(scale=1.0, nt=None, seed=777):
    """Spherical shell: r from 3387.5 km, dr = 5 m, r*dtheta = 5 m at mid radius,
    theta and phi centred on pi/4 (acoustic_sp/SConstruct:24)."""
    nr = max(int(400 * scale), 24)
    nth = nph = max(int(512 * scale), 24)
    nb = max(int(30 * scale), 4)
    nt = nt or 2000
    orr, dr = 3387500.0, H
    rmid = orr + dr * 200
    dth = H / rmid
    oth = np.pi / 4 - (nth // 2) * dth
    oph = np.pi / 4 - (nph // 2) * dth
    top = max(int(20 * scale), 2)
    rng = np.random.default_rng(seed)
    base_depth = rough_base((nph, nth), int(300 * scale), 10 * scale, max(40 * scale, 1.0),
                            int(rng.integers(1 << 30)))
    ir = np.arange(nr)[None, None, :]
    depth = (nr - 1) - ir
    vel = np.full((nph, nth, nr), V_ICE, np.float32)                    # (phi, theta, r)
    vel[depth >= base_depth[:, :, None]] = V_BASE
    vel[:, :, nr - top:] = V_AIR
    isr = nr - 1 - max(int(10 * scale), 1)
    sou = np.array([[orr + isr * dr, oph + (nph // 2) * dth, oth + (nth // 2) * dth]], np.float32)
    nrl = max(int(256 * scale), 2)
    st = max(nth // nrl, 1)
    ith, iph = np.meshgrid(np.arange(nrl) * st, np.arange(nrl) * st, indexing="ij")
    rec = np.stack([np.full(ith.size, orr + isr * dr), oph + iph.ravel() * dth,
                    oth + ith.ravel() * dth], 1).astype(np.float32)       # (r, phi, theta)
    par = dict(n=[nr, nth, nph], o=[orr, oth, oph], d=[dr, dth, dth], nt=nt, dt=DT, nb=nb,
               expl=False, free=False)
    return dict(kind="spher3d", par=par, vel=vel, sou=sou, rec=rec, wav=ricker(nt, DT), seed=seed)


def spher2d(scale=1.0, nt=None, seed=555):
    """Polar (r, theta) section of the spherical workload."""
    nr = max(int(400 * scale), 24)
    nth = max(int(512 * scale), 24)
    nb = max(int(30 * scale), 4)
    nt = nt or 2000
    orr, dr = 3387500.0, H
    dth = H / (orr + dr * 200)
    oth = np.pi / 4 - (nth // 2) * dth
    top = max(int(20 * scale), 2)
    base_depth = rough_base((nth,), int(300 * scale), 10 * scale, max(40 * scale, 1.0), seed)
    ir = np.arange(nr)[None, :]
    vel = np.full((nth, nr), V_ICE, np.float32)                          # (theta, r)
    vel[((nr - 1) - ir) >= base_depth[:, None]] = V_BASE
    vel[:, nr - top:] = V_AIR
    isr = nr - 1 - max(int(10 * scale), 1)
    sou = np.array([[orr + isr * dr, oth + (nth // 2) * dth]], np.float32)   # (r, theta)
    rec = np.stack([np.full(nth, orr + isr * dr), oth + np.arange(nth) * dth], 1).astype(np.float32)
    par = dict(n=[nr, nth], o=[orr, oth], d=[dr, dth], nt=nt, dt=DT, nb=nb, expl=False,
               free=False)
    return dict(kind="spher2d", par=par, vel=vel, sou=sou, rec=rec, wav=ricker(nt, DT), seed=seed)


def batch2d(nshot=256, nt=2000, seed=2024, scale=1.0):
    """Independent 2-D shots on grids of uneven size."""
    rng = np.random.default_rng(seed)
    shots = []
    for s in range(nshot):
        nz = int(rng.integers(int(200 * scale), int(600 * scale) + 1))
        nx = int(rng.integers(int(400 * scale), int(1200 * scale) + 1))
        mseed = int(rng.integers(1 << 30))
        col = int(rng.integers(nx // 8, nx - nx // 8))
        vel = layered_2d(nz, nx, mseed, air=min(100, nz // 4), base_mean=int(nz * 0.75),
                         amp=10.0, corr=40.0)
        row = min(20, nz // 8)
        sou = np.array([[col * H, row * H]], np.float32)
        rec = np.stack([np.arange(nx) * H, np.full(nx, row * H)], 1).astype(np.float32)
        par = dict(n=[nz, nx], o=[0.0, 0.0], d=[H, H], nt=nt, dt=DT, nb=50, dabc=True,
                   free=False, cden=True, fdorder=4, expl=False)
        shots.append(dict(kind="cart2d", par=par, vel=vel, sou=sou, rec=rec, seed=mseed))
    return dict(shots=shots, wav=ricker(nt, DT))


def adjoint2d(scale=1.0, nt=None, nshot=64, seed=4096):
    """Gradient workload: one 1000 x 2000 (z, x) grid, nb = 50, nt = 4000,
    ``nshot`` sources evenly spaced along row 20, 2000 receivers per shot
    (every column of row 20).  True model: the three-layer rough-base model
    (air in the top 100 rows, base at mean row 750, 10-node amplitude,
    40-node correlation).  Starting model: the true one smoothed with a
    20-node Gaussian (edge-replicated borders)."""
    nz, nx = max(int(1000 * scale), 24), max(int(2000 * scale), 24)
    nb = max(int(50 * scale), 4)
    nt = nt or 4000
    true = layered_2d(nz, nx, seed, air=max(int(100 * scale), 2), base_mean=int(750 * scale),
                      amp=10 * scale, corr=max(40 * scale, 1.0))
    sig = max(20 * scale, 1.0)
    r = int(np.ceil(4 * sig))
    start = _smooth(np.pad(true.astype(np.float64), r, mode="edge"), sig, (0, 1))[r:-r, r:-r]
    row = max(int(20 * scale), 1)
    cols = (np.arange(nshot) + 0.5) * nx / nshot
    shots = [np.array([[c * H, row * H]], np.float32) for c in cols]
    rec = np.stack([np.arange(nx) * H, np.full(nx, row * H)], 1).astype(np.float32)
    par = dict(n=[nz, nx], o=[0.0, 0.0], d=[H, H], nt=nt, dt=DT, nb=nb, dabc=True,
               free=False, cden=True, fdorder=4, expl=False)
    return dict(kind="cart2d", par=par, true=true, start=start.astype(np.float32), shots=shots, rec=rec,
                wav=ricker(nt, DT), ckpt=64, seed=seed)
