#!/usr/bin/env python3
"""Benchmark: the acoustic FD time loop on one MI355X.

Default workload (the headline line): plain 3-D cartesian forward modelling,
256 x 512 x 512 (z, x, y) model, h = 5 m, 40-node sponge on every face (padded
grid 336 x 592 x 592 = 117,755,904 nodes), nt = 2000 steps of dt = 6.67e-9 s,
three-layer rough-base velocity, one 2 MHz Ricker source, 128 x 128 receivers,
sponge + one-way ABC (dabc=y), FP32.

The default run also measures, under "also":
  cart2d    400 x 800 (z, x), nb = 50 (padded 500 x 900), nt = 3000, one
            source, 800 receivers
  spher3d   the spherical shell, 400 x 512 x 512 (r, theta, phi), r from
            3387.5 km, dr = 5 m, r*dtheta = 5 m at mid radius, 30-node sponge
            (padded 460 x 572 x 572 = 150,504,640 nodes), nt = 2000, one source
            10 nodes below the top, 256 x 256 receivers
  batch2d   256 independent 2-D shots of uneven size, batched graph vs the
            per-shot loop, in shots/s
  adjoint2d 64 shots on a 1000 x 2000 grid, nt = 4000, checkpoints every 64
            steps: misfit gradients per second, batched graphs vs the per-shot
            loop

One benchmark step = one complete shot (all nt time steps) with every input
already resident in HBM; ms_per_step is therefore milliseconds per shot and
ms_per_time_step the same per time step.  value = node-updates per second
over the timed shots (Gcell-updates/s).

Roofline: algorithmic bytes (16 per node update: read the previous field, the
current field and the velocity term, write the new field) over the dominant
kernel's time.  3-D: the kernel's mean duration from HIP events around every
launch on its own stream, inside the timed region.  cart2d runs as one
replayed graph: its time per step (kernel plus the graph's launch gaps, an
upper bound on the kernel) over the timed region.  batch2d and adjoint2d run
as graphs too: real bytes of their passes over the timed graphs (20 bytes per
node for a three-step pass; for a gradient the bytes of its forward,
checkpoint, recompute and adjoint passes, 32.9 per node and step with
three-step forward and recompute passes and two-step adjoint passes, 60
with one step per launch).

cpu_baseline: the reference CPU path restated in C (oracle/liborc_ref.so:
the 4-wide SSE step with schedule(static,1) at -O3, OpenMP over r for the
spherical solve; bit-identical to the checker build), timed on the host cores
on a bounded sample of the same workload.

--calibrate-hbm MIB runs only a plain device copy of MIB MiB (for scaling the
rocprofv3 FETCH_SIZE / WRITE_SIZE counters) and exits.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = 16

WORKLOADS = {
    "cart2d": dict(kernel="k_cart_step<2,2,true,false,true>",
                   traffic="cart2d_step_traffic.json",
                   desc="cart2d 400x800 (z,x) h=5m nb=50 nt=%d dabc=y fdorder=4, padded 500x900, "
                        "1 source, 800 receivers"),
    "cart3d": dict(kernel="k_cart_col<2,true,false>",
                   traffic="cart3d_step_traffic.json",
                   desc="cart3d 256x512x512 (z,x,y) h=5m nb=40 nt=%d dabc=y fdorder=4, "
                        "padded 336x592x592, 1 source, 128x128 receivers"),
    "spher3d": dict(kernel="k_spher_col3",
                    traffic="spher3d_step_traffic.json",
                    desc="spher3d 400x512x512 (r,theta,phi) r0=3387.5km dr=5m nb=30 nt=%d "
                         "expl=n, padded 460x572x572, 1 source, 256x256 receivers"),
}


def _traffic_file(name):
    if not name:
        return {}
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return {}
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _traffic(name):
    """Calibrated HBM bytes per launch of the step kernel from the committed
    counter summary."""
    return _traffic_file(name).get("hbm_bytes_per_launch")


def _traffic_scaled(name, algorithmic):
    """Calibrated HBM bytes of `algorithmic` bytes of work, from the
    traffic/algorithmic ratio a committed counter summary measured for the
    same kernel on a smaller sample (the batch kernels' traffic per cell does
    not depend on the number of shots)."""
    ratio = _traffic_file(name).get("traffic_over_algorithmic")
    return None if ratio is None else int(round(ratio * algorithmic))


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    return O


def _loop_seconds(O):
    import ctypes as C
    f = O.lib(True).orc_last_loop_seconds
    f.restype = C.c_double
    return f()


def _cores(O):
    return int(O.lib(True).orc_threads())


def cpu_baseline(case, steps):
    """The reference-form CPU loop on `steps` time steps of `case`."""
    import numpy as np

    O = _oracle()
    par = dict(case["par"], nt=steps)
    O.run(case["kind"], par, case["vel"], case["sou"], case["rec"], case["wav"][:steps], final=False,
          ref_form=True)
    secs = _loop_seconds(O)
    cells = int(np.prod([n + 2 * par["nb"] for n in par["n"]]))
    return dict(value=round(cells * steps / secs / 1e9, 4), unit="Gcell-updates/s", cores=_cores(O), kind="port",
                sample="%s padded %s, %d of %d time steps, %.1f s loop wall time (SSE step, -O3)"
                       % (case["kind"], "x".join(str(n + 2 * par["nb"]) for n in par["n"]),
                          steps, case["par"]["nt"], secs))


def cpu_baseline_batch(bat, nshot_sample):
    """The per-shot loop of the reference-form 2-D program over the first
    shots of the batch, full length."""
    import numpy as np

    O = _oracle()
    secs, cells = 0.0, 0
    for sh in bat["shots"][:nshot_sample]:
        O.run("cart2d", sh["par"], sh["vel"], sh["sou"], sh["rec"], bat["wav"], final=False, ref_form=True)
        secs += _loop_seconds(O)
        cells += int(np.prod([n + 2 * sh["par"]["nb"] for n in sh["par"]["n"]]))
    nt = bat["shots"][0]["par"]["nt"]
    return dict(value=round(nshot_sample / secs, 4), unit="shots/s", cores=_cores(O), kind="port",
                gcells_per_s=round(cells * nt / secs / 1e9, 4),
                sample="first %d of %d shots, all %d steps each, %.1f s of loop wall time (SSE step, -O3)"
                       % (nshot_sample, len(bat["shots"]), nt, secs))


def cpu_baseline_grad(case, fwd_steps=200, grad_steps=100):
    """The reference has no gradient.  Timed: one shot's reference-form
    forward loop (the reference's CPU path) and the FP64 CPU gradient of the
    same scheme (forward, adjoint and gradient), each on a truncated run of
    the full grid, scaled to nt; value = gradients/s of the FP64 CPU
    gradient."""
    import numpy as np

    O = _oracle()
    par, nt = case["par"], case["par"]["nt"]
    sou, rec = case["shots"][0], case["rec"]
    p1 = dict(par, nt=fwd_steps)
    O.run("cart2d", p1, case["start"], sou, rec, case["wav"][:fwd_steps], final=False, ref_form=True)
    fwd_s = _loop_seconds(O) * nt / fwd_steps
    p2 = dict(par, nt=grad_steps)
    t0 = time.perf_counter()
    O.grad_f64(p2, case["start"], sou, rec, case["wav"][:grad_steps])
    g_s = (time.perf_counter() - t0) * nt / grad_steps
    cells = int(np.prod([n + 2 * par["nb"] for n in par["n"]]))
    return dict(value=round(1.0 / g_s, 6), unit="gradients/s", cores=_cores(O), kind="port",
                forward_shots_per_s=round(1.0 / fwd_s, 5),
                forward_gcells_per_s=round(cells * nt / fwd_s / 1e9, 4),
                sample="one shot on the full %s padded grid: the reference-form forward loop over %d of %d steps "
                       "and the FP64 gradient over %d steps, both on %d threads and scaled to %d steps"
                       % ("x".join(str(n + 2 * par["nb"]) for n in par["n"]), fwd_steps, nt, grad_steps,
                          _cores(O), nt))


def measure(workload, nt, steps, warmup):
    """Time `steps` full shots of `workload` after `warmup` untimed ones."""
    import numpy as np

    from wavefield_imaging_amd import awefd, synth

    w = WORKLOADS[workload]
    case = getattr(synth, workload)(nt=nt)
    m = awefd.Model(workload, case["par"])
    m.set_model(case["vel"])
    m.set_sources(case["sou"])
    m.set_receivers(case["rec"])
    m.set_wavelet(case["wav"])
    cells = int(np.prod(m.padded))
    print("bench: %s padded %s, nt=%d, %d receivers" % (workload, m.padded, nt, m.nr),
          file=sys.stderr, flush=True)
    for _ in range(warmup):
        m.run_dev()
    # 3-D: the kernel's own events bracket every launch inside the timed
    # region.  2-D: the loop runs as one captured graph (a step is a few
    # microseconds, launch overhead would dominate); its time per step over
    # the timed region bounds the kernel's.
    eager_timed = m.ndim == 3
    m.profile(eager_timed)
    t0 = time.perf_counter()
    for _ in range(steps):
        m.run_dev()
    t1 = time.perf_counter()
    secs = t1 - t0
    if eager_timed:
        launches, kern_ms = m.kernel_stats()
        avg_ms = kern_ms / max(launches, 1)
        timing = "HIP events around every launch of the timed shots"
    else:
        launches = nt * steps
        avg_ms = secs / launches * 1e3
        timing = "replayed graph: time per step over the timed shots (kernel plus in-graph launch gaps)"
    m.profile(False)
    m.close()
    achieved = BYTES_PER_UPDATE * cells / (avg_ms * 1e-3) / 1e9
    out = dict(workload=w["desc"] % nt, cells_per_step=cells, time_steps_per_shot=nt,
               value=round(cells * nt * steps / secs / 1e9, 3),
               ms_per_step=round(secs / steps * 1e3, 3),
               ms_per_time_step=round(secs / steps / nt * 1e3, 5),
               roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": _traffic(w["traffic"]), "kernel": w["kernel"],
                         "kernel_mean_ms": round(avg_ms, 5), "launches": launches, "timing": timing})
    return out, case


def measure_batch(nt, nshot=256, steps=1, tune_steps=20, per_shot=True, fixed=None, cpu_shots=0):
    """256 independent 2-D shots of uneven size (nz ~ U[200,600], nx ~
    U[400,1200], nb=50): the batched graph-captured run (autotuned tile,
    shot grouping and steps per launch, or `fixed` = (rx, group_mb, depth))
    and the plain per-shot loop of the single-grid kernel, in shots/s."""
    import numpy as np

    from wavefield_imaging_amd import awefd, synth

    bat = synth.batch2d(nshot=nshot, nt=nt)
    ms = []
    for sh in bat["shots"]:
        m = awefd.Model("cart2d", sh["par"])
        m.set_model(sh["vel"])
        m.set_sources(sh["sou"])
        m.set_receivers(sh["rec"])
        m.set_wavelet(bat["wav"])
        ms.append(m)
    cells = int(sum(np.prod(m.padded) for m in ms))
    print("bench: batch2d %d shots, %.3g padded cells, nt=%d" % (nshot, cells, nt), file=sys.stderr, flush=True)
    b = awefd.Batch(ms)
    if fixed:
        rx, grp, depth = fixed
        b.set_tile(rx, grp)
        b.set_depth(depth)
        tune_ms = {}
    else:
        rx, grp, depth, tune_ms = b.tune_depth(steps=tune_steps)
    b.run()                                   # warm-up (graph capture + launch)
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run()
    t1 = time.perf_counter()
    secs = (t1 - t0) / steps
    batched = nshot / secs
    # bytes of the run: passes of `depth` steps read the two input fields and
    # V and write the last two fields (20 B/node; one step: 16 B/node)
    npass, nsingle = (nt - 1) // depth if depth > 1 else 0, nt - ((nt - 1) // depth) * depth if depth > 1 else nt
    bytes_run = (npass * 20 + nsingle * 16) * cells
    achieved = bytes_run / secs / 1e9
    if depth == 3:
        kern = "k_cart_batch_t3<2,true,%d,%d>" % (rx, 32 // rx)     # 32 x rows per block
    else:
        kern = {1: "k_cart_batch2<2,true,%d,4>", 2: "k_cart_batch_t2<2,true,%d,4>"}[depth] % rx
    out = dict(workload="batch2d %d shots nz~U[200,600] nx~U[400,1200] nb=50 nt=%d, %d padded cells"
                        % (nshot, nt, cells),
               shots_per_s=round(batched, 3), gcells_per_s=round(cells * nt / secs / 1e9, 3),
               s_per_batch=round(secs, 4), ms_per_time_step=round(secs / nt * 1e3, 5), tile_rx=rx, group_mb=grp,
               steps_per_launch=depth,
               roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": _traffic_scaled("batch2d_pass_traffic.json", 20 * cells) if depth == 3 else None,
                         "traffic_per": "one pass of three steps over all shots", "kernel": kern,
                         "bytes_per_run": bytes_run,
                         "timing": "the timed graph: %d passes of %d steps at 20 B/node and %d single steps at "
                                   "16 B/node over the run's wall time (fields of one shot fit the 256 MB "
                                   "Infinity Cache, so HBM is not the only level serving them)"
                                   % (npass, depth, nsingle)},
               tune_ms_per_step={"d%d_rx%d_g%d" % k: round(v, 4) for k, v in tune_ms.items()})
    b.close()
    if per_shot:
        # warm-up of the single-grid kernel, on the per-launch path (a
        # profiled run is never captured); each shot then runs once, as a
        # plain loop over shots does
        ms[0].profile(True)
        ms[0].run_dev()
        ms[0].profile(False)
        t0 = time.perf_counter()
        for m in ms:
            m.run_dev()
        t1 = time.perf_counter()
        loop = nshot / (t1 - t0)
        out.update(per_shot_loop_shots_per_s=round(loop, 3), speedup_vs_per_shot_loop=round(batched / loop, 2))
    for m in ms:
        m.close()
    if cpu_shots:
        out["cpu_baseline"] = cpu_baseline_batch(bat, cpu_shots)
    return out


def grad_bytes_per_node(nt, K, depth, adepth=2):
    """Algorithmic bytes per padded node of one shot's checkpointed gradient,
    following the batch driver's schedule: the forward run (a single step
    reads the previous and current field and (v dt)^2 and writes one field,
    16 B; a three-step pass writes its last two fields, 20 B), the checkpoint
    copies (two fields read and written, 16 B each), the segment recompute
    (16 B a step; a three-step pass keeps all three fields, 24 B) and the
    adjoint steps (one step reads alpha_{n+1}, alpha_{n+2}, V, c_n and dJ/dV
    and writes alpha_n and dJ/dV, 28 B; a two-step pass also reads c_{n-1}
    and writes alpha_{n-1}, 36 B).  depth 3: passes while one ends inside its
    segment (and, in the forward run, a step follows it), single steps after;
    adepth 2: pairs of adjoint steps inside each segment, a single step for
    an odd one left."""
    fwd = rcp = adj = 0
    nck = (nt + K - 1) // K
    for k in range(nck):
        n0, n1 = k * K, min(k * K + K, nt)
        it = n0
        while depth == 3 and it + 3 <= n1 and it + 3 < nt:
            fwd += 20
            it += 3
        fwd += 16 * (n1 - it)
        n = n0
        while depth == 3 and n + 3 <= n1 - 1:
            rcp += 24
            n += 3
        rcp += 16 * max(n1 - 1 - n, 0)
        steps = n1 - n0
        adj += 36 * (steps // 2) + 28 * (steps % 2) if adepth == 2 else 28 * steps
    return dict(forward=fwd, checkpoints=16 * nck, recompute=rcp, adjoint=adj)


def measure_grad(nt=4000, nshot=64, sample=8, batched=True, rx=8, group_mb=0, cpu=False):
    """Misfit gradients w.r.t. squared slowness, 1000 x 2000 grid, nb=50,
    nt=4000, checkpoints every 64 steps, 2000 receivers per shot: the per-shot
    loop (forward with checkpoints, residual, adjoint) timed on `sample`
    shots, and the batched graphs (every shot's checkpointed forward in one
    graph, residuals, every shot's backward pass in one graph summing the
    gradient), in gradients/s."""
    import ctypes as C

    import numpy as np
    import torch

    from wavefield_imaging_amd import _lib as L
    from wavefield_imaging_amd import awefd, synth

    case = synth.adjoint2d(nt=nt, nshot=nshot)
    par, rec, wav, K = case["par"], case["rec"], case["wav"], case["ckpt"]
    dev = torch.device("cuda", 0)
    nr = len(rec)
    mt = awefd.Model("cart2d", par)
    mt.set_model(case["true"])
    mt.set_receivers(rec)
    mt.set_wavelet(wav)
    cells = int(np.prod(mt.padded))
    print("bench: adjoint2d padded %s, nt=%d, %d shots, %d receivers" % (mt.padded, nt, nshot, nr),
          file=sys.stderr, flush=True)
    nobs = nshot if batched else sample
    dobs = torch.empty((nobs, nt, nr), dtype=torch.float32, device=dev)
    for s in range(nobs):
        mt.set_sources(case["shots"][s])
        torch.cuda.synchronize()
        L.check(mt.lib.awefd_run_dev(mt.h, C.c_void_p(dobs[s].data_ptr())))
    mt.close()
    m = awefd.Model("cart2d", par)
    m.set_model(case["start"])
    m.set_receivers(rec)
    m.set_wavelet(wav)
    tr = torch.empty((nt, nr), dtype=torch.float32, device=dev)
    g = torch.empty((par["n"][1], par["n"][0]), dtype=torch.float32, device=dev)
    gsum = torch.zeros_like(g)

    def one(s):
        m.set_sources(case["shots"][s])
        L.check(m.lib.awefd_run_ckpt(m.h, C.c_void_p(tr.data_ptr()), K))
        res = tr - dobs[s]
        torch.cuda.current_stream().synchronize()
        L.check(m.lib.awefd_adjoint_grad(m.h, C.c_void_p(res.data_ptr()), C.c_void_p(g.data_ptr())))
        gsum.add_(g)

    one(0)                                     # warm-up: buffers, lists
    gsum.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(sample):
        one(s)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loop = sample / (t1 - t0)
    out = dict(workload="adjoint2d 1000x2000 nb=50 nt=%d ckpt=%d, %d shots x %d receivers, padded %s"
                        % (nt, K, nshot, nr, "x".join(map(str, m.padded[:2]))),
               per_shot_loop_gradients_per_s=round(loop, 3),
               per_shot_loop_sample="%d of %d shots, %.2f s" % (sample, nshot, t1 - t0),
               per_shot_gcells_per_s_per_propagation=round(3 * cells * nt * sample / (t1 - t0) / 1e9, 3))
    m.close()
    del tr, g, gsum
    if batched:
        ms = []
        for s in range(nshot):
            mm = awefd.Model("cart2d", par)
            mm.set_model(case["start"])
            mm.set_sources(case["shots"][s])
            mm.set_receivers(rec)
            mm.set_wavelet(wav)
            ms.append(mm)
        b = awefd.Batch(ms)
        b.set_tile(rx, group_mb)
        trs = [mm.traces_tensor() for mm in ms]
        res = [torch.empty((nt, nr), dtype=torch.float32, device=dev) for _ in range(nshot)]
        gb = torch.empty((par["n"][1], par["n"][0]), dtype=torch.float32, device=dev)

        def batch_once():
            b.run_ckpt(K)
            for s in range(nshot):
                torch.sub(trs[s], dobs[s], out=res[s])
            torch.cuda.current_stream().synchronize()
            b.adjoint_grad([r.data_ptr() for r in res], gb.data_ptr())

        batch_once()                           # warm-up: buffers, graph capture
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch_once()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        bat = nshot / (t1 - t0)
        # the batch driver's passes: three steps where the tile allows it
        # (rx 4 or 8) unless AWEFD_CKPT_DEPTH=1 asks for one
        depth = 3 if rx in (4, 8) and os.environ.get("AWEFD_CKPT_DEPTH") != "1" else 1
        adepth = 1 if os.environ.get("AWEFD_ADJ_DEPTH") == "1" else 2
        parts = grad_bytes_per_node(nt, K, depth, adepth)
        bytes_grad = sum(parts.values()) * cells
        achieved = bytes_grad * nshot / (t1 - t0) / 1e9
        out.update(batched_gradients_per_s=round(bat, 3), batched_s_per_gradient_sum=round(t1 - t0, 3),
                   ms_per_time_step=round((t1 - t0) / nshot / nt * 1e3, 5),
                   batched_tile_rx=rx, batched_group_mb=group_mb,
                   batched_gcells_per_s_per_propagation=round(3 * cells * nt * nshot / (t1 - t0) / 1e9, 3),
                   speedup_vs_per_shot_loop=round(bat / loop, 2),
                   roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(achieved / HBM_PEAK_GBS, 4),
                             "traffic": _traffic_scaled("adjoint2d_traffic.json", bytes_grad * nshot),
                             "traffic_per": "one batched gradient of all shots",
                             "kernel": "%s with %s" % (
                                 "k_adj2d_batch2<2> (two adjoint steps per pass)" if adepth == 2 else
                                 "k_adj2d_batch<2>",
                                 "k_cart_batch_t3 (forward; recompute, all three fields kept)" if depth == 3 else
                                 "k_cart_batch2 (forward, recompute)"),
                             "steps_per_pass": depth,
                             "adjoint_steps_per_pass": adepth,
                             "bytes_per_gradient": bytes_grad,
                             "bytes_per_node": parts,
                             "timing": "the timed batched gradient (forward graph, residuals, backward graph): "
                                       "the algorithmic bytes of its passes per padded node (bytes_per_node) "
                                       "over its wall time"})
        b.close()
        for mm in ms:
            mm.close()
    if cpu:
        out["cpu_baseline"] = cpu_baseline_grad(case)
    return out


def calibrate_hbm(mib, reps=5):
    """A plain device copy of `mib` MiB, `reps` times: its FETCH_SIZE and
    WRITE_SIZE per dispatch scale the same counters read on the kernels."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_hbm_mib": mib, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nt", type=int, default=2000, help="time steps per shot")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="cart3d")
    ap.add_argument("--also", default="cart2d,spher3d,batch2d,adjoint2d",
                    help="comma-separated extra workloads measured after the headline one "
                         "(one shot each; empty for none)")
    ap.add_argument("--cpu-steps", type=int, default=24, help="time steps of the headline CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-hbm", type=int, metavar="MIB", default=0,
                    help="only run a plain MIB-MiB device copy (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(the time loop runs on a single GPU)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.calibrate_hbm:
        calibrate_hbm(args.calibrate_hbm)
        return

    res, case = measure(args.workload, args.nt, args.steps, args.warmup)
    out = {
        "metric": "Gcell-updates/s (%s forward, 1 GPU)" % args.workload,
        "value": res["value"],
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "ms_per_time_step": res["ms_per_time_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded three-layer rough-base model, Ricker 2 MHz)",
        "config": {"workload": res["workload"], "cells_per_step": res["cells_per_step"],
                   "time_steps_per_shot": res["time_steps_per_shot"],
                   "step": "one shot (all time steps)"},
        "roofline": res["roofline"],
    }
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(case, args.cpu_steps)
        except Exception as e:  # the baseline never hides the GPU line
            out["cpu_baseline"] = {"error": str(e)}
    del case
    also = []
    for wl in [x for x in args.also.split(",") if x and x != args.workload]:
        try:
            if wl == "batch2d":
                also.append(measure_batch(args.nt, cpu_shots=0 if args.no_cpu else 4))
            elif wl == "adjoint2d":
                also.append(measure_grad(cpu=not args.no_cpu))
            elif wl == "cart2d":
                # the 2-D forward at its own length (nt=3000), with the CPU
                # restatement timed over the whole run beside it
                r, c2 = measure(wl, 3000, 3, 2)    # 2nd warm-up run captures the graph
                if not args.no_cpu:
                    r["cpu_baseline"] = cpu_baseline(c2, 3000)
                also.append(r)
            else:
                r, c3 = measure(wl, args.nt, 1, 1)
                if not args.no_cpu:
                    r["cpu_baseline"] = cpu_baseline(c3, 20)
                also.append(r)
        except Exception as e:
            also.append({"workload": wl, "error": str(e)})
        print("bench: %s done" % wl, file=sys.stderr, flush=True)
    if also:
        out["also"] = also
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
