#!/usr/bin/env python3
"""Benchmark: the acoustic FD time loop on one MI355X.

Default workload (the headline line): plain 3-D cartesian forward modelling,
256 x 512 x 512 (z, x, y) model, h = 5 m, 40-node sponge on every face (padded
grid 336 x 592 x 592 = 117,755,904 nodes), nt = 2000 steps of dt = 6.67e-9 s,
three-layer rough-base velocity, one 2 MHz Ricker source, 128 x 128 receivers,
sponge + one-way ABC (dabc=y), FP32.

The default run also measures, under "also":
  cart2d    400 x 800 (z, x), nb = 50 (padded 500 x 900), nt = 3000, one
            source, 800 receivers; with the CPU restatement over the whole run
  spher3d   the spherical shell, 400 x 512 x 512 (r, theta, phi), r from
            3387.5 km, dr = 5 m, r*dtheta = 5 m at mid radius, 30-node sponge
            (padded 460 x 572 x 572 = 150,504,640 nodes), nt = 2000, one source
            10 nodes below the top, 256 x 256 receivers (also selectable as the
            headline with --workload spher3d)
  batch2d   256 independent 2-D shots of uneven size, batched graph vs the
            per-shot loop, in shots/s
  adjoint2d 64 shots on a 1000 x 2000 grid, nt = 4000, checkpoints every 64
            steps: misfit gradients per second, batched graphs vs the per-shot
            loop

One benchmark step = one complete shot (all nt time steps) with every input
already resident in HBM.  value = node-updates per second over the timed
steps (Gcell-updates/s).  The roofline figure is for the dominant kernel
(the fused step kernel): 16 algorithmic bytes per node-update (read the
previous field, the current field and the velocity term, write the new
field), times the padded nodes per launch, over the kernel's mean duration
measured with HIP events on the stream it runs on.

cpu_baseline: the C oracle (a restatement of the reference CPU program,
OpenMP over the host cores) on the same padded grid for a bounded number of
time steps.
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
    "cart2d": dict(kernel="k_cart_step<2,2,true>",
                   traffic=None,
                   desc="cart2d 400x800 (z,x) h=5m nb=50 nt=%d dabc=y fdorder=4, padded 500x900, "
                        "1 source, 800 receivers"),
    "cart3d": dict(kernel="k_cart_step<2,3,true>",
                   traffic="cart3d_step_traffic.json",
                   desc="cart3d 256x512x512 (z,x,y) h=5m nb=40 nt=%d dabc=y fdorder=4, "
                        "padded 336x592x592, 1 source, 128x128 receivers"),
    "spher3d": dict(kernel="k_spher_step3",
                    traffic="spher3d_step_traffic.json",
                    desc="spher3d 400x512x512 (r,theta,phi) r0=3387.5km dr=5m nb=30 nt=%d "
                         "expl=n, padded 460x572x572, 1 source, 256x256 receivers"),
}


def _traffic(name):
    """HBM bytes per launch of the step kernel from the committed PMC summary."""
    if not name:
        return None
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(case, steps):
    """Time the C oracle (reference CPU path restated) on `steps` time steps."""
    import ctypes as C

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    par = dict(case["par"], nt=steps)
    O.run(case["kind"], par, case["vel"], case["sou"], case["rec"], case["wav"][:steps], final=False)
    f = O.lib().orc_last_loop_seconds
    f.restype = C.c_double
    secs = f()
    cells = int(np.prod([n + 2 * par["nb"] for n in par["n"]]))
    return dict(value=round(cells * steps / secs / 1e9, 4), unit="Gcell-updates/s",
                cores=int(O.lib().orc_threads()), kind="port",
                sample="%s padded %s, %d of %d time steps, %.1f s loop wall time"
                       % (case["kind"], "x".join(str(n + 2 * par["nb"]) for n in par["n"]),
                          steps, case["par"]["nt"], secs))


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
    # microseconds, launch overhead would dominate), so the timed region runs
    # the graph and one more profiled (per-launch) run gives the kernel time.
    eager_timed = m.ndim == 3
    m.profile(eager_timed)
    t0 = time.perf_counter()
    for _ in range(steps):
        m.run_dev()
    t1 = time.perf_counter()
    if not eager_timed:
        m.profile(True)
        m.run_dev()
    launches, kern_ms = m.kernel_stats()
    m.profile(False)
    m.close()
    secs = t1 - t0
    avg_ms = kern_ms / max(launches, 1)
    achieved = BYTES_PER_UPDATE * cells / (avg_ms * 1e-3) / 1e9
    out = dict(workload=w["desc"] % nt, cells_per_step=cells, time_steps_per_shot=nt,
               value=round(cells * nt * steps / secs / 1e9, 3),
               ms_per_step=round(secs / steps * 1e3, 3),
               roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": _traffic(w["traffic"]), "kernel": w["kernel"],
                         "kernel_mean_ms": round(avg_ms, 4), "launches": launches})
    return out, case


def measure_batch(nt, nshot=256, steps=1, tune_steps=20, per_shot=True):
    """256 independent 2-D shots of uneven size (nz ~ U[200,600], nx ~
    U[400,1200], nb=50): the batched graph-captured run (autotuned tile,
    shot grouping and steps per launch) and the plain per-shot loop of the
    single-grid kernel, in shots/s."""
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
    rx, grp, depth, tune_ms = b.tune_depth(steps=tune_steps)
    b.run()                                   # warm-up (graph capture + launch)
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run()
    t1 = time.perf_counter()
    batched = nshot * steps / (t1 - t0)
    out = dict(workload="batch2d %d shots nz~U[200,600] nx~U[400,1200] nb=50 nt=%d, %d padded cells"
                        % (nshot, nt, cells),
               shots_per_s=round(batched, 3), gcells_per_s=round(cells * nt * steps / (t1 - t0) / 1e9, 3),
               s_per_batch=round((t1 - t0) / steps, 4), tile_rx=rx, group_mb=grp, steps_per_launch=depth,
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
    return out


def measure_grad(nt=4000, nshot=64, sample=8, batched=True, rx=8, group_mb=0):
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
        out.update(batched_gradients_per_s=round(bat, 3), batched_s_per_gradient_sum=round(t1 - t0, 3),
                   batched_tile_rx=rx, batched_group_mb=group_mb,
                   batched_gcells_per_s_per_propagation=round(3 * cells * nt * nshot / (t1 - t0) / 1e9, 3),
                   speedup_vs_per_shot_loop=round(bat / loop, 2))
        b.close()
        for mm in ms:
            mm.close()
    return out


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
    ap.add_argument("--cpu-steps", type=int, default=24, help="time steps of the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(the time loop runs on a single GPU)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    res, case = measure(args.workload, args.nt, args.steps, args.warmup)
    out = {
        "metric": "Gcell-updates/s (%s forward, 1 GPU)" % args.workload,
        "value": res["value"],
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded three-layer rough-base model, Ricker 2 MHz)",
        "config": {"workload": res["workload"], "cells_per_step": res["cells_per_step"],
                   "time_steps_per_shot": res["time_steps_per_shot"]},
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
                also.append(measure_batch(args.nt))
            elif wl == "adjoint2d":
                also.append(measure_grad())
            elif wl == "cart2d":
                # the 2-D forward at its own length (nt=3000), with the CPU
                # restatement timed over the whole run beside it
                r, c2 = measure(wl, 3000, 3, 2)    # 2nd warm-up run captures the graph
                if not args.no_cpu:
                    r["cpu_baseline"] = cpu_baseline(c2, 3000)
                also.append(r)
            else:
                r, _ = measure(wl, args.nt, 1, 1)
                also.append(r)
        except Exception as e:
            also.append({"workload": wl, "error": str(e)})
    if also:
        out["also"] = also
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
