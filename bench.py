#!/usr/bin/env python3
"""Benchmark: plain 3-D cartesian acoustic forward modelling on one MI355X.

Workload: 256 x 512 x 512 (z, x, y) model, h = 5 m, 40-node sponge on every
face (padded grid 336 x 592 x 592 = 117,755,904 nodes), nt = 2000 steps of
dt = 6.67e-9 s, three-layer rough-base velocity, one 2 MHz Ricker source,
128 x 128 receivers, sponge + one-way ABC (dabc=y), FP32.

One benchmark step = one complete shot (all nt time steps) with every input
already resident in HBM.  value = node-updates per second over the timed
steps (Gcell-updates/s).  The roofline figure is for the dominant kernel
(the fused step kernel): 16 algorithmic bytes per node-update (read the
previous field, the current field and (v dt)^2, write the new field), times
the padded nodes per launch, over the kernel's mean duration measured with
HIP events on the stream it runs on.

cpu_baseline: the C oracle (a restatement of the reference CPU program
sfAWEFDcart3Dcpu, OpenMP over the host cores) on the same padded grid for a
bounded number of time steps.
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


def _traffic():
    """HBM bytes per launch of the step kernel from the committed PMC summary."""
    path = os.path.join(ROOT, "profiles", "cart3d_step_traffic.json")
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nt", type=int, default=2000, help="time steps per shot")
    ap.add_argument("--cpu-steps", type=int, default=24, help="time steps of the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(the time loop is single-GPU; see DESIGN.md for how it would shard)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from wavefield_imaging_amd import awefd, synth

    case = synth.cart3d(nt=args.nt)
    par = case["par"]
    m = awefd.Model("cart3d", par)
    m.set_model(case["vel"])
    m.set_sources(case["sou"])
    m.set_receivers(case["rec"])
    m.set_wavelet(case["wav"])
    cells = int(np.prod(m.padded))
    print("bench: cart3d padded %s, nt=%d, %d receivers" % (m.padded, args.nt, m.nr),
          file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        m.run_dev()
    m.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m.run_dev()
    t1 = time.perf_counter()
    launches, kern_ms = m.kernel_stats()
    m.profile(False)
    secs = t1 - t0
    updates = cells * args.nt * args.steps
    value = updates / secs / 1e9
    avg_ms = kern_ms / max(launches, 1)
    achieved = BYTES_PER_UPDATE * cells / (avg_ms * 1e-3) / 1e9
    traffic = _traffic()
    out = {
        "metric": "Gcell-updates/s (cart3d forward, 1 GPU)",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded three-layer rough-base model, Ricker 2 MHz)",
        "config": {
            "workload": "cart3d 256x512x512 (z,x,y) h=5m nb=40 nt=%d dabc=y fdorder=4, "
                        "padded 336x592x592, 1 source, 128x128 receivers" % args.nt,
            "cells_per_step": cells,
            "time_steps_per_shot": args.nt,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "k_cart_step<2,3,true>",
            "kernel_mean_ms": round(avg_ms, 4),
            "launches": launches,
        },
    }
    m.close()
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(case, args.cpu_steps)
        except Exception as e:  # the baseline never hides the GPU line
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
