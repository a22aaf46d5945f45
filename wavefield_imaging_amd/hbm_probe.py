"""Plain device-to-device copy of a known size, for calibrating HBM counters.

Run under ``rocprofv3 --pmc FETCH_SIZE`` (then ``WRITE_SIZE``): the copy kernel
moves exactly ``--mib`` MiB in and out per dispatch, so the ratio of the
counter (KiB per dispatch) to that size is the scale factor applied to the
same counter read on the stencil kernel.
"""
import argparse

import torch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    n = args.mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print("copied %d MiB x %d" % (args.mib, args.reps), flush=True)


if __name__ == "__main__":
    main()
