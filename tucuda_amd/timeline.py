"""Phase timeline of the one-round matrix-core launch: 1920x1080 float32,
sigma 3, all six outputs (the workload of bench.py).

Needs the timeline build of the library (``make -C tucuda_amd/csrc
timeline``: ``libtucuda_amd_timeline.so``, the same kernels with lane 0 of
every wave writing the 100 MHz real-time counter at each phase).  Each of
``--reps`` launches follows another launch on a different buffer set, as in
a stream of launches (``--idle``: after a synchronisation and 5 ms of idle
instead); per launch the times are taken from the first wave's start, and
the table gives, per phase, the median over the launches of the 10th, 50th
and 90th percentile over the waves, in microseconds.

    python -m tucuda_amd.timeline [--reps 30] [--idle] [--json FILE]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

PHASES = [(0, "wave started"), (10, "strip loads issued"), (11, "operand tables copied"),
          (1, "barrier passed (input arrived)"), (2, "input scaled and split"),
          (3, "smooth stored"), (4, "diffH stored"), (6, "diffHH stored"), (5, "diffV stored"),
          (7, "diffHV stored"), (8, "diffVV stored"), (9, "wave ended")]
SLOTS = 12
WAVES = 2048


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--reps", type=int, default=30)
    ap.add_argument("--idle", action="store_true", help="5 ms of idle before each timed launch")
    ap.add_argument("--json", default=None, help="write every launch's percentiles here")
    args = ap.parse_args(argv)

    import torch                                # first: the library then binds to torch's HIP runtime
    from . import _capi
    _capi.LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtucuda_amd_timeline.so")
    lib = _capi.load()
    if not hasattr(lib, "tucu_timeline_read"):
        raise SystemExit(f"{_capi.LIB_PATH} has no timeline: build it with make -C tucuda_amd/csrc timeline")
    lib.tucu_timeline_read.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    lib.tucu_timeline_clear.argtypes = []

    import tucuda_amd
    from . import synth
    dev = torch.device("cuda:0")
    img = synth.edge_image(1, 1080, 1920)
    nset = 10
    ins = [torch.from_numpy(img).to(dev) for _ in range(nset)]
    outs = [[torch.zeros(1080, 1920, dtype=torch.float32, device=dev) for _ in range(6)] for _ in range(nset)]
    g = tucuda_amd.FIRGaussianConvolver2(3.0)
    for i in range(nset):
        g.convolve6(ins[i], outs[i])
    torch.cuda.synchronize()

    buf = np.zeros(WAVES * SLOTS, np.uint64)
    per_launch = []
    for rep in range(args.reps):
        i = rep % nset
        if lib.tucu_timeline_clear() != 0:
            raise SystemExit("tucu_timeline_clear failed")
        if args.idle:
            torch.cuda.synchronize()
            time.sleep(0.005)
        else:
            g.convolve6(ins[(i + 5) % nset], outs[(i + 5) % nset])
        g.convolve6(ins[i], outs[i])
        torch.cuda.synchronize()
        if lib.tucu_timeline_read(buf.ctypes.data, buf.nbytes) != 0:
            raise SystemExit("tucu_timeline_read failed")
        t = buf.reshape(WAVES, SLOTS).astype(np.int64)
        t = t[t[:, 9] > 0]                      # the waves of live strips
        rel = (t - t[:, 0].min()) * 10          # ns
        per_launch.append({s: np.percentile(rel[:, s], (10, 50, 90)).tolist() for s, _ in PHASES}
                          | {"waves": int(len(t))})
    skip = min(5, args.reps // 2)               # the first launches warm up
    print(f"{'phase':34s} {'10 %':>8s} {'50 %':>8s} {'90 %':>8s}   (us from the first wave's start;"
          f" median of {args.reps - skip} launches, {per_launch[-1]['waves']} waves)")
    table = {}
    for s, name in PHASES:
        v = [float(np.median([r[s][j] for r in per_launch[skip:]])) / 1000 for j in range(3)]
        table[name] = v
        print(f"{name:34s} {v[0]:8.2f} {v[1]:8.2f} {v[2]:8.2f}")
    if args.json:
        with open(args.json, "w") as f:
            json.dump({"idle": args.idle, "table_us": table, "launches": per_launch}, f)
    return 0


if __name__ == "__main__":
    sys.exit(main())
