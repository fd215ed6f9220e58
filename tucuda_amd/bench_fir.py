"""FIRFilter2 / general-taps and 16-bit storage measurement.

    python -m tucuda_amd.bench_fir [--cases a,b] [--families mfma,valu] [--reps 20] [--warmup 40]

Workloads: the 4096x4096 u8 image filtered by 17 x 9 general taps (N(0,1),
unit L1 norm, seeded) into float32 and into float16, and by FIRFilter2 lobes
of 9 / 5 (17 x 9 symmetric taps); a 1920x1080 float32 image by 17 / 17
lobes; and the six-output Gaussian (sigma 3) of a 1920x1080 image stored in
float32, bfloat16 and float16, into outputs of the same type.  Each is timed on
both kernel families (matrix cores, the default, and the direct FP32 vector
kernel, TUCU_FIR_PATH / TUCU_GAUSS_PATH=valu), launches replayed from a
captured HIP graph of 48 launches over 4 rotating buffer sets, after
`--warmup` untimed replays.  Prints one JSON line with
microseconds per launch and the algorithmic bytes rate (input read once,
the valid output region of each output written once).
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--cases", default="", help="comma-separated workload names (default: all)")
    ap.add_argument("--families", default="mfma,valu")
    ap.add_argument("--reps", type=int, default=20, help="graph replays timed per workload")
    ap.add_argument("--warmup", type=int, default=40,
                    help="untimed graph replays before the timed ones (the GPU's clocks settle)")
    args = ap.parse_args(argv)
    import torch
    import tucuda_amd
    from . import synth

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(7)

    def unit(n):
        t = rng.standard_normal(n).astype(np.float32)
        return t / np.abs(t).sum()

    tH, tV = unit(17), unit(9)
    lobeH = np.abs(unit(9)).astype(np.float32)
    lobeV = np.abs(unit(5)).astype(np.float32)
    cases = []
    img4k = torch.from_numpy(synth.u8_image(2, 4096, 4096)).to(dev)
    for odt in (torch.float32, torch.float16):
        cases.append(("taps17x9_4096_u8_to_" + str(odt).split(".")[-1], img4k, odt, ("taps", tH, tV), 8, 4))
    cases.append(("lobes9x5_4096_u8_to_float32", img4k, torch.float32, ("lobes", lobeH, lobeV), 8, 4))
    img1080 = torch.from_numpy(synth.edge_image(3, 1080, 1920)).to(dev)
    l17 = np.abs(unit(17)).astype(np.float32)
    cases.append(("lobes17x17_1080p_f32", img1080, torch.float32, ("lobes", l17, l17), 16, 16))
    for dt in (torch.float32, torch.bfloat16, torch.float16):
        cases.append(("gauss6_1080p_" + str(dt).split(".")[-1], img1080.to(dt), dt, ("gauss6", 3.0, None), 16, 16))
    if args.cases:
        want = set(args.cases.split(","))
        unknown = want - {c[0] for c in cases}
        if unknown:
            raise SystemExit(f"unknown cases {sorted(unknown)}")
        cases = [c for c in cases if c[0] in want]

    res = {"workloads": {}}
    for name, img, odt, (kind, h, v), rh, rv in cases:
        nout = 6 if kind == "gauss6" else 1
        outs = [[torch.zeros(img.shape, dtype=odt, device=dev) for _ in range(nout)] for _ in range(4)]
        ins = [img.clone() for _ in range(4)]
        nbytes = img.numel() * img.element_size() + \
            nout * (img.shape[0] - 2 * rv) * (img.shape[1] - 2 * rh) * outs[0][0].element_size()
        entry = {}
        for fam in args.families.split(","):
            for var in ("TUCU_FIR_PATH", "TUCU_GAUSS_PATH"):
                if fam == "valu":
                    os.environ[var] = "valu"
                else:
                    os.environ.pop(var, None)
            if kind == "taps":
                f = tucuda_amd.SeparableFilter2(h, v)
            elif kind == "gauss6":
                f = tucuda_amd.FIRGaussianConvolver2(h)
            else:
                f = tucuda_amd.FIRFilter2().initialize(h, v)

            def run(rounds):
                for _ in range(rounds):
                    for x, o in zip(ins, outs):
                        if kind == "gauss6":
                            f.convolve6(x, o)
                        else:
                            f.convolve(x, o[0])

            run(1)
            torch.cuda.synchronize()
            # 48 launches per graph (the 4 buffer sets 12 times): the cost of
            # starting a replay then adds little per launch
            rounds = 12
            graph = torch.cuda.CUDAGraph()
            st = torch.cuda.Stream(dev)
            st.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.stream(st):
                with torch.cuda.graph(graph, stream=st):
                    run(rounds)
            torch.cuda.current_stream(dev).wait_stream(st)
            for _ in range(max(1, args.warmup)):
                graph.replay()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            rs = torch.cuda.current_stream(dev)
            e0.record(rs)
            for _ in range(args.reps):
                graph.replay()
            e1.record(rs)
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) * 1000 / (args.reps * rounds * len(outs))
            entry[fam] = {"us": round(us, 2), "gb_s": round(nbytes / us / 1e3, 1)}
            del graph
        os.environ.pop("TUCU_FIR_PATH", None)
        os.environ.pop("TUCU_GAUSS_PATH", None)
        print(f"{name}: {entry}", file=sys.stderr, flush=True)
        res["workloads"][name] = entry
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    sys.exit(main())
