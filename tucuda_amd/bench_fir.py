"""FIRFilter2 / general-taps measurement (one output per launch).

    python -m tucuda_amd.bench_fir

Workloads: the 4096x4096 u8 image filtered by 17 x 9 general taps (N(0,1),
unit L1 norm, seeded) into float32 and into float16, and by FIRFilter2 lobes
of 9 / 5 (17 x 9 symmetric taps); and a 1920x1080 float32 image by 17 / 17
lobes.  Each is timed on both kernel families (matrix cores, the default,
and the direct FP32 vector kernel, TUCU_FIR_PATH=valu), launches replayed
from a captured HIP graph over 4 rotating buffer sets.  Prints one JSON line
with microseconds per launch and the algorithmic bytes rate (input read
once, the valid output region written once).
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np


def main(argv=None):
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

    res = {"workloads": {}}
    for name, img, odt, (kind, h, v), rh, rv in cases:
        outs = [torch.zeros(img.shape, dtype=odt, device=dev) for _ in range(4)]
        ins = [img.clone() for _ in range(4)]
        nbytes = img.numel() * img.element_size() + (img.shape[0] - 2 * rv) * (img.shape[1] - 2 * rh) * outs[0].element_size()
        entry = {}
        for fam in ("mfma", "valu"):
            if fam == "valu":
                os.environ["TUCU_FIR_PATH"] = "valu"
            else:
                os.environ.pop("TUCU_FIR_PATH", None)
            if kind == "taps":
                f = tucuda_amd.SeparableFilter2(h, v)
            else:
                f = tucuda_amd.FIRFilter2().initialize(h, v)

            def run():
                for x, o in zip(ins, outs):
                    f.convolve(x, o)

            run()
            torch.cuda.synchronize()
            graph = torch.cuda.CUDAGraph()
            st = torch.cuda.Stream(dev)
            st.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.stream(st):
                with torch.cuda.graph(graph, stream=st):
                    run()
            torch.cuda.current_stream(dev).wait_stream(st)
            graph.replay()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            rs = torch.cuda.current_stream(dev)
            e0.record(rs)
            for _ in range(20):
                graph.replay()
            e1.record(rs)
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) * 1000 / (20 * len(outs))
            entry[fam] = {"us": round(us, 2), "gb_s": round(nbytes / us / 1e3, 1)}
            del graph
        os.environ.pop("TUCU_FIR_PATH", None)
        res["workloads"][name] = entry
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    sys.exit(main())
