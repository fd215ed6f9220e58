"""ICIA error-moment measurement.

    python -m tucuda_amd.bench_icia [--reps 40] [--warmup 20]

Workloads: a 1920x1080 float32 image and a 3840x2160 u8 image (seeded
synthetic edges), sigma 2 and 3.  Times, from a captured HIP graph of 48
calls over 4 rotating input buffers, after `--warmup` untimed replays, (a) tucu_icia_moment -- diffH and diffV in one pass
that never writes them, reduced to the 21 (affine) or 36 (projective)
moments -- and (b) the fused diffH + diffV launch that writes both gradient
images, the first half of the reference's ICIA step (ICIA.h:344-350).
Prints one JSON line with microseconds per call and the algorithmic rate of
(a) (the input read once).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import sys

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=40, help="graph replays timed per workload")
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed graph replays before the timed ones (the GPU's clocks settle)")
    args = ap.parse_args(argv)
    import torch
    import tucuda_amd
    from . import _capi, synth
    from .filters import image_desc

    dev = torch.device("cuda:0")
    lib = _capi.load()
    fp = ctypes.POINTER(ctypes.c_float)
    res = {"workloads": {}}
    imgs = {"1080p_f32": torch.from_numpy(synth.edge_image(3, 1080, 1920)).to(dev),
            "2160p_u8": torch.from_numpy(np.clip(synth.edge_image(4, 2160, 3840), 0, 255).astype(np.uint8)).to(dev)}
    st = torch.cuda.Stream(dev)
    for name, img in imgs.items():
        ins = [img.clone() for _ in range(4)]
        for sigma in (2.0, 3.0):
            l0, l1, _ = tucuda_amd.gauss_lobes(sigma)
            entry = {}
            for model, m in (("affine", 6), ("projective", 8)):
                srcs = [image_desc(x) for x in ins]
                need = ctypes.c_size_t(0)
                assert lib.tucu_icia_moment(l0.ctypes.data_as(fp), l0.size, l1.ctypes.data_as(fp), l1.size,
                                            ctypes.byref(srcs[0]), m, None, None, ctypes.byref(need), None) == 0
                wss = [torch.empty(need.value, dtype=torch.uint8, device=dev) for _ in ins]
                outs = [torch.empty(m * (m + 1) // 2, device=dev) for _ in ins]

                def run(sh):
                    for src, ws, out in zip(srcs, wss, outs):
                        n = ctypes.c_size_t(need.value)
                        assert lib.tucu_icia_moment(l0.ctypes.data_as(fp), l0.size, l1.ctypes.data_as(fp), l1.size,
                                                    ctypes.byref(src), m, ctypes.c_void_p(out.data_ptr()),
                                                    ctypes.c_void_p(ws.data_ptr()), ctypes.byref(n), sh) == 0
                entry[model] = round(_graph_us(torch, dev, st, run, args.reps, args.warmup) / len(ins), 2)
            g = tucuda_amd.FIRGaussianConvolver2(sigma)
            grads = [(torch.empty_like(x, dtype=torch.float32), torch.empty_like(x, dtype=torch.float32)) for x in ins]

            def run_grad(sh):
                for x, (eh, ev) in zip(ins, grads):
                    g.convolve6(x, {"diffH": eh, "diffV": ev})
            entry["diffH_diffV_written"] = round(_graph_us(torch, dev, st, run_grad, args.reps, args.warmup)
                                                 / len(ins), 2)
            nbytes = img.numel() * img.element_size()
            entry["moment_gb_s"] = round(nbytes / entry["affine"] / 1e3, 1)
            res["workloads"][f"{name}_sigma{sigma:g}"] = entry
            print(f"{name} sigma {sigma}: {entry}", file=sys.stderr, flush=True)
    print(json.dumps(res), flush=True)


def _graph_us(torch, dev, st, run, reps, warmup, rounds=12):
    """Microseconds per run() from a graph holding `rounds` run()'s (so that
    the cost of starting a replay adds little per call), after `warmup`
    untimed replays."""
    sh = ctypes.c_void_p(st.cuda_stream)
    st.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(st):
        run(sh)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(st):
        with torch.cuda.graph(graph, stream=st):
            for _ in range(rounds):
                run(sh)
    torch.cuda.synchronize()
    for _ in range(max(1, warmup)):
        graph.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    rs = torch.cuda.current_stream(dev)
    e0.record(rs)
    for _ in range(reps):
        graph.replay()
    e1.record(rs)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1000 / (reps * rounds)


if __name__ == "__main__":
    sys.exit(main())
