"""Batched-mode measurement: 256 images of uneven sizes in one launch.

    python -m tucuda_amd.bench_batch [--images 256] [--reps 20] [--seed 3]

Workload: widths and heights log-uniform over [240, 2160] with the aspect
ratio clamped to [0.5, 2], per-image sigma uniform in [0.8, 4.0], lobe
radius ceil(4 sigma) (TUCU_RULE_CEIL4SIGMA), all six outputs; float32 and
bfloat16 storage.  Compared forms, all on the same resident buffers:

* ``plain``: one fused six-output launch per image (the single-image API);
* ``per_method``: six FIRGaussianConvolver2 method calls per image, the way
  the reference's API is used (FIRGaussianConvolver.h:178-242); reference-
  rule lobes, since FIRFilter2 only takes the reference's lobe lengths;
  both timed as issued from Python (``*_ms``, launch overhead included) and
  replayed from one captured HIP graph (``*_graph_ms``, no host cost);
* ``batched``: one persistent launch over every image, tile height picked by
  the autotuner, replayed from a captured HIP graph.

Images and outputs have the layout of ``cu::Array2<T>`` (rows padded to 256
bytes, ``tucuda_amd.array2``), the reference's device image; ``--packed``
uses rows packed at their width instead (most widths then start rows off
cache-line boundaries).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np


def make_sizes(n: int, seed: int):
    from . import synth
    u = synth.uniform01(seed, 3 * n).reshape(n, 3)
    lo, hi = math.log(240), math.log(2160)
    out = []
    for a, b, c in u:
        h = int(round(math.exp(lo + (hi - lo) * a)))
        w = int(round(math.exp(lo + (hi - lo) * b)))
        w = int(min(max(w, h / 2), 2 * h))
        sigma = float(np.float32(0.8 + 3.2 * c))
        out.append((h, w, sigma))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--images", type=int, default=256)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--dtypes", default="float32,bfloat16")
    ap.add_argument("--packed", action="store_true", help="rows packed at their width instead of a 256-byte pitch")
    ap.add_argument("--no-per-method", action="store_true", help="skip the method-per-output form")
    args = ap.parse_args(argv)
    print(json.dumps(run(args.images, args.reps, args.seed, args.dtypes, args.packed, not args.no_per_method)),
          flush=True)


def run(images: int = 256, reps: int = 20, seed: int = 3, dtypes: str = "float32,bfloat16",
        packed: bool = False, per_method_form: bool = True) -> dict:
    """The measurement described in the module docstring; returns the dict
    main() prints."""
    import torch
    import tucuda_amd
    from . import synth

    class _Args:
        pass
    args = _Args()
    args.images, args.reps, args.seed, args.dtypes, args.packed = images, reps, seed, dtypes, packed
    dev = torch.device("cuda:0")
    sizes = make_sizes(args.images, args.seed)
    total_px = sum(h * w for h, w, _ in sizes)
    result = {"workload": f"{args.images} images, log-uniform [240,2160], sigma U[0.8,4], radius ceil(4 sigma), 6 outputs",
              "images": args.images, "megapixels": round(total_px / 1e6, 2),
              "layout": "rows packed" if args.packed else "256-byte row pitch (cu::Array2)", "forms": {}}

    def image(h, w, dt):
        if args.packed:
            return torch.empty((h, w), dtype=dt, device=dev)
        return tucuda_amd.array2(h, w, dt, dev)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def captured(fn):
        """fn's launches captured in one HIP graph; returns its replay."""
        fn()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        st = torch.cuda.Stream(dev)
        st.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(st):
            with torch.cuda.graph(g, stream=st):
                fn()
        torch.cuda.current_stream(dev).wait_stream(st)
        keep.append(g)
        return g.replay

    keep = []
    for dname in args.dtypes.split(","):
        dt = getattr(torch, dname)
        ins, outs = [], []
        for i, (h, w, _) in enumerate(sizes):
            x = image(h, w, dt)
            x.copy_(torch.from_numpy(synth.edge_image(1000 + i, h, w)))
            ins.append(x)
            outs.append([image(h, w, dt) for _ in range(6)])
        gs = [tucuda_amd.FIRGaussianConvolver2(s, rule=tucuda_amd.TUCU_RULE_CEIL4SIGMA) for _, _, s in sizes]

        def plain():
            for g, x, o in zip(gs, ins, outs):
                g.convolve6(x, o)

        ms_plain = timed(plain, max(1, args.reps // 4))
        ms_plain_graph = timed(captured(plain), max(1, args.reps // 4))

        t0 = time.perf_counter()
        batch = tucuda_amd.GaussBatch([(x, o, g._lobes()) for g, x, o in zip(gs, ins, outs)])
        t_build = time.perf_counter() - t0
        rows, tuned_us = batch.autotune(reps=5)
        graph = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(dev)
        s.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(s):
            with torch.cuda.graph(graph, stream=s):
                batch.launch()
        torch.cuda.current_stream(dev).wait_stream(s)
        ms_batch = timed(graph.replay, args.reps)
        ms_batch_eager = timed(batch.launch, args.reps)
        form = {
            "plain_ms": round(ms_plain, 3),
            "batched_ms": round(ms_batch, 3),
            "batched_eager_ms": round(ms_batch_eager, 3),
            "plain_graph_ms": round(ms_plain_graph, 3),
            "plain_gpix_s": round(total_px / ms_plain / 1e6, 2),
            "plain_graph_gpix_s": round(total_px / ms_plain_graph / 1e6, 2),
            "batched_gpix_s": round(total_px / ms_batch / 1e6, 2),
            "speedup_vs_plain": round(ms_plain / ms_batch, 2),
            "tile_rows": rows, "tiles": batch.info()["tiles"], "grid": batch.info()["grid"],
            "table_build_s": round(t_build, 3),
        }
        if dname == "float32" and per_method_form:
            ref_gs = [tucuda_amd.FIRGaussianConvolver2(sg) for _, _, sg in sizes]

            def per_method():
                for g, x, o in zip(ref_gs, ins, outs):
                    g.smooth(x, o[0]); g.diffH(x, o[1]); g.diffV(x, o[2])
                    g.diffHH(x, o[3]); g.diffHV(x, o[4]); g.diffVV(x, o[5])

            ms_pm = timed(per_method, max(1, args.reps // 4))
            ms_pm_graph = timed(captured(per_method), max(1, args.reps // 4))
            form["per_method_ms"] = round(ms_pm, 3)
            form["per_method_graph_ms"] = round(ms_pm_graph, 3)
            form["per_method_gpix_s"] = round(total_px / ms_pm / 1e6, 2)
            form["per_method_graph_gpix_s"] = round(total_px / ms_pm_graph / 1e6, 2)
        result["forms"][dname] = form
        del batch, graph, ins, outs
        keep.clear()
        torch.cuda.empty_cache()
    f32 = result["forms"].get("float32", {})
    bf = result["forms"].get("bfloat16", {})
    if f32 and bf:
        result["bf16_batched_vs_f32_plain"] = round(f32["plain_ms"] / bf["batched_ms"], 2)
        result["bf16_batched_vs_f32_plain_graph"] = round(f32["plain_graph_ms"] / bf["batched_ms"], 2)
        if "per_method_ms" in f32:
            result["bf16_batched_vs_f32_per_method"] = round(f32["per_method_ms"] / bf["batched_ms"], 2)
            result["bf16_batched_vs_f32_per_method_graph"] = round(f32["per_method_graph_ms"] / bf["batched_ms"], 2)
    return result


if __name__ == "__main__":
    sys.exit(main())
