#!/usr/bin/env python3
"""Benchmark: 1920x1080 float32 Gaussian smoothing + five derivatives on one
MI355X, fused into one launch per image.

A step is one call of the fused kernel over one 1920x1080 image (sigma 3,
lobe radius 16 for all three lobes, all six outputs).  Inputs and outputs
live in HBM before the timed region starts; the step rotates through
``--sets`` buffer sets (about 58 MB each) so that the 256 MB Infinity Cache
cannot hold the working set.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gpix/s & % HBM roofline (1080p Gauss+5 derivs); p50/p99 ms served stream"
HBM_PEAK_GBS = 8000.0
NROW, NCOL, SIGMA = 1080, 1920, 3.0
LAUNCHES_PER_GRAPH = 50                 # at least this many launches per graph replay


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(nrow: int, ncol: int, radii, in_size: int = 4, out_size: int = 4) -> int:
    """Input read once plus each output's valid region written once."""
    total = nrow * ncol * in_size
    for rv, rh in radii:
        total += (nrow - 2 * rv) * (ncol - 2 * rh) * out_size
    return total


def cpu_baseline(img, sigma: float, budget_s: float = 12.0, budget_1core_s: float = 4.0):
    """The oracle's FP32 restatement (H pass then V pass per output, FP32
    intermediate, no FMA contraction) on the host cores, over whole images,
    as many images as fit in about ``budget_s`` seconds (at least one); then
    on one core for about ``budget_1core_s`` seconds."""
    import oracle
    import numpy as np
    try:
        cores = len(os.sched_getaffinity(0))          # every core this process may run on
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, cores)
    lobes = oracle.gauss_lobes(sigma)
    outs = {m: np.zeros(img.shape, np.float32) for m in oracle.METHOD_ORDER}

    def rate(nthreads, budget):
        def one():
            for m in oracle.METHOD_ORDER:
                oracle.gauss(m, sigma, img, outs[m], lobes=lobes, nthreads=nthreads)

        one()                                         # warm-up
        n, t0 = 0, time.perf_counter()
        while True:
            one()
            n += 1
            el = time.perf_counter() - t0
            if el >= budget or (n >= 3 and el * (n + 1) / n > 2 * budget):
                return n, el

    n, el = rate(cores, budget_s)
    n1, el1 = rate(1, budget_1core_s)
    return {"value": round(img.size * n / el / 1e9, 5), "unit": "Gpix/s", "cores": cores, "kind": "port",
            "cpu_count": os.cpu_count(),
            "sample": f"{n} whole 1920x1080 images, 6 outputs each ({el:.1f} s on {cores} threads, "
                      f"one per core of the process's affinity set)",
            "value_1core": round(img.size * n1 / el1 / 1e9, 5),
            "sample_1core": f"{n1} whole images on 1 thread ({el1:.1f} s)"}


def hot_kernel_us(g, ins, outs, launches: int = LAUNCHES_PER_GRAPH, reps: int = 8) -> float:
    """Microseconds per launch with one buffer set used by every launch (the
    input and outputs stay in the 256 MB Infinity Cache between launches):
    the cache-hot figure beside the rotating one."""
    import torch
    dev = ins.device
    stream = torch.cuda.current_stream(dev)
    graph = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream(dev)
    cap.wait_stream(stream)
    with torch.cuda.stream(cap):
        with torch.cuda.graph(graph, stream=cap):
            for _ in range(launches):
                g.convolve6(ins, outs)
    stream.wait_stream(cap)
    for _ in range(2):
        graph.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        graph.replay()
    e1.record(stream)
    torch.cuda.synchronize()
    del graph
    return e0.elapsed_time(e1) * 1000 / (reps * launches)


# Sources that define the headline kernel (sepfir_mfma_kernel<16, 2, float,
# float, MODE_GAUSS>) and its launch: the recorded HBM traffic is reported
# only while they hash as they did when the counters were taken.
KERNEL_SOURCES = ("tucuda_amd/csrc/mfma.h", "tucuda_amd/csrc/sepfir.h", "tucuda_amd/csrc/batch.h",
                  "tucuda_amd/csrc/mfma_f32.hip", "tucuda_amd/csrc/inst_mfma.inc", "tucuda_amd/csrc/capi.hip",
                  "tucuda_amd/csrc/Makefile", "include/tucu.h")


def kernel_source_hash() -> str:
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(rel.encode() + b"\0" + f.read() + b"\0")
    return h.hexdigest()


def kernel_label() -> str:
    """The single-image kernel the library picks for this run: the
    matrix-core strip kernel (tile rows from TUCU_MFMA_ROWS, default 64),
    one workgroup per tile while the tiles fit the resident workgroups
    (1080p: 510 tiles), else the resident workgroups looping over the
    tiles; or the direct FP32 vector kernel under TUCU_GAUSS_PATH=valu."""
    if os.environ.get("TUCU_GAUSS_PATH") == "valu":
        return "direct FP32 vector kernel"
    rows = os.environ.get("TUCU_MFMA_ROWS", "64")
    rows = rows if rows in ("16", "32", "64") else "64"
    cols = {"16": 256, "32": 128, "64": 64}[rows]
    return (f"matrix-core strips, {rows} x {cols} tiles: one workgroup per tile while the tiles fit "
            "the resident workgroups, else the resident workgroups looping over the tiles")


def recorded_traffic():
    """(bytes per launch, stale) from the FETCH_SIZE / WRITE_SIZE record."""
    tfile = os.path.join(ROOT, "profiles", "gauss6_1080p_f32_traffic.json")
    if not os.path.exists(tfile):
        return None, True
    with open(tfile) as f:
        rec = json.load(f)
    if rec.get("source_sha256") != kernel_source_hash():
        return None, True
    return rec.get("bytes_per_launch"), False


def _membench():
    import ctypes
    path = os.path.join(ROOT, "tucuda_amd", "libtucuda_membench.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (make -C tucuda_amd/csrc)")
    L = ctypes.CDLL(path)
    L.tucu_membench_fill.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_void_p]
    L.tucu_membench_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    return L


def write_ceiling(dev, out_bytes: int) -> dict:
    """Reference rates of the memory system with the filters' store form
    (16-byte non-temporal stores), from a hand-written fill and copy kernel
    (tucuda_amd/csrc/membench.hip): the fill of the six outputs' bytes,
    replayed from a HIP graph over 10 rotating buffers (a launch of this
    kernel's size, ramp and tail included), and a sustained fill and copy of
    1 GiB.  GB/s."""
    import ctypes
    import torch
    L = _membench()
    stream = torch.cuda.current_stream(dev)
    sh = ctypes.c_void_p(stream.cuda_stream)

    def events(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) * 1e-3 / reps

    nb = (out_bytes + 15) // 16 * 16
    bufs = [torch.empty(nb // 4, dtype=torch.float32, device=dev) for _ in range(10)]
    graph = torch.cuda.CUDAGraph()
    st = torch.cuda.Stream(dev)
    st.wait_stream(stream)
    with torch.cuda.stream(st):
        with torch.cuda.graph(graph, stream=st):
            for b in bufs:
                assert L.tucu_membench_fill(ctypes.c_void_p(b.data_ptr()), nb, 1.0, 1, 0,
                                            ctypes.c_void_p(st.cuda_stream)) == 0
    stream.wait_stream(st)
    t_small = events(graph.replay, 20) / len(bufs)
    del graph, bufs
    big = 1 << 30
    a = torch.empty(big // 4, dtype=torch.float32, device=dev)
    b = torch.empty(big // 4, dtype=torch.float32, device=dev)
    t_fill = events(lambda: L.tucu_membench_fill(ctypes.c_void_p(a.data_ptr()), big, 2.0, 1, 0, sh), 10)
    t_copy = events(lambda: L.tucu_membench_copy(ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(a.data_ptr()),
                                                 big, sh), 10)
    del a, b
    torch.cuda.empty_cache()
    return {"fill_outputs_size_gbs": round(nb / t_small / 1e9, 1), "fill_outputs_size_us": round(t_small * 1e6, 2),
            "fill_1gib_gbs": round(big / t_fill / 1e9, 1),
            "copy_1gib_gbs": round(2 * big / t_copy / 1e9, 1),
            "kernel": "tucuda_amd/csrc/membench.hip: 16-byte non-temporal stores (and loads, copy)"}


def batched_leg(reps: int) -> dict:
    """The batched workload (256 images, log-uniform [240, 2160], sigma
    U[0.8, 4], radius ceil(4 sigma), six outputs) in one graph-replayed
    launch, float32 and bfloat16, against the plain per-image fused launches
    replayed from one captured graph (tucuda_amd.bench_batch)."""
    from tucuda_amd import bench_batch
    r = bench_batch.run(images=256, reps=reps, seed=3, per_method_form=False)
    f32, bf = r["forms"]["float32"], r["forms"]["bfloat16"]
    return {"workload": r["workload"], "megapixels": r["megapixels"],
            "f32_batched_ms": f32["batched_ms"], "bf16_batched_ms": bf["batched_ms"],
            "f32_plain_graph_ms": f32["plain_graph_ms"], "bf16_plain_graph_ms": bf["plain_graph_ms"],
            "f32_batched_gpix_s": f32["batched_gpix_s"], "bf16_batched_gpix_s": bf["batched_gpix_s"],
            "bf16_batched_vs_f32_plain_graph": r["bf16_batched_vs_f32_plain_graph"],
            "tile_rows": {"float32": f32["tile_rows"], "bfloat16": bf["tile_rows"]}}


def sigma_sweep(dev, nrow: int, ncol: int, sigmas=(0.8, 1.2, 1.6, 2.0, 2.4, 2.8, 3.2, 4.0)) -> dict:
    """Microseconds per fused six-output launch (radius ceil(4 sigma)) for
    each sigma, float32 and bfloat16 images, replayed from a HIP graph over
    6 rotating buffer sets."""
    import torch
    import tucuda_amd
    res = {"workload": f"{ncol}x{nrow}, six outputs, radius ceil(4 sigma), graph replay", "us": {}}
    for dt in (torch.float32, torch.bfloat16):
        xs = [torch.rand(nrow, ncol, device=dev).mul(255).to(dt) for _ in range(6)]
        outs = [[torch.empty(nrow, ncol, device=dev, dtype=dt) for _ in range(6)] for _ in range(6)]
        for sigma in sigmas:
            g = tucuda_amd.FIRGaussianConvolver2(sigma, rule=tucuda_amd.TUCU_RULE_CEIL4SIGMA)

            def run():
                for x, o in zip(xs, outs):
                    g.convolve6(x, o)

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
            for _ in range(30):
                graph.replay()
            e1.record(rs)
            torch.cuda.synchronize()
            res["us"][f"{str(dt).split('.')[-1]} sigma={sigma}"] = round(e0.elapsed_time(e1) / 30 / 6 * 1000, 2)
            del graph
        del xs, outs
        torch.cuda.empty_cache()
    return res


def served_leg(dev, rates: str, seconds: float) -> dict:
    """p50/p99 latency of the continuous-batching server under Poisson
    arrivals: 256 resident bfloat16 images (20% 3840x2160, the rest
    log-uniform in [240, 2160]), random sigma in [0.8, 4] and output subset."""
    import torch
    from tucuda_amd import serve_bench
    from tucuda_amd.serve import Server
    images = serve_bench.setup(256, 5, dev)
    torch.cuda.synchronize()
    out = {"workload": "Poisson arrivals; 256 resident bf16 images, 20% 3840x2160, rest log-uniform "
                       "[240,2160]; sigma U[0.8,4], radius ceil(4 sigma); random output subset; bf16 out",
           "seconds_per_rate": seconds, "latency": "completion - scheduled arrival", "rates": []}
    with Server(16 << 30) as srv:
        serve_bench.run_stream(srv, images, 500.0, 0.5, 104, check_every=0)
        for k, rate in enumerate(float(r) for r in rates.split(",") if r):
            res = serve_bench.run_stream(srv, images, rate, seconds, 110 + k)
            log(f"served {rate:.0f} req/s: p50 {res['p50_ms']} ms p99 {res['p99_ms']} ms, "
                f"mean batch {res['mean_batch']}, {res['mismatched']} of {res['checked']} checks differ")
            out["rates"].append({k2: res[k2] for k2 in (
                "rate_req_s", "requests", "served_req_s", "served_gpix_s", "p50_ms", "p99_ms", "max_ms",
                "mean_batch", "checked", "mismatched")})
        _, _, cap = serve_bench.run_capacity(srv, images, 2000, 177)
        out["capacity"] = cap
        log(f"served capacity: {cap['served_gpix_s']} Gpix/s")
    del images
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=3000,
                    help="untimed steps before the timed ones, run as the timed ones are (graph replays "
                         "back to back), so that the timed region starts at the GPU's steady clocks")
    ap.add_argument("--sets", type=int, default=10, help="rotating buffer sets (>= 5 defeats the 256 MB MALL)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--persistent", type=int, default=0, metavar="ROWS",
                    help="run each image through the persistent batched kernel with ROWS-row tiles "
                         "(4, 8 or 16) instead of one workgroup per tile")
    ap.add_argument("--eager", action="store_true",
                    help="launch every step from Python instead of replaying a captured HIP graph")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--serve-seconds", type=float, default=30.0,
                    help="seconds of Poisson arrivals per served-stream rate (0 skips the served leg)")
    ap.add_argument("--batch-reps", type=int, default=10,
                    help="graph replays of the batched leg (0 skips it)")
    ap.add_argument("--serve-rates", default="200,1000,4000", help="served-stream arrival rates, req/s")
    ap.add_argument("--sigma-sweep", action="store_true",
                    help="instead: time one fused six-output 1080p launch for several sigmas "
                         "(radius ceil(4 sigma)), float32 and bfloat16, and print one JSON line")
    ap.add_argument("--nrow", type=int, default=NROW, help="image rows (the headline workload is 1080)")
    ap.add_argument("--ncol", type=int, default=NCOL, help="image columns (the headline workload is 1920)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (reference point for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(images shard across GPUs with no communication; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import tucuda_amd
    from tucuda_amd import synth

    if args.sigma_sweep:
        print(json.dumps(sigma_sweep(torch.device("cuda:0"), args.nrow, args.ncol)), flush=True)
        return

    dev = torch.device("cuda:0")
    nrow, ncol = args.nrow, args.ncol
    img = synth.edge_image(1, nrow, ncol)
    g = tucuda_amd.FIRGaussianConvolver2(SIGMA)
    l0, l1, l2 = (x.size for x in tucuda_amd.gauss_lobes(SIGMA))
    r = [l0 - 1, l1, l2 - 1]
    radii = [(r[v], r[h]) for h, v in (tucuda_amd.METHOD_LOBES[m] for m in tucuda_amd.METHOD_ORDER)]
    alg_bytes = algorithmic_bytes(nrow, ncol, radii)

    d_img = torch.from_numpy(img).to(dev)
    sets = []
    for s in range(args.sets):
        ins = d_img.clone()
        outs = [torch.empty((nrow, ncol), dtype=torch.float32, device=dev) for _ in range(6)]
        sets.append((ins, outs))
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream(dev)
    if args.persistent:
        batches = []
        for ins, outs in sets:
            b = tucuda_amd.GaussBatch([(ins, outs, g._lobes())])
            b.set_tile_rows(args.persistent)
            batches.append(b)

        class _One:
            def __init__(self, b):
                self.b = b

            def convolve6(self, ins, outs):
                self.b.launch()
        by_set = {id(s[0]): _One(b) for s, b in zip(sets, batches)}

        class _Dispatch:
            def convolve6(self, ins, outs):
                by_set[id(ins)].convolve6(ins, outs)
        g = _Dispatch()
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    # one call per set from Python first: the operand tables are then
    # prebuilt when the graph is captured
    for k in range(min(args.warmup, args.sets)):
        ins, outs = sets[k]
        g.convolve6(ins, outs)
    torch.cuda.synchronize()

    # Timed region.  Default: one HIP graph holding the buffer sets' launches
    # in turn, at least 50 of them (the sets repeated), replayed until `steps`
    # launches have run (graph replay removes the per-call host overhead of
    # Python + ctypes; 50 launches per replay keep the cost of starting a
    # replay, ~8 us, below 0.2 us per launch).  Events on the launch stream
    # bracket each replay; the kernel-average duration is replay time /
    # launches per replay.
    if args.eager:
        for k in range(args.sets, args.warmup):
            ins, outs = sets[k % args.sets]
            g.convolve6(ins, outs)
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
        t_start = torch.cuda.Event(enable_timing=True)
        t_end = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t_start.record(stream)
        for k in range(args.steps):
            ins, outs = sets[k % args.sets]
            starts[k].record(stream)
            g.convolve6(ins, outs)
            ends[k].record(stream)
        t_end.record(stream)
        torch.cuda.synchronize()
        total_ms = t_start.elapsed_time(t_end)
        kern_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
        steps = args.steps
    else:
        per = args.sets * max(1, (LAUNCHES_PER_GRAPH + args.sets - 1) // args.sets)
        graph = torch.cuda.CUDAGraph()
        cap = torch.cuda.Stream(dev)
        cap.wait_stream(stream)
        with torch.cuda.stream(cap):
            with torch.cuda.graph(graph, stream=cap):
                for k in range(per):
                    ins, outs = sets[k % args.sets]
                    g.convolve6(ins, outs)
        stream.wait_stream(cap)
        # the untimed steps: replays back to back, as in the timed region
        for _ in range(max(2, (args.warmup - min(args.warmup, args.sets) + per - 1) // per)):
            graph.replay()
        torch.cuda.synchronize()
        reps = max(1, args.steps // per)
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
        t_start = torch.cuda.Event(enable_timing=True)
        t_end = torch.cuda.Event(enable_timing=True)
        rs = torch.cuda.current_stream(dev)
        t_start.record(rs)
        for r in range(reps):
            starts[r].record(rs)
            graph.replay()
            ends[r].record(rs)
        t_end.record(rs)
        torch.cuda.synchronize()
        total_ms = t_start.elapsed_time(t_end)
        kern_ms = [s.elapsed_time(e) / per for s, e in zip(starts, ends)]
        steps = reps * per
    avg_kern_ms = float(np.mean(kern_ms))
    med_kern_ms = float(np.median(kern_ms))
    ms_per_step = total_ms / steps
    gpix = nrow * ncol / (ms_per_step * 1e-3) / 1e9
    achieved = alg_bytes / (avg_kern_ms * 1e-3) / 1e9
    log(f"kernel avg {avg_kern_ms*1e3:.2f} us median {med_kern_ms*1e3:.2f} us; "
        f"wall per step {ms_per_step*1e3:.2f} us; {achieved:.0f} GB/s algorithmic "
        f"({'eager' if args.eager else 'graph'} launches)")

    traffic, traffic_stale = recorded_traffic()
    if traffic_stale:
        log("recorded HBM traffic is for other kernel sources: roofline.traffic = null")

    line = {
        "metric": METRIC,
        "value": round(gpix, 4),
        "unit": "Gpix/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded uniform [0,255] + 8 step edges)",
        "config": {
            "workload": f"{ncol}x{nrow} float32 image, sigma=3 (lobes 17/16/17, radius 16), "
                        "smooth+diffH+diffV+diffHH+diffHV+diffVV fused in one launch",
            "nrow": nrow, "ncol": ncol, "sigma": SIGMA, "outputs": 6,
            "buffer_sets": args.sets, "shift": True,
            "launch": "eager" if args.eager else f"hip-graph replay, {per} launches per graph",
            "kernel": f"persistent, {args.persistent}-row tiles" if args.persistent else kernel_label(),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_stale": traffic_stale,
            "kernel_us_avg": round(avg_kern_ms * 1e3, 3),
            "algorithmic_bytes": alg_bytes,
        },
    }
    # The cache-hot figure (one buffer set for every launch) beside the
    # rotating one that `achieved` uses.
    hot_us = hot_kernel_us(g, *sets[0])
    line["roofline"]["kernel_us_hot_single_buffer"] = round(hot_us, 3)
    log(f"cache-hot (one buffer set): {hot_us:.2f} us per launch")
    # Context for `frac`: what a plain write stream of the outputs' size, and
    # a sustained one, reach (the outputs are 85% of the kernel's bytes).
    wc = write_ceiling(dev, 6 * (nrow - 32) * (ncol - 32) * 4)
    line["roofline"]["write_ceiling"] = wc
    log(f"write ceiling: fill of the outputs' size {wc['fill_outputs_size_gbs']} GB/s "
        f"({wc['fill_outputs_size_us']} us), 1 GiB fill {wc['fill_1gib_gbs']} GB/s, 1 GiB copy {wc['copy_1gib_gbs']} GB/s")
    if args.batch_reps > 0:
        log("batched leg ...")
        line["batched"] = batched_leg(args.batch_reps)
        log(f"batched: bf16 {line['batched']['bf16_batched_ms']} ms, f32 plain (graph) "
            f"{line['batched']['f32_plain_graph_ms']} ms: {line['batched']['bf16_batched_vs_f32_plain_graph']}x")
    if args.serve_seconds > 0:
        line["served"] = served_leg(dev, args.serve_rates, args.serve_seconds)
    if not args.no_cpu:
        log("timing the CPU baseline ...")
        line["cpu_baseline"] = cpu_baseline(img, SIGMA, args.cpu_budget)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
