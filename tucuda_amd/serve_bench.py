"""Served-stream measurement: Poisson arrivals into the continuous-batching server.

    python -m tucuda_amd.serve_bench [--rates 200,1000,4000] [--duration 30]

Workload: a resident set of bfloat16 input images, 80% with widths and
heights log-uniform over [240, 2160] (aspect ratio clamped to [0.5, 2]) and
20% of 3840x2160; each request takes one of them at random, a sigma uniform
in [0.8, 4.0] (lobe radius ceil(4 sigma)) and a random non-empty subset of
the six outputs.  Requests arrive as a Poisson process at each rate for
``--duration`` seconds; latency is completion time minus the scheduled
arrival time, so it includes any lag of the submitting loop.  Outputs are
released as soon as they are polled.

Images and outputs have rows padded to 256 bytes (the ``cu::Array2``
layout; the server's pool pads output rows to 128 bytes).

A capacity run then submits ``--capacity-requests`` requests in one call to
a server whose pool (``--capacity-pool-gib``) holds the whole burst's
outputs, so the scheduler, not the submitting or draining Python loop, sets
the pace; it reports the served throughput, latency and batch sizes under
that saturation next to the plain per-image forms on the same requests,
float32 and bfloat16 images:
one fused launch per request, and one launch per requested output (the
method-per-output calls of FIRGaussianConvolver2); all launches captured in
one HIP graph, the best case for per-image launches (no host cost between
them).

A few served requests per run are kept and compared, bit for bit on the
region the filter writes, with the single-image launch of the same request.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np


def make_images(n: int, seed: int, frac_4k: float = 0.2):
    """(nrow, ncol) of the resident input set."""
    from . import synth
    u = synth.uniform01(seed, 2 * n).reshape(n, 2)
    lo, hi = math.log(240), math.log(2160)
    n4k = int(round(frac_4k * n))
    out = []
    for i, (a, b) in enumerate(u):
        if i < n4k:
            out.append((2160, 3840))
            continue
        h = int(round(math.exp(lo + (hi - lo) * a)))
        w = int(round(math.exp(lo + (hi - lo) * b)))
        out.append((h, int(min(max(w, h / 2), 2 * h))))
    return out


def device_image(seed: int, h: int, w: int, dev, dtype, nedges: int = 8, amplitude: float = 64.0):
    """Uniform [0,255) image plus ``nedges`` step edges, generated on the
    device, in the layout of ``cu::Array2`` (rows padded to 256 bytes)."""
    import torch
    import tucuda_amd
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    img = torch.rand((h, w), generator=g, device=dev, dtype=torch.float32) * 255.0
    par = torch.rand((nedges, 3), generator=g, device=dev, dtype=torch.float64).cpu().numpy()
    yy = torch.arange(h, device=dev, dtype=torch.float32).view(h, 1)
    xx = torch.arange(w, device=dev, dtype=torch.float32).view(1, w)
    for t, cx, cy in par:
        th = 2.0 * math.pi * t
        img += amplitude * (((xx - cx * w) * math.cos(th) + (yy - cy * h) * math.sin(th)) > 0).float()
    out = tucuda_amd.array2(h, w, dtype, dev)
    out.copy_(img)
    return out


def _array2_like(x, dtype):
    import tucuda_amd
    y = tucuda_amd.array2(x.shape[0], x.shape[1], dtype, x.device)
    y.copy_(x)
    return y


def make_requests(n: int, nimages: int, seed: int):
    """(image index, sigma, mask) of n requests."""
    from . import synth
    u = synth.uniform01(seed, 3 * n).reshape(n, 3)
    idx = np.minimum((u[:, 0] * nimages).astype(np.int64), nimages - 1)
    sigma = (0.8 + 3.2 * u[:, 1]).astype(np.float32)
    mask = 1 + np.minimum((u[:, 2] * 63).astype(np.int64), 62)
    return [(int(i), float(s), int(m)) for i, s, m in zip(idx, sigma, mask)]


def _pct(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if len(xs) else float("nan")


class _Checker:
    """Keeps a few served requests unreleased and compares them with the
    single-image launch afterwards."""

    def __init__(self, srv, every: int):
        self.srv, self.every, self.kept, self.count = srv, every, [], 0

    def take(self, c, req) -> bool:
        self.count += 1
        if self.every and self.count % self.every == 0 and len(self.kept) < 16:
            self.kept.append((c, req))
            return True
        return False

    def verify(self, images):
        import torch
        import tucuda_amd
        bad = 0
        for c, (ii, sigma, mask) in self.kept:
            x = images[ii]
            g = tucuda_amd.FIRGaussianConvolver2(sigma, rule=tucuda_amd.TUCU_RULE_CEIL4SIGMA)
            for k, m in enumerate(tucuda_amd.METHOD_ORDER):
                if not mask >> k & 1:
                    continue
                ref = torch.full_like(x, float("nan"))
                getattr(g, m)(x, ref)
                got = torch.empty_like(x)
                self.srv.copy_output(c, k, got)
                valid = ~torch.isnan(ref)
                if not torch.equal(got[valid], ref[valid]):
                    bad += 1
            self.srv.release(c["id"])
        n = len(self.kept)
        self.kept = []
        return n, bad


def run_stream(srv, images, rate: float, duration: float, seed: int, check_every: int = 500):
    """Poisson arrivals at ``rate`` req/s for ``duration`` s; returns a dict."""
    rng = np.random.default_rng(seed)
    n = max(1, int(rng.poisson(rate * duration)))
    gaps = rng.exponential(1.0 / rate, n)
    reqs = make_requests(n, len(images), seed + 1)
    st0 = srv.stats()
    t0 = srv.now() + 0.02
    arrival = t0 + np.cumsum(gaps) - gaps[0]
    lat, lag, pix, done = [], [], 0, 0
    by_id = {}
    chk = _Checker(srv, check_every)

    def drain():
        nonlocal pix, done
        for c in srv.poll():
            req = by_id.pop(c["id"])
            if c["status"] != 0:
                raise RuntimeError(f"request {c['id']} failed with status {c['status']}")
            lat.append(c["completed"] - c["arrival"])
            h, w = images[req[0]].shape
            pix += h * w
            done += 1
            if not chk.take(c, req):
                srv.release(c["id"])

    i = 0
    while i < n:
        now = srv.now()
        while i < n and arrival[i] <= now:
            ii, sigma, mask = reqs[i]
            rid = srv.submit(images[ii], sigma, mask, arrival=float(arrival[i]))
            lag.append(srv.now() - arrival[i])
            by_id[rid] = reqs[i]
            i += 1
        drain()
        if i < n:
            wait = arrival[i] - srv.now()
            if wait > 4e-4:
                time.sleep(wait - 3e-4)
    t_last = float(arrival[-1])
    deadline = time.monotonic() + 60.0
    while by_id and time.monotonic() < deadline:
        drain()
        time.sleep(1e-4)
    if by_id:
        raise RuntimeError(f"{len(by_id)} requests still pending after 60 s")
    t_end = srv.now()
    checked, bad = chk.verify(images)
    st1 = srv.stats()
    batches = st1["batches"] - st0["batches"]
    span = t_end - float(arrival[0])
    return {
        "rate_req_s": rate, "duration_s": duration, "requests": n,
        "served_req_s": round(n / span, 1),
        "served_gpix_s": round(pix / span / 1e9, 3),
        "p50_ms": round(_pct(lat, 50) * 1e3, 3),
        "p99_ms": round(_pct(lat, 99) * 1e3, 3),
        "mean_ms": round(float(np.mean(lat)) * 1e3, 3),
        "max_ms": round(float(np.max(lat)) * 1e3, 3),
        "submit_lag_p99_ms": round(_pct(lag, 99) * 1e3, 3),
        "drain_after_last_arrival_ms": round((t_end - t_last) * 1e3, 3),
        "batches": batches, "mean_batch": round(n / max(batches, 1), 2),
        "pool_peak_gib": round(st1["pool_peak"] / 2**30, 3),
        "checked": checked, "mismatched": bad,
    }


def run_capacity(srv, images, n: int, seed: int):
    """n requests submitted at once in one call (tucu_server_submit_many,
    all stamped with the submission time): the scheduler sees a full queue,
    so the served rate is the GPU's and the server's, not the submitting
    thread's; served throughput in Gpix/s of input, and the latency and batch
    sizes under that saturation."""
    reqs = make_requests(n, len(images), seed)
    pix = sum(images[ii].shape[0] * images[ii].shape[1] for ii, _, _ in reqs)
    pending = set()
    lat = []
    st0 = srv.stats()

    def drain():
        for c in srv.poll():
            if c["status"] != 0:
                raise RuntimeError(f"request {c['id']} failed with status {c['status']}")
            pending.discard(c["id"])
            lat.append(c["completed"] - c["arrival"])
            srv.release(c["id"])

    t0 = srv.now()
    pending.update(srv.submit_many([(images[ii], sigma, mask) for ii, sigma, mask in reqs], arrival=t0))
    deadline = time.monotonic() + 120.0
    while pending and time.monotonic() < deadline:
        drain()
        time.sleep(5e-5)
    if pending:
        raise RuntimeError(f"{len(pending)} requests still pending after 120 s")
    dt = srv.now() - t0
    st1 = srv.stats()
    batches = st1["batches"] - st0["batches"]
    return reqs, pix, {"requests": n, "seconds": round(dt, 4), "served_req_s": round(n / dt, 1),
                       "served_gpix_s": round(pix / dt / 1e9, 2),
                       "p50_ms": round(_pct(lat, 50) * 1e3, 3), "p99_ms": round(_pct(lat, 99) * 1e3, 3),
                       "batches": batches, "mean_batch": round(n / max(batches, 1), 1),
                       "max_batch": st1["max_batch"], "pool_peak_gib": round(st1["pool_peak"] / 2**30, 3)}


def plain_per_image(images, reqs, dtype, reps: int = 3, per_method: bool = False):
    """ms for one fused launch per request (outputs of the request's mask),
    or with ``per_method`` one launch per requested output (the way the
    method-per-output API is called), every launch captured in one HIP
    graph; float32 or bfloat16 images, inputs and outputs with rows padded
    to 256 bytes like the served form's."""
    import torch
    import tucuda_amd
    dev = images[0].device
    xs = [_array2_like(x, dtype) for x in images]
    maxpx = max(x.stride(0) * x.shape[0] for x in xs)
    flat = [torch.empty(maxpx, dtype=dtype, device=dev) for _ in range(6)]
    convs = {}
    calls = []
    for ii, sigma, mask in reqs:
        x = xs[ii]
        h, w = x.shape
        if sigma not in convs:
            convs[sigma] = tucuda_amd.FIRGaussianConvolver2(sigma, rule=tucuda_amd.TUCU_RULE_CEIL4SIGMA)
        ld = x.stride(0)
        outs = {m: flat[k][: h * ld].view(h, ld)[:, :w] for k, m in enumerate(tucuda_amd.METHOD_ORDER)
                if mask >> k & 1}
        if per_method:
            calls.extend((convs[sigma], x, {m: t}) for m, t in outs.items())
        else:
            calls.append((convs[sigma], x, outs))

    def run():
        for g, x, outs in calls:
            g.convolve6(x, outs)

    run()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream(dev)
    s.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(s):
        with torch.cuda.graph(graph, stream=s):
            run()
    torch.cuda.current_stream(dev).wait_stream(s)
    graph.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        graph.replay()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del graph, flat, xs
    return ms


def setup(nimages: int, seed: int, dev):
    import torch
    sizes = make_images(nimages, seed)
    return [device_image(seed * 1000 + i, h, w, dev, torch.bfloat16) for i, (h, w) in enumerate(sizes)]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--rates", default="200,1000,4000")
    ap.add_argument("--duration", type=float, default=30.0)
    ap.add_argument("--images", type=int, default=256, help="resident input images")
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--pool-gib", type=float, default=16.0)
    ap.add_argument("--tile-rows", type=int, default=0, help="batched tile rows (0: the library default)")
    ap.add_argument("--max-inflight", type=int, default=2)
    ap.add_argument("--capacity-requests", type=int, default=2000)
    ap.add_argument("--capacity-pool-gib", type=float, default=48.0,
                    help="output pool of the capacity run's server (room for the whole burst's outputs)")
    ap.add_argument("--max-batch-mpix", type=int, default=0, help="input megapixels per launch (0: the library default)")
    ap.add_argument("--no-plain", action="store_true", help="skip the per-image comparison")
    args = ap.parse_args(argv)

    import torch
    from .serve import Server

    dev = torch.device("cuda:0")
    images = setup(args.images, args.seed, dev)
    torch.cuda.synchronize()
    result = {"workload": f"{args.images} resident bf16 images (20% 3840x2160, rest log-uniform [240,2160]), "
                          f"sigma U[0.8,4], radius ceil(4 sigma), random output subset",
              "pool_gib": args.pool_gib, "tile_rows": args.tile_rows, "max_inflight": args.max_inflight,
              "streams": []}
    opts = dict(tile_rows=args.tile_rows, max_inflight=args.max_inflight, max_batch_mpix=args.max_batch_mpix)
    with Server(int(args.pool_gib * 2**30), **opts) as srv:
        run_stream(srv, images, 500.0, 0.5, args.seed + 99, check_every=0)        # warm-up
        for k, rate in enumerate(float(r) for r in args.rates.split(",") if r):
            res = run_stream(srv, images, rate, args.duration, args.seed + 10 * k)
            result["streams"].append(res)
            print(json.dumps({"progress": res}), file=sys.stderr, flush=True)
    if args.capacity_requests > 0:
        with Server(int(args.capacity_pool_gib * 2**30), **opts) as srv:
            run_capacity(srv, images, 200, args.seed + 98)                      # warm-up
            reqs, pix, cap = run_capacity(srv, images, args.capacity_requests, args.seed + 77)
            cap["pool_gib"] = args.capacity_pool_gib
            result["capacity"] = cap
    if args.capacity_requests > 0 and not args.no_plain:
        cap = result["capacity"]
        for form, pm in (("fused", False), ("per_method", True)):
            for dname, dt in (("float32", torch.float32), ("bfloat16", torch.bfloat16)):
                ms = plain_per_image(images, reqs, dt, per_method=pm)
                cap[f"plain_{form}_{dname}_graph_ms"] = round(ms, 3)
                cap[f"plain_{form}_{dname}_gpix_s"] = round(pix / ms / 1e6, 2)
                cap[f"served_bf16_vs_plain_{form}_{dname}"] = round(cap["served_gpix_s"] / (pix / ms / 1e6), 2)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    sys.exit(main())
