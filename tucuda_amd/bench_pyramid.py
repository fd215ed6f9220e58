"""Measurement of cu::subsample and of the pyramid step (smooth, then halve).

    python -m tucuda_amd.bench_pyramid

Workloads (seeded synthetic images, inputs resident in HBM):
  * subsample of a 1920x1080 float32 image with identity1x1 (plain
    decimation), diffHV3x3 and gaussian5x5, and of a 3840x2160 u_char image
    with identity1x1 and a 3840x2160 float32 image with gaussian5x5;
  * a four-level Gaussian pyramid of the 1920x1080 float32 image: at each
    level FIRGaussianConvolver2(1.0).smooth into a full-size buffer, then
    subsample(identity1x1) into the next level (the reference's pyramid
    step, Testprog/subsampleTest/main.cu after a smooth); and the same
    pyramid with the fused step (smoothSubsample, one launch per level);
  * one fused step on 1920x1080 float32 (sigma 1 and 3), 3840x2160 float32
    (sigma 3) and 3840x2160 u_char (sigma 2) images; algorithmic bytes: the
    input once plus the quarter-size float32 output once.
Launches are replayed from a captured HIP graph that goes 12 times over 4
rotating buffer sets, after 40 untimed replays.
Algorithmic bytes of a subsample: the input rows the operator reads (every
second row for the 1-row operators, all rows otherwise) once, plus the
nrow/2 x ncol/2 output once.  Prints one JSON line: microseconds per launch
and GB/s, plus the fraction of the 8 TB/s HBM peak.
"""
from __future__ import annotations

import json
import sys
import time

HBM_PEAK_GB_S = 8000.0
ONE_ROW_OPS = ("identity1x1", "diffH1x3", "diffHH1x3")


def _graph_us(run, nsets, dev, reps=20, rounds=12, warmup=40):
    """Microseconds per launch (per set) from a graph holding `rounds` run()'s,
    so that the cost of starting a replay adds little per launch, after
    `warmup` untimed replays (the GPU's clocks settle)."""
    import torch
    run()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    st = torch.cuda.Stream(dev)
    st.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(st):
        with torch.cuda.graph(graph, stream=st):
            for _ in range(rounds):
                run()
    torch.cuda.current_stream(dev).wait_stream(st)
    for _ in range(warmup):
        graph.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    rs = torch.cuda.current_stream(dev)
    e0.record(rs)
    for _ in range(reps):
        graph.replay()
    e1.record(rs)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1000 / (reps * rounds * nsets)


def subsample_bytes(nrow, ncol, in_es, out_es, op):
    rows = nrow // 2 if op in ONE_ROW_OPS else nrow
    return rows * ncol * in_es + (nrow // 2) * (ncol // 2) * out_es


def main(argv=None):
    import torch
    import tucuda_amd
    from . import synth

    dev = torch.device("cuda:0")
    f1080 = torch.from_numpy(synth.edge_image(3, 1080, 1920)).to(dev)
    f4k = torch.from_numpy(synth.edge_image(4, 2160, 3840)).to(dev)
    u4k = torch.from_numpy(synth.u8_image(5, 2160, 3840)).to(dev)
    cases = [("1080p_f32_identity1x1", f1080, "identity1x1"),
             ("1080p_f32_diffHV3x3", f1080, "diffHV3x3"),
             ("1080p_f32_gaussian5x5", f1080, "gaussian5x5"),
             ("2160p_u8_identity1x1", u4k, "identity1x1"),
             ("2160p_f32_gaussian5x5", f4k, "gaussian5x5")]
    res = {"subsample": {}, "pyramid": {}}
    nsets = 4
    # about 50 ms of launches first, so that the first timed case does not
    # run while the clocks ramp up from idle
    warm = tucuda_amd.array2(1080 // 2, 1920 // 2, torch.float32, dev)
    torch.cuda.synchronize()
    t_end = time.perf_counter() + 0.05
    while time.perf_counter() < t_end:
        for _ in range(20):
            tucuda_amd.subsample(f1080, warm, "gaussian5x5")
        torch.cuda.synchronize()
    for name, img, op in cases:
        nrow, ncol = img.shape
        ins = [tucuda_amd.array2(nrow, ncol, img.dtype, dev) for _ in range(nsets)]
        for x in ins:
            x.copy_(img)
        outs = [tucuda_amd.array2(nrow // 2, ncol // 2, torch.float32, dev) for _ in range(nsets)]

        def run():
            for x, o in zip(ins, outs):
                tucuda_amd.subsample(x, o, op)

        us = _graph_us(run, nsets, dev)
        nbytes = subsample_bytes(nrow, ncol, img.element_size(), 4, op)
        res["subsample"][name] = {"us": round(us, 2), "gb_s": round(nbytes / us / 1e3, 1),
                                  "frac": round(nbytes / us / 1e3 / HBM_PEAK_GB_S, 3), "bytes": nbytes}

    # four-level pyramid: smooth (sigma 1, reference lobes), then halve
    g = tucuda_amd.FIRGaussianConvolver2(1.0)
    levels = [(1080 >> k, 1920 >> k) for k in range(4)]
    sets = []
    for _ in range(nsets):
        lv = [tucuda_amd.array2(*levels[0], torch.float32, dev)]
        lv[0].copy_(f1080)
        sm = []
        for k in range(3):
            sm.append(tucuda_amd.array2(*levels[k], torch.float32, dev))
            lv.append(tucuda_amd.array2(*levels[k + 1], torch.float32, dev))
        sets.append((lv, sm))

    def pyramid():
        for lv, sm in sets:
            for k in range(3):
                g.smooth(lv[k], sm[k])
                tucuda_amd.subsample(sm[k], lv[k + 1], "identity1x1")

    us = _graph_us(pyramid, nsets, dev)
    res["pyramid"]["1080p_f32_sigma1_4levels"] = {
        "us": round(us, 2), "launches": 6,
        "levels": ["x".join(map(str, s[::-1])) for s in levels]}

    # the same pyramid with the fused step: smoothSubsample, three launches
    def pyramid_fused():
        for lv, _ in sets:
            for k in range(3):
                g.smoothSubsample(lv[k], lv[k + 1])

    us = _graph_us(pyramid_fused, nsets, dev)
    res["pyramid"]["1080p_f32_sigma1_4levels_fused"] = {"us": round(us, 2), "launches": 3}

    # one fused step on larger images: input read once, quarter-size output written once
    for name, img, sigma in (("1080p_f32_sigma1", f1080, 1.0), ("1080p_f32_sigma3", f1080, 3.0),
                             ("2160p_f32_sigma3", f4k, 3.0), ("2160p_u8_sigma2", u4k, 2.0)):
        gs = tucuda_amd.FIRGaussianConvolver2(sigma)
        nrow, ncol = img.shape
        ins = [tucuda_amd.array2(nrow, ncol, img.dtype, dev) for _ in range(nsets)]
        for x in ins:
            x.copy_(img)
        outs = [tucuda_amd.array2(nrow // 2, ncol // 2, torch.float32, dev) for _ in range(nsets)]

        def step():
            for x, o in zip(ins, outs):
                gs.smoothSubsample(x, o)

        us = _graph_us(step, nsets, dev)
        nbytes = nrow * ncol * img.element_size() + (nrow // 2) * (ncol // 2) * 4
        res["pyramid"]["step_" + name] = {"us": round(us, 2), "gb_s": round(nbytes / us / 1e3, 1),
                                           "frac": round(nbytes / us / 1e3 / HBM_PEAK_GB_S, 3), "bytes": nbytes}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    sys.exit(main())
