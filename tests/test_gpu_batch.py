"""Batched mode: many Gaussian jobs of uneven sizes and radii in one
persistent launch.  Each job must get exactly (bit for bit) what the
single-image fused launch writes for it, which in turn matches the oracle;
repeated launches and HIP-graph replays must not depend on leftover state;
the autotuner must leave the results unchanged."""
import numpy as np
import pytest

import oracle
from tests._util import assert_f32_parity, to_host
from tucuda_amd import synth

pytestmark = pytest.mark.gpu


def _jobs(dev, n, seed, dtype, rule, sizes=(40, 420)):
    import torch
    import tucuda_amd
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        h = int(np.exp(rng.uniform(np.log(sizes[0]), np.log(sizes[1]))))
        w = int(np.clip(np.exp(rng.uniform(np.log(sizes[0]), np.log(sizes[1]))), h / 2, 2 * h))
        sigma = float(np.float32(rng.uniform(0.8, 4.0)))
        img = synth.edge_image(seed * 100 + i, h, w)
        d_in = torch.from_numpy(img).to(dev).to(dtype)
        mask = int(rng.integers(1, 64)) if i % 3 else 0x3F
        outs = {m: torch.full((h, w), -7.0, dtype=dtype, device=dev)
                for k, m in enumerate(tucuda_amd.METHOD_ORDER) if mask >> k & 1}
        lobes = tucuda_amd.gauss_lobes(sigma, rule)
        jobs.append((d_in, outs, lobes, bool(i % 2 == 0), img, sigma))
    return jobs


def _single(dev, job):
    import torch
    import tucuda_amd
    d_in, outs, lobes, shift = job[:4]
    g = tucuda_amd.FIRGaussianConvolver2(1.0)
    g._lobe0, g._lobe1, g._lobe2 = lobes
    ref = {m: torch.full_like(t, -7.0) for m, t in outs.items()}
    g.convolve6(d_in, ref, shift)
    return ref


@pytest.mark.parametrize("rule", [0, 1])
def test_batch_matches_single_and_oracle(dev, rule):
    import torch
    import tucuda_amd
    jobs = _jobs(dev, 14, 3 + rule, torch.float32, rule)
    b = tucuda_amd.GaussBatch([j[:4] for j in jobs])
    info = b.info()
    assert info["jobs"] == 14 and info["tiles"] > 0
    for rep in range(3):                          # counters reset themselves between launches
        for j in jobs:
            for t in j[1].values():
                t.fill_(-7.0)
        b.launch()
        torch.cuda.synchronize()
        for j in jobs:
            ref = _single(dev, j)
            for m, t in j[1].items():
                assert torch.equal(t, ref[m]), (rep, m)
    # single-image launches against the oracle
    for j in jobs:
        d_in, outs, lobes, shift, img, sigma = j
        for m, t in outs.items():
            h, v = oracle.METHODS[m]
            want = np.full(img.shape, -7.0, np.float32)
            if rule == 0:
                oracle.fir2(lobes[h], lobes[v], img, want, shift)
            else:
                oracle.gauss_general(m, oracle.gauss_lobes(sigma, "f32", "ceil4"), img, want, shift)
            assert_f32_parity(to_host(t), want, f"sigma={sigma} {m} shift={shift}")
    b.close()


def test_batch_bf16_equals_single(dev):
    import torch
    import tucuda_amd
    jobs = _jobs(dev, 10, 7, torch.bfloat16, 1)
    b = tucuda_amd.GaussBatch([j[:4] for j in jobs])
    b.launch()
    torch.cuda.synchronize()
    for j in jobs:
        ref = _single(dev, j)
        for m, t in j[1].items():
            assert torch.equal(t, ref[m]), m


def test_batch_graph_replay(dev):
    """Capture one batched launch, replay it several times with eager work
    and synchronisations in between; every replay must reproduce the eager
    result (no memset nodes are recorded: the kernel resets its counters)."""
    import torch
    import tucuda_amd
    jobs = _jobs(dev, 8, 11, torch.float32, 0)
    b = tucuda_amd.GaussBatch([j[:4] for j in jobs])
    b.launch()
    torch.cuda.synchronize()
    eager = [{m: t.clone() for m, t in j[1].items()} for j in jobs]
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream(dev)
    s.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(s):
        with torch.cuda.graph(graph, stream=s):
            b.launch()
    torch.cuda.current_stream(dev).wait_stream(s)
    for rep in range(4):
        for j in jobs:
            for t in j[1].values():
                t.fill_(-7.0)               # the border is never written; the valid region must be
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for j, e in zip(jobs, eager):
            for m, t in j[1].items():
                assert torch.equal(t, e[m]), (rep, m)


def test_batch_autotune_and_cache(dev, tmp_path):
    import torch
    import tucuda_amd
    jobs = _jobs(dev, 12, 13, torch.float32, 1, sizes=(100, 600))
    b = tucuda_amd.GaussBatch([j[:4] for j in jobs])
    allowed = tucuda_amd.batch_tile_rows()
    rows, us = b.autotune(reps=3)
    assert rows in allowed and us > 0
    assert b.info()["tile_rows"] == rows
    for j in jobs:
        for t in j[1].values():
            t.fill_(-7.0)
    b.launch()
    torch.cuda.synchronize()
    for j in jobs:
        ref = _single(dev, j)
        for m, t in j[1].items():
            assert torch.equal(t, ref[m]), m
    path = str(tmp_path / "tune.jsonl")
    tucuda_amd.tune_cache_save(path)
    text = open(path).read()
    assert f'"tile_rows": {rows}' in text
    tucuda_amd.tune_cache_load(path)
    b2 = tucuda_amd.GaussBatch([j[:4] for j in jobs])     # same shape: the cached height is used
    assert b2.info()["tile_rows"] == rows
    for r in (allowed[0], allowed[2]):
        b2.set_tile_rows(r)
        assert b2.info()["tile_rows"] == r
