"""tucu_batch_run_static: the batched launch that allocates nothing (its
tables go to a static device scratch array).  Called through the C ABI, it
must give what tucu_batch_create + tucu_batch_launch give for the same jobs,
bit for bit, on two calls in a row (the second smaller, so leftovers of the
first in the shared scratch would show), and refuse a batch whose tables
exceed the scratch before launching anything."""
import ctypes

import numpy as np
import pytest

from tucuda_amd import synth

pytestmark = pytest.mark.gpu


def _jobs(dev, n, seed, sizes):
    import torch
    import tucuda_amd
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        h = int(rng.integers(*sizes))
        w = int(rng.integers(*sizes))
        sigma = float(np.float32(rng.uniform(0.8, 4.0)))
        x = torch.from_numpy(synth.edge_image(seed * 1000 + i, h, w)).to(dev)
        mask = int(rng.integers(1, 64)) if i % 3 else 0x3F
        outs = [torch.full((h, w), -3.0, device=dev) if mask >> k & 1 else None for k in range(6)]
        lobes = tucuda_amd.gauss_lobes(sigma, tucuda_amd.TUCU_RULE_CEIL4SIGMA)
        jobs.append((x, outs, lobes, i % 2 == 0))
    return jobs


def _job_array(jobs):
    from tucuda_amd import _capi
    from tucuda_amd.filters import image_desc
    arr = (_capi.TucuGaussJob * len(jobs))()
    keep = []
    for j, (x, outs, lobes, shift) in zip(arr, jobs):
        lob = [np.ascontiguousarray(l, np.float32) for l in lobes]
        keep.append(lob)
        j.in_ = image_desc(x)
        mask = 0
        for k, t in enumerate(outs):
            if t is not None:
                j.out[k] = image_desc(t)
                mask |= 1 << k
        j.lobe0, j.lobe1, j.lobe2 = (l.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for l in lob)
        j.n0, j.n1, j.n2 = (l.size for l in lob)
        j.mask, j.shift = mask, int(shift)
    return arr, keep


def _run_static(dev, jobs):
    import torch
    from tucuda_amd import _capi
    arr, keep = _job_array(jobs)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _capi.check(_capi.load().tucu_batch_run_static(arr, len(jobs), stream))


def _run_batch(dev, jobs):
    import torch
    from tucuda_amd import _capi
    arr, keep = _job_array(jobs)
    h = ctypes.c_void_p()
    _capi.check(_capi.load().tucu_batch_create(ctypes.byref(h), arr, len(jobs)))
    try:
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _capi.check(_capi.load().tucu_batch_launch(h, stream))
        torch.cuda.synchronize()
    finally:
        _capi.load().tucu_batch_destroy(h)


def test_static_batch_equals_allocating_batch(dev):
    import torch
    for n, seed, sizes in ((12, 1, (90, 700)), (5, 2, (40, 300))):
        jobs = _jobs(dev, n, seed, sizes)
        _run_static(dev, jobs)
        torch.cuda.synchronize()
        got = [[t.clone() if t is not None else None for t in j[1]] for j in jobs]
        for j in jobs:
            for t in j[1]:
                if t is not None:
                    t.fill_(-3.0)
        _run_batch(dev, jobs)
        for j, g in zip(jobs, got):
            for t, u in zip(j[1], g):
                if t is not None:
                    assert torch.equal(t, u), (n, seed)
                    assert not torch.equal(t, torch.full_like(t, -3.0))


def test_static_batch_refuses_oversized_tables(dev):
    import torch
    import tucuda_amd
    from tucuda_amd import _capi
    x = torch.zeros((48, 48), device=dev)
    o = torch.full((48, 48), -3.0, device=dev)
    lobes = tucuda_amd.gauss_lobes(1.0)
    jobs = [(x, [o, None, None, None, None, None], lobes, True)] * 6000
    with pytest.raises(_capi.TucuError, match="static scratch"):
        _run_static(dev, jobs)
    torch.cuda.synchronize()
    assert torch.equal(o, torch.full_like(o, -3.0))          # nothing was launched
