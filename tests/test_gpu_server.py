"""Continuous-batching server: every served output equals, bit for bit on
the region the filter writes, what the single-image launch writes for the
request -- whether the request ran alone or mixed into batches with others,
and whether or not the output pool throttled admission.  Served bfloat16
results are also held to the float64 oracle (normwise max error <= 1e-2)."""
import threading
import time

import numpy as np
import pytest

import oracle
from tests._util import bf16_bits_to_f32, to_host
from tucuda_amd import synth

pytestmark = pytest.mark.gpu

DTYPE_NAMES = ("bfloat16", "float32")


def _requests(dev, n, seed, dtype, sizes=(72, 360)):
    import torch
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        h = int(np.exp(rng.uniform(np.log(sizes[0]), np.log(sizes[1]))))
        w = int(np.clip(np.exp(rng.uniform(np.log(sizes[0]), np.log(sizes[1]))), h / 2, 2 * h))
        img = synth.edge_image(seed * 1000 + i, h, w)
        sigma = float(np.float32(rng.uniform(0.8, 4.0)))
        mask = int(rng.integers(1, 64))
        out.append((torch.from_numpy(img).to(dev).to(dtype), sigma, mask, img))
    return out


def _wait(srv, ids, timeout=60.0):
    got = {}
    t_end = time.monotonic() + timeout
    while len(got) < len(ids) and time.monotonic() < t_end:
        for c in srv.poll():
            got[c["id"]] = c
        time.sleep(1e-4)
    assert len(got) == len(ids), f"{len(ids) - len(got)} requests did not complete"
    return got


def _outputs(srv, c, x, mask):
    import torch
    res = {}
    for k in range(6):
        if mask >> k & 1:
            t = torch.empty_like(x)
            srv.copy_output(c, k, t)
            res[k] = t
    return res


def _single(x, sigma, mask):
    """Single-image launch of one request; unwritten pixels stay NaN."""
    import torch
    import tucuda_amd
    g = tucuda_amd.FIRGaussianConvolver2(sigma, rule=tucuda_amd.TUCU_RULE_CEIL4SIGMA)
    outs = {tucuda_amd.METHOD_ORDER[k]: torch.full_like(x, float("nan")) for k in range(6) if mask >> k & 1}
    g.convolve6(x, outs)
    torch.cuda.synchronize()
    return {tucuda_amd.METHOD_ORDER.index(m): t for m, t in outs.items()}


def _same(got, ref, what):
    import torch
    valid = ~torch.isnan(ref)
    assert int(valid.sum()) > 0, what
    assert torch.equal(got[valid], ref[valid]), what


@pytest.mark.parametrize("dname", DTYPE_NAMES)
def test_alone_equals_mix(dev, dname):
    import torch
    from tucuda_amd.serve import Server
    dtype = getattr(torch, dname)
    reqs = _requests(dev, 36, 21, dtype)
    with Server(1 << 30, in_dtype=dtype, out_dtype=dtype) as srv:
        alone = []
        for x, sigma, mask, _ in reqs:                        # one at a time
            rid = srv.submit(x, sigma, mask)
            c = _wait(srv, [rid])[rid]
            assert c["status"] == 0
            alone.append(_outputs(srv, c, x, mask))
            srv.release(rid)
        b0 = srv.stats()["batches"]
        assert b0 == len(reqs)

        # the mix: a burst, then requests trickling in from another thread
        # while the first batches are on the device
        ids = [None] * len(reqs)
        for i in range(len(reqs) // 2):
            x, sigma, mask, _ = reqs[i]
            ids[i] = srv.submit(x, sigma, mask)

        def trickle():
            for i in range(len(reqs) // 2, len(reqs)):
                x, sigma, mask, _ = reqs[i]
                ids[i] = srv.submit(x, sigma, mask)
                time.sleep(2e-4)

        th = threading.Thread(target=trickle)
        th.start()
        th.join()
        got = _wait(srv, ids)
        st = srv.stats()
        assert st["max_batch"] > 1, "the burst was not batched"
        assert st["batches"] - b0 < len(reqs)
        for i, (x, sigma, mask, _) in enumerate(reqs):
            c = got[ids[i]]
            assert c["status"] == 0
            assert c["arrival"] <= c["admitted"] <= c["launched"] <= c["completed"]
            mixed = _outputs(srv, c, x, mask)
            ref = _single(x, sigma, mask)
            for k in mixed:
                _same(alone[i][k], ref[k], f"alone req {i} out {k} sigma={sigma}")
                _same(mixed[k], ref[k], f"mixed req {i} out {k} sigma={sigma}")
            srv.release(ids[i])
        assert srv.stats()["pool_used"] == 0


def test_pool_backpressure(dev):
    """A pool that holds about two requests: the rest wait for memory, are
    admitted as outputs are released, and still get the same results."""
    import torch
    from tucuda_amd.serve import Server
    reqs = _requests(dev, 12, 33, torch.bfloat16, sizes=(200, 300))
    per_req = max(6 * ((x.shape[1] * 2 + 127) // 128 * 128) * x.shape[0] for x, *_ in reqs)
    pool = 2 * per_req + 4096
    with Server(pool, max_pool_bytes=pool) as srv:           # a fixed pool: no growth
        ids = [srv.submit(x, s, 0x3F) for x, s, _, _ in reqs]
        done = {}
        t_end = time.monotonic() + 60
        while len(done) < len(ids) and time.monotonic() < t_end:
            for c in srv.poll():
                assert c["status"] == 0
                i = ids.index(c["id"])
                x, sigma, _, _ = reqs[i]
                got = _outputs(srv, c, x, 0x3F)
                ref = _single(x, sigma, 0x3F)
                for k in got:
                    _same(got[k], ref[k], f"req {i} out {k}")
                srv.release(c["id"])
                done[c["id"]] = c
            time.sleep(1e-4)
        assert len(done) == len(ids)
        st = srv.stats()
        assert st["pool_peak"] <= pool
        assert st["pool_used"] == 0
        # admission is first in, first out: requests are admitted in submission order
        order = sorted(done.values(), key=lambda c: (c["admitted"], c["id"]))
        assert [c["id"] for c in order] == sorted(ids)


def test_served_bf16_against_oracle(dev):
    """Served bfloat16 results against the float64 oracle run on the stored
    (bfloat16) input with the same lobes: normwise max error <= 1e-2, the
    bound the precision mode states for the error the filter adds."""
    import torch
    import tucuda_amd
    from tucuda_amd.serve import Server
    reqs = _requests(dev, 8, 45, torch.bfloat16, sizes=(72, 200))
    with Server(1 << 28) as srv:
        ids = [srv.submit(x, s, 0x3F) for x, s, _, _ in reqs]
        got = _wait(srv, ids)
        for rid, (x, sigma, _, _) in zip(ids, reqs):
            outs = _outputs(srv, got[rid], x, 0x3F)
            xin = bf16_bits_to_f32(to_host(x))
            lobes64 = oracle.gauss_lobes(sigma, "f64", "ceil4")      # FP64 coefficients too
            for k, m in enumerate(tucuda_amd.METHOD_ORDER):
                want = np.full(xin.shape, np.nan, np.float32)
                oracle.gauss_general(m, lobes64, xin, want, True, precision="f64")
                valid = ~np.isnan(want)
                got_k = bf16_bits_to_f32(to_host(outs[k])).astype(np.float64)
                d = got_k[valid] - want[valid]
                err = np.abs(d).max() / np.abs(want[valid].astype(np.float64)).max()
                assert err <= 1e-2, (m, sigma, err)
            srv.release(rid)


def test_server_errors(dev):
    import torch
    import tucuda_amd
    from tucuda_amd.serve import Server
    x = torch.zeros((64, 64), dtype=torch.bfloat16, device=dev)
    with Server(1 << 20, max_pool_bytes=1 << 20) as srv:
        with pytest.raises(tucuda_amd.TucuError):
            srv.submit(x.float(), 2.0, 1)                     # dtype differs from the server's
        with pytest.raises(tucuda_amd.TucuError):
            srv.submit(x, 2.0, 0)                             # nothing requested
        with pytest.raises(tucuda_amd.TucuError):
            srv.submit(x, 2.0, 1 << 6)                        # unknown output
        big = torch.zeros((1024, 1024), dtype=torch.bfloat16, device=dev)
        with pytest.raises(tucuda_amd.TucuError):
            srv.submit(big, 2.0, 0x3F)                        # larger than the pool may grow
        with pytest.raises(tucuda_amd.TucuError):
            srv.release(12345)                                # unknown id
        # an image too small for the filter completes with nothing written
        tiny = torch.zeros((4, 4), dtype=torch.bfloat16, device=dev)
        rid = srv.submit(tiny, 3.0, 0x3F)
        c = _wait(srv, [rid])[rid]
        assert c["status"] == 0
        srv.release(rid)
        with pytest.raises(tucuda_amd.TucuError):
            srv.release(rid)                                  # already released


def test_release_refuses_unfinished(dev):
    """Releasing a request that has not completed (here: queued behind the
    pool, which a completed request still holds) is refused and leaves the
    request intact; it completes normally once memory frees up."""
    import torch
    import tucuda_amd
    from tucuda_amd.serve import Server
    x = torch.zeros((256, 512), dtype=torch.bfloat16, device=dev)
    with Server(1 << 20, max_pool_bytes=1 << 20) as srv:     # 1 MiB of outputs, fixed
        ra = srv.submit(x, 2.0, 0x03)                         # 2 outputs: 512 KiB
        _wait(srv, [ra])
        rb = srv.submit(x, 2.0, 0x07)                         # 3 outputs: 768 KiB, waits for ra's memory
        time.sleep(0.05)
        assert srv.poll() == []
        with pytest.raises(tucuda_amd.TucuError, match="not completed"):
            srv.release(rb)
        srv.release(ra)
        cb = _wait(srv, [rb])[rb]
        assert cb["status"] == 0
        srv.release(rb)
        assert srv.stats()["pool_used"] == 0


def test_submit_many(dev):
    """Requests submitted in one call run like ones submitted one by one
    (same outputs as their single-image launches, ids in order), and a call
    holding one invalid request queues none of them."""
    import torch
    import tucuda_amd
    from tucuda_amd.serve import Server
    reqs = _requests(dev, 10, 57, torch.bfloat16, sizes=(64, 300))
    with Server(1 << 28) as srv:
        bad = [(x, s, m) for x, s, m, _ in reqs[:3]] + [(reqs[3][0], 2.0, 0)]
        with pytest.raises(tucuda_amd.TucuError):
            srv.submit_many(bad)                              # the last one asks for no output
        assert srv.stats()["requests_launched"] == 0
        ids = srv.submit_many([(x, s, m) for x, s, m, _ in reqs])
        assert ids == sorted(ids) and len(set(ids)) == len(reqs)
        got = _wait(srv, ids)
        assert srv.stats()["max_batch"] > 1
        for rid, (x, sigma, mask, _) in zip(ids, reqs):
            c = got[rid]
            assert c["status"] == 0
            outs = _outputs(srv, c, x, mask)
            ref = _single(x, sigma, mask)
            for k in outs:
                _same(outs[k], ref[k], f"req {rid} out {k}")
            srv.release(rid)
        assert srv.stats()["pool_used"] == 0
        assert srv.submit_many([]) == []


def test_pool_grows_by_free_hbm(dev):
    """A pool that starts empty grows in chunks while the device's free
    memory (less the reserve) allows; a reserve larger than the device keeps
    it from growing, and a request that then can never fit fails with
    TUCU_ERR_BAD_ALLOC instead of waiting forever."""
    import torch
    import tucuda_amd
    from tucuda_amd.serve import Server
    reqs = _requests(dev, 6, 71, torch.bfloat16, sizes=(300, 400))
    with Server(0, grow_bytes=1 << 20) as srv:
        assert srv.stats()["pool_bytes"] == 0
        ids = [srv.submit(x, s, 0x3F) for x, s, _, _ in reqs]
        got = _wait(srv, ids)
        st = srv.stats()
        assert st["pool_bytes"] >= st["pool_peak"] > 0
        for rid, (x, sigma, mask, _) in zip(ids, reqs):
            c = got[rid]
            assert c["status"] == 0
            outs = _outputs(srv, c, x, 0x3F)
            ref = _single(x, sigma, 0x3F)
            for k in outs:
                _same(outs[k], ref[k], f"req {rid} out {k}")
            srv.release(rid)
    total = torch.cuda.get_device_properties(dev).total_memory
    with Server(0, reserve_bytes=2 * total) as srv:
        rid = srv.submit(reqs[0][0], 2.0, 0x01)
        c = _wait(srv, [rid])[rid]
        assert c["status"] == 4                               # TUCU_ERR_BAD_ALLOC
        assert srv.stats()["pool_bytes"] == 0
