"""Python front end of the continuous-batching server (C ABI tucu_server_*).

    srv = Server(pool_bytes=8 << 30)              # bf16 in / out, ceil(4 sigma) lobes
    rid = srv.submit(img, sigma=2.0, mask=0x3F)   # img: 2-D bf16 device tensor (a view is fine)
    for c in srv.poll():                          # completed requests
        srv.copy_output(c, 0, dst)                # smooth -> dst (device tensor)
        srv.release(c["id"])                      # output memory back to the pool

The input tensor must stay alive until its request completes.  Timestamps
(arrival, admitted, launched, completed) are seconds on the server clock,
``srv.now()``.
"""
from __future__ import annotations

import ctypes

import torch

from . import _capi
from ._capi import TucuCompletion, TucuRequest, TucuServerConfig, TucuServerCounters, check
from .filters import _DTYPES, image_desc


class Server:
    def __init__(self, pool_bytes: int, in_dtype=torch.bfloat16, out_dtype=torch.bfloat16,
                 rule: int = _capi.TUCU_RULE_CEIL4SIGMA, tile_rows: int = 0, max_batch_jobs: int = 1024,
                 max_inflight: int = 2, max_tiles: int = 0, max_batch_mpix: int = 0,
                 max_pool_bytes: int = 0, reserve_bytes: int = 0, grow_bytes: int = 0):
        """``pool_bytes`` of output memory are reserved at once; the pool grows
        while the device's free memory less ``reserve_bytes`` (0: 1 GiB)
        allows, up to ``max_pool_bytes`` (0: no cap; pass pool_bytes for a
        fixed pool)."""
        cfg = TucuServerConfig(int(pool_bytes), _DTYPES[in_dtype], _DTYPES[out_dtype], int(rule),
                               int(tile_rows), int(max_batch_jobs), int(max_inflight), int(max_tiles),
                               int(max_batch_mpix), 0, int(max_pool_bytes), int(reserve_bytes), int(grow_bytes))
        h = ctypes.c_void_p()
        check(_capi.load().tucu_server_create(ctypes.byref(h), ctypes.byref(cfg)))
        self._h = h
        self._buf = (TucuCompletion * 4096)()

    def now(self) -> float:
        return _capi.load().tucu_server_now(self._h)

    def submit(self, img: torch.Tensor, sigma: float, mask: int = 0x3F, shift: bool = True,
               arrival: float | None = None) -> int:
        req = TucuRequest(image_desc(img), float(sigma), int(mask), int(bool(shift)), 0,
                          self.now() if arrival is None else float(arrival))
        rid = ctypes.c_uint64(0)
        check(_capi.load().tucu_server_submit(self._h, ctypes.byref(req), ctypes.byref(rid)))
        return rid.value

    def submit_many(self, requests, arrival: float | None = None) -> list:
        """Submits ``(img, sigma, mask)`` or ``(img, sigma, mask, shift)``
        tuples in one call (all checked first; none queued if one is
        invalid); returns their ids in order."""
        n = len(requests)
        arr = (TucuRequest * max(n, 1))()
        t = self.now() if arrival is None else float(arrival)
        for i, r in enumerate(requests):
            img, sigma, mask = r[0], r[1], r[2]
            shift = r[3] if len(r) > 3 else True
            arr[i] = TucuRequest(image_desc(img), float(sigma), int(mask), int(bool(shift)), 0, t)
        ids = (ctypes.c_uint64 * max(n, 1))()
        check(_capi.load().tucu_server_submit_many(self._h, arr, n, ids))
        return list(ids[:n])

    def poll(self, max_items: int = 4096):
        n = ctypes.c_int(0)
        check(_capi.load().tucu_server_poll(self._h, self._buf, min(max_items, len(self._buf)), ctypes.byref(n)))
        out = []
        for i in range(n.value):
            c = self._buf[i]
            out.append({"id": c.id, "status": c.status, "arrival": c.arrival, "admitted": c.admitted,
                        "launched": c.launched, "completed": c.completed,
                        "out": [(c.out[k].data, c.out[k].stride, c.out[k].nrow, c.out[k].ncol, c.out[k].dtype)
                                for k in range(6)]})
        return out

    def copy_output(self, completion: dict, k: int, dst: torch.Tensor) -> None:
        """Device-to-device copy of output k of a completed request into dst
        (a 2-D tensor of the output's element type and shape)."""
        data, stride, nrow, ncol, dt = completion["out"][k]
        if not data:
            raise ValueError(f"output {k} was not requested")
        if tuple(dst.shape) != (nrow, ncol) or _DTYPES[dst.dtype] != dt:
            raise ValueError("destination shape / dtype mismatch")
        es = dst.element_size()
        torch.cuda.synchronize(dst.device)
        check(_capi.load().tucu_memcpy2d(ctypes.c_void_p(dst.data_ptr()), dst.stride(0) * es,
                                         ctypes.c_void_p(data), stride * es, ncol * es, nrow, 2))

    def release(self, rid: int) -> None:
        check(_capi.load().tucu_server_release(self._h, ctypes.c_uint64(rid)))

    def stats(self) -> dict:
        st = TucuServerCounters()
        check(_capi.load().tucu_server_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _capi.load().tucu_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
