"""Write and copy rates of the memory system with the filters' store form.

    python -m tucuda_amd.membench

A grid-stride fill with 16-byte stores (non-temporal, the form the filter
kernels use, and ordinary) at several grid sizes, and a copy, from
tucuda_amd/csrc/membench.hip: the reference point for what a write stream
of a given size reaches on this GPU.  Prints one line per point and one
JSON line.
"""
from __future__ import annotations

import ctypes
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    path = os.path.join(HERE, "libtucuda_membench.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (make -C tucuda_amd/csrc)")
    L = ctypes.CDLL(path)
    L.tucu_membench_fill.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_void_p]
    L.tucu_membench_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    return L


def main(argv=None):
    import torch
    dev = torch.device("cuda:0")
    L = load()
    st = torch.cuda.current_stream(dev)
    sh = ctypes.c_void_p(st.cuda_stream)

    def ev(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            fn()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) * 1e-3 / reps

    res = {}
    for size in (48 << 20, 1 << 30, 4 << 30):
        nbuf = 10 if size < (1 << 29) else 2
        reps = 20 if size < (1 << 29) else 5
        bufs = [torch.empty(size // 4, device=dev) for _ in range(nbuf)]
        for nt in (1, 0):
            for pc in (2, 4, 8, 16):
                k = [0]

                def fill():
                    b = bufs[k[0] % nbuf]
                    k[0] += 1
                    assert L.tucu_membench_fill(ctypes.c_void_p(b.data_ptr()), size, 1.0, nt, pc, sh) == 0

                gbs = size / ev(fill, reps) / 1e9
                res[f"fill_{size >> 20}MiB_nt{nt}_wg{pc}"] = round(gbs, 1)
                print(f"fill {size >> 20} MiB nt={nt} wg/cu={pc}: {gbs:.1f} GB/s", file=sys.stderr, flush=True)
        if nbuf >= 2:
            a, b = bufs[0], bufs[1]

            def copy():
                assert L.tucu_membench_copy(ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(a.data_ptr()),
                                            size, sh) == 0

            gbs = 2 * size / ev(copy, reps) / 1e9
            res[f"copy_{size >> 20}MiB"] = round(gbs, 1)
            print(f"copy {size >> 20} MiB: {gbs:.1f} GB/s (read + written)", file=sys.stderr, flush=True)
        del bufs
        torch.cuda.empty_cache()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    sys.exit(main())
