"""stk_stats_noalloc, the statistics entry point that allocates nothing, against stk_stats.

The test plays the caller: it allocates the device inputs and the output itself, beforehand, with
the HIP runtime the library is linked against (loaded by path, so it is the same runtime instance),
and the library's call enqueues on the context stream, so the test waits for the device before
reading the output.  Two calls in a row on two meshes, the second smaller, then the first again:
the fixed scratch array keeps the previous call's contents, which must not matter.
"""
import ctypes as C

import numpy as np
import pytest

from dimspers_amd import _lib
from dimspers_amd.inputs import make_inputs, scaled_workload

pytestmark = pytest.mark.gpu

HIP_H2D, HIP_D2H = 1, 2


class _Hip:
    def __init__(self):
        _lib.load()                                   # the library first: its runtime is the one loaded
        self.h = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
            getattr(self.h, name).restype = C.c_int
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), a.nbytes) == 0 and p.value
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, HIP_H2D) == 0
        return p

    def download(self, p, n):
        out = np.empty(n)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, HIP_D2H) == 0
        return out


def _case(nr, nt, seed):
    inp = make_inputs(scaled_workload(3, nr, nt))
    ctx = _lib.Context.from_mesh(inp.mesh)
    rng = np.random.default_rng(seed)
    U = rng.standard_normal(inp.mesh.nU)
    return ctx, U, np.asarray(inp.T, dtype=np.float64)


def test_stats_without_allocation_match_stk_stats():
    hip = _Hip()
    cases = [_case(32, 256, 1), _case(8, 64, 2)]
    bufs = []
    for ctx, U, T in cases + cases[:1]:
        print(f"[case] stk_stats_noalloc nU {ctx.nU}", flush=True)
        want = ctx.stats(T, U)
        assert U.size == ctx.nU and T.size == ctx.nU // 2
        Ud, Td = hip.upload(U), hip.upload(T)
        out = hip.upload(np.full(4, np.nan))
        bufs += [Ud, Td, out]
        assert hip.h.hipDeviceSynchronize() == 0
        rc = ctx.lib.stk_stats_noalloc(ctx.h, Ud, Td, out)
        assert rc == _lib.STK_OK, ctx.lib.stk_last_error(ctx.h)
        got = hip.download(out, 4)
        print(f"[result] noalloc {got} stk_stats {np.asarray(want)}", flush=True)
        assert np.allclose(got, want, rtol=1e-14, atol=0)
    for ctx, _, _ in cases:
        # NULL device pointers are refused
        assert ctx.lib.stk_stats_noalloc(ctx.h, None, None, bufs[2]) == _lib.STK_E_ARG
        ctx.close()
    for p in bufs:
        assert hip.h.hipFree(p) == 0
