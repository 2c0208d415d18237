"""hgpu_c2p_solve_static(): the compacted con-to-prim solve whose queues live in
the library's static device array, called through the C ABI and compared with
hgpu_c2p_solve() (caller workspace) on the same zones: primitives, return
codes and Newton iterations bit for bit, on two calls in a row, the second with
a smaller input (the array keeps the first call's contents), with the plain
and the learned guess.  Sizes beyond the array's capacity are refused."""
import ctypes

import numpy as np
import pytest

from coolharm3d_amd import _lib
from coolharm3d_amd.c2p import Con2PrimBatch

pytestmark = pytest.mark.gpu

A, GAM = 0.9, 5.0 / 3.0
STATIC_MAX = 1 << 18


def _static_fn(lib):
    fn = lib.hgpu_c2p_solve_static
    vp, ll, d, i = ctypes.c_void_p, ctypes.c_longlong, ctypes.c_double, ctypes.c_int
    fn.restype = i
    fn.argtypes = [vp, ll, d, d, vp, vp, vp, i, i, i, vp, vp, vp]
    return fn


def _solve_static(B, fn, d, mlp, budget_first=0, budget_next=0):
    n = d["U"].shape[1]
    out = (B.empty((8, n), np.float64), B.empty((n,), np.int32), B.empty((n,), np.int32))
    rc = fn(None, n, A, GAM, d["U"].ptr, d["rth"].ptr, B.ptr(mlp), 1, budget_first, budget_next,
            out[0].ptr, out[1].ptr, out[2].ptr)
    _lib.check(B.lib, rc, "hgpu_c2p_solve_static")
    return [o.numpy() for o in out]


def _solve_work(B, d, mlp, budget_first=0, budget_next=0):
    out = B.solve(d["U"], d["rth"], A, GAM, mlp=mlp, plain_guess="cold", compact=True,
                  budget_first=budget_first, budget_next=budget_next)
    return [o.numpy() for o in out]


def _assert_same(got, ref):
    for g, r, what in zip(got, ref, ("prim", "code", "iters")):
        assert g.dtype == r.dtype and g.shape == r.shape, what
        assert np.array_equal(g.view(np.uint8), r.view(np.uint8)), f"{what} differs"


def test_static_scratch_solve_equals_workspace_solve():
    B = Con2PrimBatch()
    fn = _static_fn(B.lib)
    w, _ = B.train_guess(1 << 14, 99, A, GAM, steps=200, batch=1024, lr=3e-3)
    for mlp in (None, w):
        # full capacity first, then a smaller call over the array's stale contents
        for n, seed, budgets in ((STATIC_MAX, 11, (0, 0)), (50_000, 12, (4, 3))):
            print(f"static scratch solve: n={n} seed={seed} learned={mlp is not None} budgets={budgets}", flush=True)
            d = B.samples(n, seed, A, GAM)
            got = _solve_static(B, fn, d, mlp, *budgets)
            ref = _solve_work(B, d, mlp, *budgets)
            _assert_same(got, ref)
            assert (got[2] > 0).all()


def test_static_scratch_refuses_oversize():
    B = Con2PrimBatch()
    fn = _static_fn(B.lib)
    d = B.samples(8, 1, A, GAM)
    code, iters = B.empty((8,), np.int32), B.empty((8,), np.int32)
    # the size check comes before any use of the (too small) arrays
    rc = fn(None, STATIC_MAX + 1, A, GAM, d["U"].ptr, d["rth"].ptr, None, 1, 0, 0, d["prim"].ptr, code.ptr,
            iters.ptr)
    assert rc == -3
    assert b"static scratch capacity" in B.lib.hgpu_last_error()
