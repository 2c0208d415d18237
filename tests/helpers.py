"""Shared comparison helpers of the parity tests."""
from __future__ import annotations

import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
NG = 3
VARS = ["rho", "u", "u1", "u2", "u3", "B1", "B2", "B3"]


def golden(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


def interior(a):
    return a[NG:-NG, NG:-NG, NG:-NG]


def rel_l1(ref, got):
    """Per-variable relative L1 error sum|got-ref| / sum|ref| over zones."""
    ref = np.asarray(ref, dtype=np.float64).reshape(-1, 8)
    got = np.asarray(got, dtype=np.float64).reshape(-1, 8)
    num = np.abs(got - ref).sum(axis=0)
    den = np.abs(ref).sum(axis=0)
    out = np.where(den > 0, num / np.where(den > 0, den, 1.0), num)
    return out


def assert_rel_l1(ref, got, tol, what=""):
    e = rel_l1(ref, got)
    bad = [f"{VARS[v]}={e[v]:.3e}" for v in range(8) if not (e[v] <= tol)]
    assert not bad, f"{what}: relative L1 above {tol:g}: " + ", ".join(bad)


def nfail_equal_modulo_cooling_init(ref_nfail, got_nfail, n1t):
    """(ok, info): nfail maps [15][N1T][N2T][N3T] equal, allowing for the
    reference's one-time attribution of one zone's fail() counts to zone
    (N1TOT-1, N2S, N3S): the first call of cooling_func_hr_disk() runs
    get_coord() over the x1 line (src/phys.c:697-698), and get_coord() sets
    icurr/jcurr/kcurr (src/macros.h:613), which fail() uses as the zone index
    (src/diag.c:172).  The counts that zone received must appear, in the GPU
    map, on exactly one interior zone."""
    r = np.array(ref_nfail, dtype=np.int64, copy=True)
    g = np.array(got_nfail, dtype=np.int64)
    if np.array_equal(r, g):
        return True, "identical"
    ghost = (n1t - 1, NG, NG)
    for l in range(r.shape[0]):
        moved = r[(l,) + ghost] - g[(l,) + ghost]
        if moved:
            r[(l,) + ghost] -= moved
            excess = np.argwhere((g[l] - r[l]) == moved)
            if len(excess) != 1:
                return False, f"counter {l}: {moved} counts at {ghost} not matched by one zone"
            r[(l,) + tuple(excess[0])] += moved
    ok = np.array_equal(r, g)
    return ok, ("identical after moving the first cooled zone's counts" if ok
                else f"differ at {np.argwhere(r != g)[:5].tolist()}")
