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


# --------------------------------------------------------------------------
# failure fixtures (tests/golden/c0f_*, c1f_*: zones overwritten so that step
# one runs every branch of the recovery tree and both interpolating fixups)

FAILURE_FIXTURES = {"c0f": "c0f_balsara1024_failures.npz", "c1f": "c1f_torus24x24x16_failures.npz"}
NFAIL = {n: l for l, n in enumerate(["FLOOR", "TMAX", "GAMMA_MAX", "GAMMA_CALC", "USE_FULL_INV", "USE_EE",
                                     "UTOPRIM", "UTOPRIM_EE", "USE_INTERP_PRIM", "USE_INTERP_V",
                                     "INTERP_PRIM", "INTERP_V"])}
BRANCH = {n: s for s, n in enumerate(["Utoprim_2d_fast", "Utoprim_1d", "Utoprim_1d_ee", "Utoprim_1d_ee2",
                                      "fixup_interp_v", "fixup_interp_prim"])}


def failure_grid(which):
    """(GridSpec, keyword arguments of the run, fixture dict) of a failure fixture."""
    from coolharm3d_amd import balsara1_grid, torus_grid
    G = golden(FAILURE_FIXTURES[which])
    if which == "c0f":
        return balsara1_grid(1024, cour=float(G["cour"])), {}, G
    n = tuple(int(x) for x in G["n"])
    return torus_grid(*n), {"r_isco": float(G["r_isco"])}, G


def zone_errors(ref, got, rtol, atol_frac):
    """Zones where |got - ref| > rtol |ref| + atol_frac max|ref_v| (per
    variable v; no absolute part for the density and internal energy, which
    are positive): a per-zone check, so a wrong answer in a single
    low-density zone cannot hide inside a grid-wide L1 sum."""
    ref = np.asarray(ref, dtype=np.float64).reshape(-1, 8)
    got = np.asarray(got, dtype=np.float64).reshape(-1, 8)
    scale = np.abs(ref).max(axis=0)
    scale[:2] = 0.0
    bad = np.abs(got - ref) > rtol * np.abs(ref) + atol_frac * scale
    return np.argwhere(bad)


def check_failure_run(H, G, steps, advance, n1t, l1_tol=1e-12, zone_rtol=1e-9, exact=False):
    """Advances H (a Harm3D or OracleRun) through the recorded steps of a
    failure fixture and compares state, pflag and nfail at each one: state to
    relative L1 <= l1_tol per variable and per zone to zone_rtol (bit for bit
    if exact); pflag bit for bit; nfail identical up to the reference's
    first-call misattribution."""
    done = 0
    for s in steps:
        advance(s - done)
        done = s
        got = interior(H.prims())
        ref = G[f"s{s}_prims"]
        assert np.isfinite(got).all(), f"non-finite state after {s} steps"
        if exact:
            assert np.array_equal(got, ref), f"state after {s} steps not bit-identical"
        assert_rel_l1(ref, got, l1_tol, f"after {s} steps")
        bad = zone_errors(ref, got, zone_rtol, 1e-13)
        assert len(bad) == 0, f"after {s} steps: {len(bad)} zone values off, first {bad[:4].tolist()}"
        assert np.array_equal(interior(H.pflag()), G[f"s{s}_pflag"].astype(np.int32)), f"pflag after {s}"
        ok, info = nfail_equal_modulo_cooling_init(G[f"s{s}_nfail"], H.nfail(), n1t)
        assert ok, f"nfail after {s} steps: {info}"
