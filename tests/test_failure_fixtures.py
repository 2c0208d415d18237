"""The failure fixtures reach the recovery tree's branches, and the oracle
and the host build of the kernel bodies reproduce the reference on them.

tests/golden/c0f_balsara1024_failures.npz and c1f_torus24x24x16_failures.npz
are the shock tube and the torus with a few zones of the initial state
overwritten (tests/golden/make_golden.py).  The reference build counted, on
step one, the calls and successful returns of each solver of
recover_primitives() (src/step_ch.c:801-1024) and fixup_entropy_eq()
(src/fixup.c:313-454) and the calls of the interpolating fixups
(src/fixup.c:499-761); those counters and the recorded flag maps show which
branches the fixture exercises, so a match on the fixture pins them.
"""
import os

import numpy as np
import pytest

from helpers import (golden, interior, FAILURE_FIXTURES, NFAIL, BRANCH, failure_grid, check_failure_run,
                     ROOT)

HOSTLIB = os.path.join(ROOT, "coolharm3d_amd", "libhgpu_hostcheck.so")


def _nfail_step1(G):
    return G["s1_nfail"].reshape(15, -1).sum(axis=1)


@pytest.mark.parametrize("which", sorted(FAILURE_FIXTURES))
def test_fixture_reaches_every_branch_on_step_one(which):
    G = golden(FAILURE_FIXTURES[which])
    b = G["s1_branch"]
    nf = _nfail_step1(G)
    pf = G["s1_pflag"]
    # the flag map holds both kinds of flagged zones
    assert (pf == 1).any() and (pf == 2).any()
    # the 2D solver failed (or returned an unphysical Lorentz factor) somewhere:
    # the 1D fallback ran, and succeeded at least once
    assert b[BRANCH["Utoprim_1d"]][0] > 0 and b[BRANCH["Utoprim_1d"]][1] > 0
    assert b[BRANCH["Utoprim_1d"]][1] < b[BRANCH["Utoprim_1d"]][0]
    # the entropy fix ran both solvers, each with successes and failures
    for s in ("Utoprim_1d_ee", "Utoprim_1d_ee2"):
        assert 0 < b[BRANCH[s]][1] < b[BRANCH[s]][0], s
    # both interpolating fixups ran
    assert b[BRANCH["fixup_interp_v"]][0] > 0 and b[BRANCH["fixup_interp_prim"]][0] > 0
    # the counters of every branch of the tree are non-zero
    for name in ("FLOOR", "GAMMA_MAX", "USE_FULL_INV", "UTOPRIM", "UTOPRIM_EE", "USE_INTERP_V",
                 "INTERP_PRIM", "INTERP_V"):
        assert nf[NFAIL[name]] > 0, name
    codes = {int(c): int(n) for c, n in G["s1_codes"][BRANCH["Utoprim_2d_fast"]] if n}
    assert codes.get(-100, 0) > 0, "no conserved density <= 0 reached the 2D solver"


def test_torus_fixture_also_interpolates_primitives_and_uses_the_entropy_solution():
    nf = _nfail_step1(golden(FAILURE_FIXTURES["c1f"]))
    assert nf[NFAIL["USE_INTERP_PRIM"]] > 0 and nf[NFAIL["USE_EE"]] > 0


@pytest.mark.parametrize("which", sorted(FAILURE_FIXTURES))
def test_oracle_matches_reference_on_failures(which):
    from pyoracle import OracleRun
    grid, kw, G = failure_grid(which)
    O = OracleRun(grid, **kw)
    O.set_prims(G["p0"])
    O.set_time(0.0, float(G["dt0"]))
    check_failure_run(O, G, (1, 2, 10), O.step, grid.n1 + 6, exact=(which == "c0f"))


@pytest.mark.skipif(not os.path.exists(HOSTLIB), reason="host check build missing")
@pytest.mark.parametrize("which", sorted(FAILURE_FIXTURES))
def test_host_build_matches_reference_on_failures(which):
    from coolharm3d_amd import Harm3D
    grid, kw, G = failure_grid(which)
    H = Harm3D(grid, hostcheck=True, **kw)
    H.set_prims(G["p0"])
    H.set_time(0.0, float(G["dt0"]))
    check_failure_run(H, G, (1, 2, 10), H.step_ch, grid.n1 + 6, exact=(which == "c0f"))


def test_zone_check_catches_what_the_l1_sum_hides():
    """A 1e-6 relative error in one low-density zone passes the grid-wide
    relative L1 bar of 1e-12 but not the per-zone check."""
    from helpers import rel_l1, zone_errors
    G = golden(FAILURE_FIXTURES["c1f"])
    ref = G["s1_prims"]
    got = ref.copy()
    i, j, k = np.unravel_index(np.argmin(ref[..., 0]), ref.shape[:3])
    got[i, j, k, 1] *= 1 + 1e-6
    assert rel_l1(ref, got).max() <= 1e-12
    bad = zone_errors(ref, got, 1e-9, 1e-13)
    assert bad.tolist() == [[np.ravel_multi_index((i, j, k), ref.shape[:3]), 1]]


def test_oracle_balsara_to_t04_bit_identical():
    from pyoracle import OracleRun
    from coolharm3d_amd import balsara1_grid
    G = golden("c0_balsara1024_t04.npz")
    O = OracleRun(balsara1_grid(1024, cour=float(G["cour"])))
    O.set_prims(G["p0"])
    O.set_time(0.0, float(G["dt0"]))
    n = 0
    while O.time()["t"] < float(G["t_end"]):
        O.step(1)
        n += 1
    assert n == int(G["nstep"])
    assert np.array_equal(interior(O.prims()), G["end_prims"])
    assert np.array_equal(O.nfail(), G["end_nfail"])
    assert O.time()["t"] == float(G["end_t"])


@pytest.mark.skipif(not os.path.exists(HOSTLIB), reason="host check build missing")
def test_host_build_balsara_to_t04_bit_identical():
    from coolharm3d_amd import Harm3D, balsara1_grid
    G = golden("c0_balsara1024_t04.npz")
    H = Harm3D(balsara1_grid(1024, cour=float(G["cour"])), hostcheck=True)
    H.set_prims(G["p0"])
    H.set_time(0.0, float(G["dt0"]))
    n = 0
    while H.time()[0] < float(G["t_end"]):
        H.step_ch(1)
        n += 1
    assert n == int(G["nstep"])
    assert np.array_equal(interior(H.prims()), G["end_prims"])
    assert np.array_equal(H.nfail(), G["end_nfail"])
    assert H.time()[0] == float(G["end_t"])
