"""Kernel bodies (host build, libhgpu_hostcheck.so) against the reference's
golden outputs.  These run the exact per-thread code of the GPU kernels in
host loops, so the arithmetic of every kernel is pinned on a machine without
a GPU; the GPU tests repeat the comparison on the device.

Tolerances: per-variable relative L1 <= 1e-12 after 1 and 10 steps, <= 1e-9
after 100 steps; pflag bit-identical; nfail identical up to the reference's
first-call misattribution (helpers.nfail_equal_modulo_cooling_init).
"""
import os

import numpy as np
import pytest

from helpers import golden, interior, assert_rel_l1, nfail_equal_modulo_cooling_init, ROOT

HOSTLIB = os.path.join(ROOT, "coolharm3d_amd", "libhgpu_hostcheck.so")
pytestmark = pytest.mark.skipif(not os.path.exists(HOSTLIB), reason="host check build missing")

TOL = {1: 1e-12, 10: 1e-12, 100: 1e-9}


def _run(H, G, steps, n1t, check_nfail=True, done=0):
    for s in steps:
        H.step_ch(s - done)
        done = s
        got = interior(H.prims())
        assert np.isfinite(got).all()
        assert_rel_l1(G[f"s{s}_prims"], got, TOL[s], f"after {s} steps")
        t, dt, _ = H.time()
        assert t == pytest.approx(float(G[f"s{s}_t"]), rel=1e-14)
        assert dt == pytest.approx(float(G[f"s{s}_dt"]), rel=1e-10)
        assert np.array_equal(interior(H.pflag()), G[f"s{s}_pflag"].astype(np.int32)), f"pflag after {s}"
        if check_nfail:
            ok, info = nfail_equal_modulo_cooling_init(G[f"s{s}_nfail"], H.nfail(), n1t)
            assert ok, f"nfail after {s} steps: {info}"


def test_balsara1_matches_reference_bit_for_bit():
    from coolharm3d_amd import Harm3D, balsara1_grid
    G = golden("c0_balsara1024.npz")
    H = Harm3D(balsara1_grid(1024, cour=float(G["cour"])), hostcheck=True)
    H.set_prims(G["p0"])
    H.set_time(0.0, float(G["dt0"]))
    H.step_ch(1)
    assert np.array_equal(interior(H.prims()), G["s1_prims"]), "step 1 not bit-identical"
    _run(H, G, (10, 100), 1030, done=1)
    # the whole 100-step state is bit-identical for this configuration
    assert np.array_equal(interior(H.prims()), G["s100_prims"])


def test_balsara1_initial_data_and_bounds_match_reference():
    """problems.balsara1 + hgpu_fixup_bounds reproduces the reference's
    init() (src/init.balsara1.c:93-135: fixup(p); bounds(p,0)) bit for bit."""
    from coolharm3d_amd import Harm3D, balsara1_grid, balsara1
    G = golden("c0_balsara1024.npz")
    g = balsara1_grid(1024)
    H = Harm3D(g, hostcheck=True)
    H.set_prims(balsara1(g))
    H.fixup_bounds()
    assert np.array_equal(H.prims(), G["p0"])


@pytest.mark.slow
def test_torus_matches_reference():
    from coolharm3d_amd import Harm3D, torus_grid
    G = golden("c1_torus24x24x16.npz")
    n = tuple(int(x) for x in G["n"])
    H = Harm3D(torus_grid(*n), hostcheck=True, r_isco=float(G["r_isco"]))
    H.set_prims(G["p0"])
    H.set_time(0.0, float(G["dt0"]))
    _run(H, G, (1, 10, 100), n[0] + 6)


@pytest.mark.slow
def test_torus_diag3_matches_reference():
    """The torus on the COORD_DIAGONAL3 theta map (the reference's binary-disk
    set-up's, src/coord.c:270-276) against the reference build of it."""
    from coolharm3d_amd import Harm3D, torus_grid
    G = golden("c1d3_torus24x24x16.npz")
    n = tuple(int(x) for x in G["n"])
    H = Harm3D(torus_grid(*n, coords="diag3"), hostcheck=True, r_isco=float(G["r_isco"]))
    H.set_prims(G["p0"])
    H.set_time(0.0, float(G["dt0"]))
    _run(H, G, (1, 10, 100), n[0] + 6)


def test_torus_initial_fixup_matches_reference():
    """hgpu_fixup_bounds on the generated torus equals the reference's
    fixup(p); bounds(p,0) (src/fixup.c:8, src/bounds.c:492-579)."""
    from coolharm3d_amd import Harm3D, torus_grid, torus
    G = golden("c1_torus24x24x16.npz")
    g = torus_grid(*(int(x) for x in G["n"]))
    H = Harm3D(g, hostcheck=True)
    H.set_prims(torus(g))
    H.fixup_bounds()
    assert np.array_equal(H.prims(), G["p0"])
