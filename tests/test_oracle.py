"""Pins the CPU restatement (oracle/harm_oracle.c) before it is trusted as the
checker of the GPU library.

Each part of the oracle is checked against something other than its own
formula:
  * the reference's outputs (tests/golden/, written by the reference build):
    Balsara test 1 bit for bit for 100 steps (state, pflag, nfail, t, dt);
    the Kerr torus to 1e-12 (1, 10 steps) / 1e-9 (100 steps) with identical
    pflag maps;
  * closed forms: special-relativistic MHD conserved variables and the
    at-rest characteristic speeds in flat space; exact face values of linear
    data for MC and PPM; no new extrema at a discontinuity; HLL/LF
    consistency F(U,U) = F(U);
  * invariants: g_cov g^con = 1 and sqrt(-det g) for the metric records; the
    connection's symmetry, its agreement with finite differences of a
    textbook Kerr-Schild metric, and the Killing identities
    T^mu_nu Gamma^nu_{t mu} = T^mu_nu Gamma^nu_{phi mu} = 0; con-to-prim round
    trips; div B at round-off under constrained transport; the boundary
    symmetries; the CFL time step formula;
  * the reference build itself (oracle/_ref, when built in this container) on
    a grid and seed that no fixture covers.
"""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import golden, interior, assert_rel_l1, nfail_equal_modulo_cooling_init, rel_l1, ROOT

import pyoracle
from pyoracle import OracleRun

FACE1, FACE2, FACE3, CENT = 0, 1, 2, 3


def _geom_struct(g42):
    gcon = g42[0:16].reshape(4, 4)
    gcov = g42[16:32].reshape(4, 4)
    return dict(gcon=gcon, gcov=gcov, g=g42[32], g_inv=g42[33], alpha=g42[34], beta=g42[35:38], ncon=g42[38:42])


# --------------------------------------------------------------------------
# reference outputs

def test_oracle_balsara_bit_identical_to_reference():
    from coolharm3d_amd.problems import balsara1_grid
    G = golden("c0_balsara1024.npz")
    O = OracleRun(balsara1_grid(1024, cour=float(G["cour"])))
    O.set_prims(G["p0"])
    O.set_time(0.0, float(G["dt0"]))
    done = 0
    for s in (1, 10, 100):
        O.step(s - done)
        done = s
        assert np.array_equal(interior(O.prims()), G[f"s{s}_prims"]), f"state after {s} steps"
        assert np.array_equal(interior(O.pflag()), G[f"s{s}_pflag"].astype(np.int32))
        assert np.array_equal(O.nfail(), G[f"s{s}_nfail"])
        tm = O.time()
        assert tm["t"] == float(G[f"s{s}_t"]) and tm["dt"] == float(G[f"s{s}_dt"])


@pytest.mark.slow
def test_oracle_torus_matches_reference():
    from coolharm3d_amd.problems import torus_grid
    G = golden("c1_torus24x24x16.npz")
    n = tuple(int(x) for x in G["n"])
    O = OracleRun(torus_grid(*n), r_isco=float(G["r_isco"]))
    O.set_prims(G["p0"])
    O.set_time(0.0, float(G["dt0"]))
    done = 0
    for s, tol in ((1, 1e-12), (10, 1e-12), (100, 1e-9)):
        O.step(s - done)
        done = s
        assert_rel_l1(G[f"s{s}_prims"], interior(O.prims()), tol, f"torus after {s} steps")
        assert np.array_equal(interior(O.pflag()), G[f"s{s}_pflag"].astype(np.int32)), f"pflag after {s}"
        ok, info = nfail_equal_modulo_cooling_init(G[f"s{s}_nfail"], O.nfail(), n[0] + 6)
        assert ok, info
        assert O.time()["dt"] == pytest.approx(float(G[f"s{s}_dt"]), rel=1e-10)


@pytest.mark.slow
def test_oracle_torus_diag3_matches_reference():
    """The torus on the reference's binary-disk theta map (COORD_DIAGONAL3,
    src/coord.c:270-276) against the reference build of that configuration."""
    from coolharm3d_amd.problems import torus_grid
    G = golden("c1d3_torus24x24x16.npz")
    n = tuple(int(x) for x in G["n"])
    O = OracleRun(torus_grid(*n, coords="diag3"), r_isco=float(G["r_isco"]))
    O.set_prims(G["p0"])
    O.set_time(0.0, float(G["dt0"]))
    done = 0
    for s, tol in ((1, 1e-12), (10, 1e-12), (100, 1e-9)):
        O.step(s - done)
        done = s
        assert_rel_l1(G[f"s{s}_prims"], interior(O.prims()), tol, f"diag3 torus after {s} steps")
        assert np.array_equal(interior(O.pflag()), G[f"s{s}_pflag"].astype(np.int32)), f"pflag after {s}"
        ok, info = nfail_equal_modulo_cooling_init(G[f"s{s}_nfail"], O.nfail(), n[0] + 6)
        assert ok, info


def test_oracle_torus_first_step_flags():
    from coolharm3d_amd.problems import torus_grid
    G = golden("c1_torus24x24x16.npz")
    n = tuple(int(x) for x in G["n"])
    O = OracleRun(torus_grid(*n), r_isco=float(G["r_isco"]))
    O.set_prims(G["p0"])
    O.set_time(0.0, float(G["dt0"]))
    O.step(1)
    assert_rel_l1(G["s1_prims"], interior(O.prims()), 1e-12, "torus step 1")
    assert np.array_equal(interior(O.pflag()), G["s1_pflag"].astype(np.int32))


# --------------------------------------------------------------------------
# flat-space closed forms

def _flat_geom():
    from coolharm3d_amd.problems import balsara1_grid
    O = OracleRun(balsara1_grid(16))
    return O.geom(3, 3, CENT)


def _srmhd_conserved(p, gam):
    """Special-relativistic MHD: D, tau + D, S_i, B^i in the reference's
    convention U = (rho u^t, T^t_t + rho u^t, T^t_i, B^i) (flat, sqrt(-g) = 1)."""
    rho, u, ut_ = p[0], p[1], np.array(p[2:5])
    B = np.array(p[5:8])
    gamma = math.sqrt(1.0 + ut_ @ ut_)
    ucon = np.concatenate([[gamma], ut_])                  # u^i = u~^i in flat space
    eta = np.diag([-1.0, 1, 1, 1])
    ucov = eta @ ucon
    bt = B @ ut_
    bcon = np.concatenate([[bt], (B + bt * ut_) / gamma])
    bcov = eta @ bcon
    bsq = bcon @ bcov
    pg = (gam - 1) * u
    w = rho + u + pg + bsq
    T_t = w * ucon[0] * ucov - bcon[0] * bcov
    T_t[0] += pg + 0.5 * bsq
    return np.array([rho * gamma, T_t[0] + rho * gamma, T_t[1], T_t[2], T_t[3], B[0], B[1], B[2]])


def test_oracle_conserved_variables_flat_closed_form():
    g = _flat_geom()
    rng = np.random.default_rng(11)
    for _ in range(200):
        p = np.concatenate([rng.uniform(0.1, 2, 2), rng.normal(0, 2, 3), rng.normal(0, 1, 3)])
        U = pyoracle.prim_to_flux(g, 5.0 / 3.0, p, 0)
        ref = _srmhd_conserved(p, 5.0 / 3.0)
        assert np.allclose(U, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_oracle_wave_speeds_at_rest():
    g = _flat_geom()
    gam = 4.0 / 3.0
    for B in ((0, 0, 0), (0.3, 0.0, 0.0), (0.1, 0.7, -0.2)):
        p = np.array([1.3, 0.7, 0, 0, 0, *B], float)
        bsq = float(np.dot(B, B))
        cms2 = (gam * (gam - 1) * p[1] + bsq) / (p[0] + p[1] + (gam - 1) * p[1] + bsq)
        for d in (1, 2, 3):
            cmax, cmin = pyoracle.vchar(g, gam, p, d)
            assert cmax == pytest.approx(math.sqrt(cms2), rel=1e-14)
            assert cmin == pytest.approx(-math.sqrt(cms2), rel=1e-14)


@pytest.mark.parametrize("flux", [0, 1])
def test_oracle_flux_consistency(flux):
    """F(U, U) = F(U): HLL up to its SMALL regulariser, Lax-Friedrichs exactly."""
    g = _flat_geom()
    rng = np.random.default_rng(3 + flux)
    for _ in range(50):
        p = np.concatenate([rng.uniform(0.1, 2, 2), rng.normal(0, 1, 3), rng.normal(0, 1, 3)])
        for d in (1, 2, 3):
            F, _ = pyoracle.face_flux(g, 5.0 / 3.0, flux, d, p, p)
            ref = pyoracle.prim_to_flux(g, 5.0 / 3.0, p, d)
            assert np.allclose(F, ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("recon", [2, 5])
def test_oracle_reconstruction_exact_on_linear_data(recon):
    y = 0.3 + 1.7 * np.arange(20)
    left, right = pyoracle.recon_line(recon, y)
    faces = 0.3 + 1.7 * (np.arange(20) - 0.5)
    sl = slice(3, 18)
    assert np.allclose(left[sl], faces[sl], rtol=0, atol=1e-12)
    assert np.allclose(right[sl], faces[sl], rtol=0, atol=1e-12)


@pytest.mark.parametrize("recon", [2, 5])
def test_oracle_reconstruction_monotone_at_a_jump(recon):
    y = np.where(np.arange(24) < 12, 1.0, 0.125) + 0.01 * np.sin(np.arange(24))
    left, right = pyoracle.recon_line(recon, y)
    for f in range(3, 21):
        lo, hi = min(y[f - 1], y[f]), max(y[f - 1], y[f])
        assert lo - 1e-15 <= left[f] <= hi + 1e-15 or abs(left[f] - y[f - 1]) < 1e-15
        assert lo - 1e-15 <= right[f] <= hi + 1e-15 or abs(right[f] - y[f]) < 1e-15


# --------------------------------------------------------------------------
# Kerr-Schild geometry

@pytest.fixture(scope="module")
def torus_small():
    from coolharm3d_amd.problems import torus_grid
    g = torus_grid(24, 24, 16)
    return g, OracleRun(g)


def test_oracle_metric_records(torus_small):
    g, O = torus_small
    for (i, j) in ((0, 0), (3, 3), (10, 7), (29, 29), (17, 20)):
        for pos in (FACE1, FACE2, FACE3, CENT, 4):
            G = _geom_struct(O.geom(i, j, pos))
            assert np.allclose(G["gcov"] @ G["gcon"], np.eye(4), atol=1e-12)
            sdet = math.sqrt(-np.linalg.det(G["gcov"]))
            assert G["g"] == pytest.approx(sdet, rel=1e-11)
            assert G["alpha"] == pytest.approx(1.0 / math.sqrt(-G["gcon"][0, 0]), rel=1e-15)


def _ks_mks_gcov(grid, x1, x2):
    """Textbook Kerr-Schild metric (M = 1) in the grid's numerical coordinates
    r = R0 + e^x1 and th(x2): th_length x2 + th_beg + h sin(2 pi x2) (modified
    Kerr-Schild) or pi/2 (1 + h f + c f^n), f = 2 x2 - 1 (COORD_DIAGONAL3),
    written out here from those formulas."""
    a = grid.a
    r = grid.R0 + math.exp(x1)
    dr = math.exp(x1)
    if grid.coords == "diag3":
        n, c, f = grid.diag3_exponent, grid.diag3_factor, 2 * x2 - 1
        th = 0.5 * math.pi * (1 + grid.h_slope * f + c * f ** n)
        dth = math.pi * (grid.h_slope + n * c * f ** (n - 1))
    else:
        th = grid.th_length * x2 + grid.th_beg + grid.h_slope * math.sin(2 * math.pi * x2)
        dth = grid.th_length + 2 * math.pi * grid.h_slope * math.cos(2 * math.pi * x2)
    s, c = math.sin(th), math.cos(th)
    sig = r * r + a * a * c * c
    z = 2 * r / sig
    g = np.zeros((4, 4))
    g[0, 0] = -(1 - z)
    g[0, 1] = g[1, 0] = z
    g[0, 3] = g[3, 0] = -z * a * s * s
    g[1, 1] = 1 + z
    g[1, 3] = g[3, 1] = -a * (1 + z) * s * s
    g[2, 2] = sig
    g[3, 3] = s * s * (r * r + a * a + z * a * a * s * s)
    J = np.diag([1.0, dr, dth, 1.0])
    return J @ g @ J


@pytest.fixture(scope="module")
def torus_small_diag3():
    from coolharm3d_amd.problems import torus_grid
    g = torus_grid(24, 24, 16, coords="diag3")
    return g, OracleRun(g)


@pytest.mark.parametrize("coords", ["mks", "diag3"])
def test_oracle_connection(coords, torus_small, torus_small_diag3):
    g, O = torus_small if coords == "mks" else torus_small_diag3
    rng = np.random.default_rng(5)
    for (i, j) in ((3, 3), (10, 12), (26, 26), (15, 4)):
        G = O.conn(i, j)
        # symmetric in the lower indices
        assert np.allclose(G, np.transpose(G, (0, 2, 1)), rtol=0, atol=1e-13 * np.abs(G).max())
        # finite differences of the textbook metric
        x1 = g.startx[1] + (i + 0.5) * g.dx[1]
        x2 = g.startx[2] + (j + 0.5) * g.dx[2]
        h = 1e-5
        dg = np.zeros((4, 4, 4))
        dg[1] = (_ks_mks_gcov(g, x1 + h, x2) - _ks_mks_gcov(g, x1 - h, x2)) / (2 * h)
        dg[2] = (_ks_mks_gcov(g, x1, x2 + h) - _ks_mks_gcov(g, x1, x2 - h)) / (2 * h)
        ginv = np.linalg.inv(_ks_mks_gcov(g, x1, x2))
        low = 0.5 * (np.einsum("acb->cab", dg) + np.einsum("bca->cab", dg) - dg)   # Gamma_{c ab}
        ref = np.einsum("dc,cab->dab", ginv, low)
        assert np.allclose(G, ref, rtol=0, atol=1e-7 * np.abs(ref).max())
        # Killing vectors d_t and d_phi: T^mu_nu Gamma^nu_{k mu} = 0 for symmetric T^{mu nu}
        gcov = _geom_struct(O.geom(i, j, CENT))["gcov"]
        for _ in range(5):
            A = rng.normal(size=(4, 4))
            Tup = A + A.T
            Tmix = Tup @ gcov                     # T^mu_nu
            scale = np.abs(Tmix).max() * np.abs(G).max()
            for kv in (0, 3):
                val = np.einsum("mn,nm->", Tmix, G[:, kv, :])
                assert abs(val) <= 1e-12 * scale


def test_oracle_con2prim_round_trip(torus_small):
    g, O = torus_small
    rng = np.random.default_rng(21)
    gam = g.gam
    nfail = 0
    for n in range(300):
        i, j = int(rng.integers(3, 27)), int(rng.integers(3, 27))
        G42 = O.geom(i, j, CENT)
        gc = _geom_struct(G42)["gcov"]
        rho = 10 ** rng.uniform(-3, 0)
        # velocities and fields of order one in the local orthonormal frame,
        # b^2 / rho <= ~1 (the magnetically dominated regime is covered by
        # the batched con-to-prim tests)
        ut = rng.normal(0, 0.5, 3) / np.sqrt(np.diag(gc)[1:])
        B = rng.normal(0, 0.3, 3) * math.sqrt(rho) / np.sqrt(np.diag(gc)[1:])
        p = np.array([rho, rho * 10 ** rng.uniform(-2, 0.5), *ut, *B])
        U = pyoracle.prim_to_flux(G42, gam, p, 0)
        guess = p * (1 + 0.1 * rng.uniform(-1, 1, 8))
        ret, q, gamma, bsq, it = pyoracle.utoprim_2d(G42, gam, U, guess)
        if ret:
            nfail += 1
            continue
        # the solver stops on a 1e-10 relative step in W = gamma^2 w (plus two
        # extra iterations); u, the small difference w - rho - p, is bounded
        # relative to the total energy density (which includes b^2)
        W = (abs(U[1]) + abs(U[0])) / G42[32]
        assert abs(q[0] - p[0]) <= 1e-7 * p[0], (n, p, q)
        assert abs(q[1] - p[1]) <= 1e-8 * W, (n, p, q)
        assert np.allclose(q[2:5], p[2:5], rtol=1e-7, atol=1e-8 * np.abs(p[2:5]).max()), (n, p, q)
        # B^i = U^i / sqrt(-g); U^i = sqrt(-g) (b^i u^t - b^t u^i) carries the cancellation
        assert np.allclose(q[5:], p[5:], rtol=1e-10, atol=1e-15)
        ret1, q1, _, _ = pyoracle.utoprim_1d(G42, gam, U, guess)
        if not ret1:
            assert abs(q1[0] - p[0]) <= 1e-7 * p[0] and abs(q1[1] - p[1]) <= 1e-8 * W
    assert nfail <= 3
    ret, *_ = pyoracle.utoprim_2d(O.geom(5, 5, CENT), gam, -np.ones(8), np.ones(8))
    assert ret == -100


def test_oracle_boundaries_and_timestep(torus_small):
    from coolharm3d_amd.problems import torus
    g, _ = torus_small
    O = OracleRun(g)
    O.set_prims(torus(g))
    O.fixup_bounds()
    p = O.prims()
    NG, n2, n3 = 3, g.n2, g.n3
    for gg in range(NG):
        lo, src = p[:, NG - 1 - gg, NG:NG + n3], p[:, NG + gg, NG:NG + n3].copy()
        src[..., [3, 6]] *= -1
        assert np.array_equal(lo, src)                    # x2 reflection, U2 and B2 flipped
        assert np.array_equal(p[:, :, NG - 1 - gg], p[:, :, NG + n3 - 1 - gg])    # x3 periodic
        assert np.array_equal(p[:, :, NG + n3 + gg], p[:, :, NG + gg])
    dt0 = 1.0e-5
    O.set_time(0.0, dt0)
    O.step(1)
    v = O.vmax()
    d = g.dx
    ref = g.cour * math.sqrt((1.0 / 3.0) / (v[1] ** 2 / d[1] ** 2 + v[2] ** 2 / d[2] ** 2 + v[3] ** 2 / d[3] ** 2))
    assert O.time()["dt"] == min(ref, 1.3 * dt0)


def test_oracle_divb_at_roundoff(torus_small):
    from coolharm3d_amd.problems import torus
    g, _ = torus_small
    O = OracleRun(g)
    O.set_prims(torus(g))
    O.fixup_bounds()
    O.set_time(0.0, 1.0e-5)
    p = O.prims()
    gdet = np.array([[_geom_struct(O.geom(i, j, CENT))["g"] for j in range(p.shape[1])] for i in range(p.shape[0])])
    scale = max(np.abs(p[..., 5 + d] * gdet[..., None]).max() / g.dx[1 + d] for d in range(3))
    d0 = O.divb_max()
    O.step(10)
    assert d0 <= 1e-12 * scale
    assert O.divb_max() <= 1e-12 * scale


# --------------------------------------------------------------------------
# against the reference build on a grid and seed no fixture covers

REF_SCRIPT = r"""
import sys, numpy as np
sys.path[:0] = [{root!r}, {oracle!r}]
from refwrap import RefRun, torus_params
from coolharm3d_amd.problems import torus, torus_grid
g = torus_grid(20, 18, 12)
R = RefRun("c1", 20, 18, 12, params=torus_params(g))
R.set_prims(torus(g, seed=7))
R.fixup_bounds(); R.reset_nfail()
R.set_time(0.0, 1.0e-5, g.cour)
p0 = R.prims()
conn = np.array([R.conn(i, j, 3) for (i, j) in ((3, 3), (10, 9), (22, 20))])
R.step(5)
np.savez({out!r}, p0=p0, p5=R.prims(), pflag=R.pflag(), nfail=R.nfail(), conn=conn, r_isco=R.scalars()["r_isco"])
"""


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "c1", "libref.so")),
                    reason="reference build oracle/_ref/c1 not present")
def test_oracle_against_reference_build_new_grid(tmp_path):
    from coolharm3d_amd.problems import torus_grid
    out = str(tmp_path / "ref.npz")
    code = REF_SCRIPT.format(root=ROOT, oracle=os.path.join(ROOT, "oracle"), out=out)
    subprocess.run([sys.executable, "-c", code], check=True, stdout=subprocess.DEVNULL)
    R = dict(np.load(out))
    g = torus_grid(20, 18, 12)
    O = OracleRun(g, r_isco=float(R["r_isco"]))
    for n, (i, j) in enumerate(((3, 3), (10, 9), (22, 20))):
        c = O.conn(i, j)
        assert np.allclose(c, R["conn"][n], rtol=0, atol=1e-14 * np.abs(c).max())
    O.set_prims(R["p0"])
    O.set_time(0.0, 1.0e-5)
    O.step(5)
    assert_rel_l1(interior(R["p5"]), interior(O.prims()), 1e-12, "new torus grid, 5 steps")
    assert np.array_equal(interior(O.pflag()), interior(R["pflag"]))
    ok, info = nfail_equal_modulo_cooling_init(R["nfail"], O.nfail(), 26)
    assert ok, info
