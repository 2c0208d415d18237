"""Binary-black-hole metric (the circumbinary-disk configuration).

The reference has no body for its post-Newtonian binary metrics
(src/metric.h:104-171 declares them only), so this row is parity-unpinned:
the metric is held to what the mathematics fixes -- the inverse and
determinant, the flat far field, the single-hole limit (harmonic Schwarzschild
to 1PN order), the pi-rotation symmetry of an equal-mass binary, the helical
symmetry of a circular orbit, and the connection against finite differences
of the metric (time derivatives included).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from coolharm3d_amd import _lib
from coolharm3d_amd.pn import PnMetric, bbh_block, CENT, FACE1


def one_zone(X1, X2, X3, **kw):
    c = bbh_block(1, 1, 1, ghosts=0, **kw)
    for l, x in ((1, X1), (2, X2), (3, X3)):
        c.startx[l] = x - 0.5 * c.dx[l]
    return c


def geom(X, t=0.0, hostcheck=True, **kw):
    return PnMetric(one_zone(*X, **kw), hostcheck=hostcheck).geometry(t, CENT).reshape(42)


def gcov(g42):
    return g42[16:32].reshape(4, 4)


def gcon(g42):
    return g42[0:16].reshape(4, 4)


POINTS = [(math.log(30.0), 0.5, 0.1), (math.log(45.0), 0.27, 2.2), (math.log(200.0), 0.81, 4.0),
          (math.log(259.0), 0.05, 6.0)]


def test_inverse_determinant_lapse():
    for X in POINTS:
        for t in (0.0, 250.0):
            g = geom(X, t)
            a, b = gcov(g), gcon(g)
            assert np.allclose(a @ b, np.eye(4), atol=1e-13 * np.abs(a).max() * np.abs(b).max())
            assert np.isclose(g[32], math.sqrt(-np.linalg.det(a)), rtol=1e-12)
            assert np.isclose(g[34], 1.0 / math.sqrt(-b[0, 0]), rtol=1e-15)
            assert np.allclose(g[38:42], [1.0 / g[34], -g[34] * b[0, 1], -g[34] * b[0, 2], -g[34] * b[0, 3]],
                               rtol=1e-14, atol=1e-300)


def test_single_hole_is_harmonic_schwarzschild():
    """m2 = 0: the 1PN truncation of Schwarzschild in harmonic coordinates,
    g00 = -1 + 2m/r - 2m^2/r^2, g_ij = (1 + 2m/r) delta_ij, no shift."""
    for X in POINTS:
        r, th = math.exp(X[0]), math.pi * X[1] + 0.3 * math.sin(2 * math.pi * X[1])
        dth = math.pi + 2 * math.pi * 0.3 * math.cos(2 * math.pi * X[1])
        g = gcov(geom(X, 0.0, m1=1.0, m2=0.0))
        f = 1 + 2 / r
        expect = np.diag([-1 + 2 / r - 2 / r ** 2, f * r * r, f * r * r * dth * dth, f * (r * math.sin(th)) ** 2])
        assert np.allclose(g, expect, rtol=1e-13, atol=1e-13 * r * r)


def test_flat_far_field():
    X = (math.log(1e7), 0.4, 1.0)
    r = 1e7
    g = gcov(geom(X))
    flat = gcov(geom(X, m1=0.0, m2=1e-300))
    scale = np.sqrt(np.abs(np.outer(np.diag(flat), np.diag(flat))))
    assert np.all(np.abs(g - flat) <= 3.0 / r * scale)
    assert np.isclose(flat[0, 0], -1.0, rtol=1e-12)


def test_equal_mass_pi_rotation():
    """Rotating an equal-mass binary by pi about its axis maps it onto itself."""
    for X in POINTS:
        for t in (0.0, 123.0):
            a = gcov(geom(X, t))
            b = gcov(geom((X[0], X[1], X[2] + math.pi), t))
            assert np.allclose(a, b, rtol=1e-12, atol=1e-13 * np.abs(a).max())


def test_circular_orbit_helical_symmetry():
    """Fixed circular orbit: g(t + s, phi + omega s) = g(t, phi)."""
    omega = math.sqrt(1.0 / 20.0 ** 3)
    for X in POINTS[:2]:
        a = gcov(geom(X, 10.0, inspiral=False))
        s = 37.0
        b = gcov(geom((X[0], X[1], X[2] + omega * s), 10.0 + s, inspiral=False))
        assert np.allclose(a, b, rtol=1e-11, atol=1e-12 * np.abs(a).max())


@pytest.mark.parametrize("inspiral,coords", [(True, "mks"), (False, "mks"), (True, "diag3")])
def test_connection_against_finite_differences(inspiral, coords):
    """The connection (metric derivatives by forward-mode differentiation)
    against central differences of the metric in t and X, on the modified
    Kerr-Schild and the COORD_DIAGONAL3 theta maps (h_slope 0.0729...,
    exponent 9: the reference's binary-disk values)."""
    X, t = (math.log(40.0), 0.37, 1.3), 50.0
    kw = dict(inspiral=inspiral, coords=coords)
    if coords == "diag3":
        kw["h_slope"] = 0.0729419909586
    conn = PnMetric(one_zone(*X, **kw), hostcheck=True).connection(t).reshape(4, 4, 4)
    h = [1e-2, 1e-5, 1e-5, 1e-5]

    def g_at(dt=0.0, dX=(0.0, 0.0, 0.0)):
        return gcov(geom((X[0] + dX[0], X[1] + dX[1], X[2] + dX[2]), t + dt, **kw))

    dg = np.empty((4, 4, 4))   # dg[c][a][b] = d_c g_ab
    dg[0] = (g_at(h[0]) - g_at(-h[0])) / (2 * h[0])
    for c in range(1, 4):
        e = [0.0, 0.0, 0.0]
        e[c - 1] = h[c]
        m = [-x for x in e]
        dg[c] = (g_at(0.0, e) - g_at(0.0, m)) / (2 * h[c])
    gi = np.linalg.inv(g_at())
    # low[c][a][b] = Gamma_{cab} = (d_a g_cb + d_b g_ca - d_c g_ab) / 2
    low = np.empty((4, 4, 4))
    for c in range(4):
        for a in range(4):
            for b in range(4):
                low[c, a, b] = 0.5 * (dg[a, c, b] + dg[b, c, a] - dg[c, a, b])
    fd = np.einsum("dc,cab->dab", gi, low)
    scale = np.abs(fd).max()
    assert np.allclose(conn, fd, rtol=1e-5, atol=1e-7 * scale)
    assert np.allclose(conn, conn.transpose(0, 2, 1), rtol=0, atol=1e-15 * scale)
    # the time derivative enters: Gamma^t_tt differs between a moving and a frozen metric
    if inspiral:
        assert abs(conn[0, 0, 0]) > 0


def test_source_from_metric_derivatives_equals_connection_form():
    """The step's source term for the binary metric contracts the stress
    tensor with the metric's derivatives, T^{mu lambda} d_nu g_{mu lambda} / 2;
    for a metric connection that equals T^mu_kappa Gamma^kappa_{nu mu}.  Here
    with this metric's Gamma (time derivatives included) and d g formed from
    it, for random symmetric stress tensors (the identity needs T^{mu lambda}
    symmetric, as the MHD stress tensor is)."""
    rng = np.random.default_rng(5)
    X, t = (math.log(40.0), 0.37, 1.3), 50.0
    g42 = geom(X, t)
    gcv, gcn = gcov(g42), gcon(g42)
    G = PnMetric(one_zone(*X), hostcheck=True).connection(t).reshape(4, 4, 4)   # G[k][nu][mu]
    dg = np.einsum("ak,kbc->cab", gcv, G) + np.einsum("bk,kac->cab", gcv, G)    # dg[c][a][b] = d_c g_ab
    # the spatial block of this metric is diagonal: its off-diagonal derivatives vanish
    for c in range(4):
        for a, b in ((1, 2), (1, 3), (2, 3)):
            assert abs(dg[c, a, b]) <= 1e-12 * np.abs(dg[c]).max()
    for _ in range(5):
        S = rng.normal(size=(4, 4))
        Tuu = S + S.T                                      # T^{mu lambda}
        Tmix = Tuu @ gcv                                   # T^mu_kappa
        a = np.einsum("mk,knm->n", Tmix, G)
        b = 0.5 * np.einsum("ml,nml->n", Tuu, dg)
        assert np.allclose(a, b, rtol=1e-10, atol=1e-12 * np.abs(a).max())


def test_entry_point_errors():
    lib = _lib.load(hostcheck=True)
    c = one_zone(*POINTS[0])
    c.b0 = 0.0
    out = np.zeros(42)
    rc = lib.hgpu_pn_geometry(None, c, 0.0, CENT, out.ctypes.data)
    assert rc < 0 and b"separation" in lib.hgpu_last_error()
    c = one_zone(*POINTS[0])
    rc = lib.hgpu_pn_geometry(None, c, 0.0, 7, out.ctypes.data)
    assert rc < 0 and b"pos" in lib.hgpu_last_error()


@pytest.mark.gpu
def test_gpu_metric_matches_host():
    c = bbh_block(30, 16, 40)
    G, H = PnMetric(c), PnMetric(c, hostcheck=True)
    for pos in (FACE1, CENT):
        a, b = G.geometry(75.0, pos), H.geometry(75.0, pos)
        assert np.allclose(a, b, rtol=1e-13, atol=1e-13 * np.abs(b).max(axis=-1, keepdims=True))
    a, b = G.connection(75.0), H.connection(75.0)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-13 * np.abs(b).max())


# ---------------------------------------------------------------------------
# the advance on the binary metric (geometry levels t, t + dt/2, t + dt per step)

def _bbh_run(hostcheck, precision="fp64", steps=3, graph=False, shape=(20, 16, 24), in_kernels=None,
             coords="mks"):
    from coolharm3d_amd import Harm3D, bbh_grid, bbh_disk
    from coolharm3d_amd.problems import bbh_gdet
    g = bbh_grid(*shape, coords=coords)
    p0 = bbh_disk(g, bbh_gdet(g, PnMetric, hostcheck=True))
    H = Harm3D(g, hostcheck=hostcheck, precision=precision, track_nfail=1)
    if in_kernels is not None:
        H.set_metric_eval(in_kernels)
    H.set_prims(p0)
    H.fixup_bounds()
    H.set_time(0.0, 1.0e-3)
    if graph:
        H.set_graph(True)
    d0 = H.divb_max()
    H.step_ch(steps)
    return H, d0


@pytest.mark.parametrize("precision,coords", [("fp64", "mks"), ("mixed", "mks"), ("fp64", "diag3")])
def test_bbh_step_host_build(precision, coords):
    """A few steps of the circumbinary disk on the moving metric: finite state,
    constrained transport keeps div B at round-off, the clock advances (also
    on the reference's binary-disk theta map, COORD_DIAGONAL3)."""
    H, d0 = _bbh_run(True, precision, coords=coords)
    p = H.prims()
    assert np.isfinite(p).all()
    bmax = np.abs(p[3:-3, 3:-3, 3:-3, 5:8]).max()
    assert H.divb_max() <= max(10 * d0, 1e-12 * bmax / min(H.grid.dx[1:]))
    t, dt, _ = H.time()
    assert t > 0 and dt > 0
    assert (p[3:-3, 3:-3, 3:-3, 0] > 0).all() and (p[3:-3, 3:-3, 3:-3, 1] > 0).all()


@pytest.mark.gpu
@pytest.mark.parametrize("precision,coords", [("fp64", "mks"), ("mixed", "mks"), ("fp64", "diag3")])
def test_gpu_bbh_step_matches_host(precision, coords):
    """GPU vs the host build of the same bodies on the moving metric.  In FP64
    they differ only through the device's sin/cos/exp/pow in the metric; the
    mixed face solve also uses the device's single-instruction FP32
    reciprocal, square root, log2 and exp2 (about 1 ulp each, the host build
    the correctly rounded libm forms), so its bar is FP32 rounding."""
    G, _ = _bbh_run(False, precision, coords=coords)
    H, _ = _bbh_run(True, precision, coords=coords)
    a, b = G.prims()[3:-3, 3:-3, 3:-3].reshape(-1, 8), H.prims()[3:-3, 3:-3, 3:-3].reshape(-1, 8)
    l1 = np.abs(a - b).sum(axis=0) / np.maximum(np.abs(b).sum(axis=0), 1e-300)
    print(f"[bbh] {precision} GPU vs host rel L1 after 3 steps: {l1}", flush=True)
    assert np.all(l1 <= (1e-9 if precision == "fp64" else 1e-6))
    assert G.time() == pytest.approx(H.time(), rel=1e-12)


def test_metric_eval_modes_host_build():
    """The host build reads stored levels only: asking it to evaluate the
    metric inside the kernels is an error, and it reports the stored mode."""
    from coolharm3d_amd import Harm3D, bbh_grid
    from coolharm3d_amd._lib import HgpuError
    H = Harm3D(bbh_grid(12, 8, 8), hostcheck=True)
    assert H.metric_eval() is False
    with pytest.raises(HgpuError, match="evaluated in the kernels"):
        H.set_metric_eval(True)
    H.set_metric_eval(False)
    # the stored mode holds its three record levels from the start
    ng, _ = H.geometry_counts()
    assert H.device_bytes() > 3 * 42 * 8 * ng


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_bbh_metric_in_kernels_matches_stored(precision):
    """Face and update records evaluated inside the kernels from the level's
    orbit give the same state, flags and counters, bit for bit, as the
    stored levels (the same functions at the same times)."""
    A, _ = _bbh_run(False, precision, steps=3, in_kernels=True)
    B, _ = _bbh_run(False, precision, steps=3, in_kernels=False)
    assert A.metric_eval() and not B.metric_eval()
    # the evaluated mode has not allocated the stored levels: three FP64
    # record levels (42 doubles a record) and two FP32 face copies (43 floats:
    # the same fields and 1 - alpha; one record each outside mixed precision)
    ng, _ = A.geometry_counts()
    levels = 3 * 42 * 8 * ng + 2 * 43 * 4 * (ng if precision == "mixed" else 1)
    print(f"[bbh] device bytes: evaluated {A.device_bytes()}, stored {B.device_bytes()}", flush=True)
    assert B.device_bytes() - A.device_bytes() == levels
    assert np.array_equal(A.prims(), B.prims())
    assert np.array_equal(A.pflag(), B.pflag())
    assert np.array_equal(A.nfail(), B.nfail())
    assert A.time() == B.time()
    assert A.divb_max() == B.divb_max()
    # the evaluated mode keeps no centre level: the download forms the state's
    ga, da = A.geometry()
    gb, db = B.geometry()
    cent = slice(3 * A.nt, 4 * A.nt)   # position 3 (zone centres) of [5][NT]
    assert np.array_equal(ga[cent], gb[cent])
    assert np.array_equal(da, db)
    assert A.device_bytes() == B.device_bytes()   # the download allocated them
    # and the evaluated mode still steps the same after the allocation
    A.step_ch(1)
    B.step_ch(1)
    assert np.array_equal(A.prims(), B.prims())


@pytest.mark.gpu
def test_gpu_bbh_switch_to_stored_between_stages():
    """Leaving the evaluated mode between the predictor and the corrector: the
    corrector reads the step's stored levels, filled at the switch, and the
    step equals one taken wholly in the evaluated mode, bit for bit."""
    A, _ = _bbh_run(False, "fp64", steps=2, in_kernels=True)
    B, _ = _bbh_run(False, "fp64", steps=2, in_kernels=True)
    for H, switch in ((A, False), (B, True)):
        H.advance(0)
        if switch:
            H.set_metric_eval(False)
        H.advance(1)
        H.end_step()
    assert not B.metric_eval()
    assert np.array_equal(A.prims(), B.prims())
    assert np.array_equal(A.pflag(), B.pflag())
    assert np.array_equal(A.nfail(), B.nfail())
    assert A.time() == B.time()
    A.step_ch(1)
    B.step_ch(1)
    assert np.array_equal(A.prims(), B.prims())


@pytest.mark.gpu
def test_gpu_bbh_autotune_keeps_results():
    """Launch shapes on the binary metric (evaluated records): the autotuner
    leaves the state unchanged and the tuned steps equal the untuned ones."""
    E, _ = _bbh_run(False, "mixed", steps=2)
    T, _ = _bbh_run(False, "mixed", steps=2)
    before = T.prims()
    rep = T.autotune(reps=1)
    print(f"[bbh] autotune picked {rep}", flush=True)
    assert np.array_equal(before, T.prims())
    E.step_ch(2)
    T.step_ch(2)
    assert np.array_equal(E.prims(), T.prims())
    assert E.time() == T.time()


@pytest.mark.gpu
def test_gpu_bbh_graph_replay():
    E, _ = _bbh_run(False, "mixed", steps=4)
    G, _ = _bbh_run(False, "mixed", steps=4, graph=True)
    assert np.array_equal(E.prims(), G.prims())
    assert E.time() == G.time()


# ---------------------------------------------------------------------------
# The binary's own terms, pinned by routes independent of the closed form the
# library evaluates.

def _orbit(t, m1, m2, b, phi0):
    """Hole positions of the fixed circular orbit (Newtonian Kepler frequency,
    centre of mass at the origin)."""
    M = m1 + m2
    ph = phi0 + math.sqrt(M / b ** 3) * t
    n = np.array([math.cos(ph), math.sin(ph), 0.0])
    return (m2 / M) * b * n, -(m1 / M) * b * n


def _numerical(x):
    """(X1, X2, X3) of Cartesian x for R0 = 0, th_beg = 0, th_length = pi,
    h_slope = 0: X1 = log r, X2 = theta / pi, X3 = phi."""
    r = float(np.linalg.norm(x))
    return (math.log(r), math.acos(x[2] / r) / math.pi, math.atan2(x[1], x[0]) % (2 * math.pi))


def _UV_library(g, r, sth):
    """(U, V) of metric records g[..., 42]: g_phiphi = (1 + 2U) (r sin theta)^2
    gives the Newtonian U, g_tt = -1 + 2V - 2U^2 then V (the time coordinate is
    the harmonic one, so g_tt is the same in numerical coordinates)."""
    U = 0.5 * (g[..., 31] / (r * sth) ** 2 - 1.0)
    return U, 0.5 * (g[..., 16] + 1.0) + U * U


def _V_library(x, t, **kw):
    """V of the library's metric at Cartesian x."""
    g = geom(_numerical(x), t, h_slope=0.0, **kw)
    r = float(np.linalg.norm(x))
    return float(_UV_library(g, r, math.sqrt(x[0] ** 2 + x[1] ** 2) / r)[1])


def _V_retarded(x, t, m1, m2, b, phi0):
    """The 1PN potential of two point masses from its definition as a retarded
    integral expanded in 1/c (Blanchet & Damour 1989; Blanchet, Faye & Ponsot
    1998): V = sum_a sigma_a / r_a + (1/2) d^2/dt^2 sum_a m_a r_a, with the
    point-mass source sigma_a = m_a (1 + 3/2 v_a^2 - m_b / r_ab).  The second
    time derivative is taken numerically along the orbit, so this route shares
    no algebra with the expanded closed form."""
    m = (m1, m2)
    h = 0.5

    def mr(s):
        y = _orbit(s, m1, m2, b, phi0)
        return sum(m[a] * np.linalg.norm(x - y[a]) for a in range(2))

    d2 = (-mr(t - 2 * h) + 16 * mr(t - h) - 30 * mr(t) + 16 * mr(t + h) - mr(t + 2 * h)) / (12 * h * h)
    y = _orbit(t, m1, m2, b, phi0)
    ya, yb = _orbit(t - 1e-3, m1, m2, b, phi0), _orbit(t + 1e-3, m1, m2, b, phi0)
    V = 0.5 * d2
    for a in range(2):
        v2 = float(np.sum(((yb[a] - ya[a]) / 2e-3) ** 2))
        V += m[a] * (1.0 + 1.5 * v2 - m[1 - a] / b) / np.linalg.norm(x - y[a])
    return V


BIN = dict(m1=0.3, m2=0.7, sep=20.0, phi0=0.4)


def _pts(t):
    y1, y2 = _orbit(t, BIN["m1"], BIN["m2"], BIN["sep"], BIN["phi0"])
    return [y1 + np.array([3.0, 1.0, 0.5]), y1 + np.array([-2.0, 0.0, 2.5]), y2 + np.array([0.5, -4.0, -1.0]),
            y2 + np.array([0.0, 2.0, 6.0]), 0.5 * (y1 + y2) + np.array([0.0, 0.0, 3.0]),
            np.array([35.0, -12.0, 20.0]), np.array([-80.0, 50.0, -30.0])]


def test_potential_equals_retarded_integral_form():
    """Every coefficient of V -- the velocity terms and both interaction
    terms, including the one from the holes' accelerations -- against the
    retarded-integral definition, for unequal masses, near each hole, between
    them and farther out: agreement to 1e-8 relative, the finite-difference
    error of the time derivative near a hole (a 1 % change of the m1 m2
    acceleration term alone moves V by ~1e-5 relative there)."""
    kw = dict(m1=BIN["m1"], m2=BIN["m2"], sep=BIN["sep"], phi0=BIN["phi0"], inspiral=False)
    for t in (0.0, 77.0):
        for x in _pts(t):
            vl = _V_library(x, t, **kw)
            vr = _V_retarded(x, t, BIN["m1"], BIN["m2"], BIN["sep"], BIN["phi0"])
            assert abs(vl - vr) <= 1e-8 * abs(vr), (t, x, vl, vr)


def test_potential_check_catches_a_wrong_interaction_coefficient():
    """The check above fails for a plausible mistake: the acceleration
    interaction term m_a m_b (r_b^2 - r_a^2 - r_ab^2)/(4 r_a r_ab^3) with the
    wrong sign -- it is the part of V that the retarded form gets from the
    second time derivative, so dropping that derivative must show up."""
    t = 0.0
    m1, m2, b, phi0 = BIN["m1"], BIN["m2"], BIN["sep"], BIN["phi0"]
    y = _orbit(t, m1, m2, b, phi0)
    for x in _pts(t)[:4]:
        vr = _V_retarded(x, t, m1, m2, b, phi0)
        r = [np.linalg.norm(x - y[a]) for a in range(2)]
        m = (m1, m2)
        acc = sum(m[a] * m[1 - a] * (r[1 - a] ** 2 - r[a] ** 2 - b * b) / (4 * r[a] * b ** 3) for a in range(2))
        assert abs(acc) > 1e-6
        vl = _V_library(x, t, m1=m1, m2=m2, sep=b, phi0=phi0, inspiral=False)
        assert abs((vl - 2 * acc) - vr) > 1e3 * 1e-8 * abs(vr)


def test_far_field_monopole_is_the_binding_energy():
    """The sphere average of V at radius R >> b is M_1PN / R with M_1PN the
    binary's total mass-energy, M + (1/2) mu v^2 - m1 m2 / b = M - m1 m2/(2b)
    on a circular orbit (up to (b/R)^2 corrections of the non-harmonic 1PN
    terms).  This fixes the remaining coefficient, the -m_a m_b/(r_a r_ab)
    term, which no vacuum field equation sees."""
    m1, m2, b = BIN["m1"], BIN["m2"], BIN["sep"]
    R = 4000.0
    n2, n3 = 96, 96
    c = bbh_block(1, n2, n3, ghosts=0, m1=m1, m2=m2, sep=b, phi0=BIN["phi0"], inspiral=False, h_slope=0.0)
    c.startx[1] = math.log(R) - 0.5 * c.dx[1]
    g = PnMetric(c, hostcheck=True).geometry(0.0, CENT).reshape(n2, n3, 42)
    th = math.pi * (np.arange(n2) + 0.5) / n2
    V = _UV_library(g, R, np.sin(th)[:, None])[1]
    w = np.sin(th)[:, None] * np.ones((1, n3))
    mono = R * float((V * w).sum() / w.sum())
    expect = m1 + m2 - m1 * m2 / (2 * b)
    assert abs(mono - expect) <= 2e-6, (mono, expect)
    # the binding term itself is ~5e-3: a 10 % change of the -m_a m_b/(r_a r_ab)
    # coefficient (2 m1 m2/b x 0.1 = 2e-3) is far outside the tolerance
    assert m1 * m2 / (2 * b) > 1e3 * 2e-6


def test_harmonic_gauge_condition():
    """The metric is in harmonic gauge to the order it is built: the time
    coordinate obeys box t = d_mu(sqrt(-g) g^{mu t}) / sqrt(-g) = 0 up to
    relative remainders of order U and v^2 (the terms the 1PN metric leaves
    out).  At leading order this is dV/dt + d_i V_i = 0, so it fixes the shift
    g_ti = -4 V_i against the time dependence of V: with -3 V_i in place of
    -4 V_i the residual would be a quarter of d_t(sqrt(-g) g^tt), against a
    bar of 1.5 U + 0.01 (< 0.09 at these points).  Finite differences in t and the numerical coordinates (in which
    sqrt(-g) carries the coordinate Jacobian, so the divergence is the same
    invariant)."""
    kw = dict(m1=0.4, m2=0.6, sep=100.0, phi0=0.2, inspiral=False, h_slope=0.0)
    y1, y2 = _orbit(30.0, 0.4, 0.6, 100.0, 0.2)
    for x in (y1 + np.array([12.0, 5.0, 3.0]), y2 + np.array([-8.0, 10.0, -6.0]), np.array([20.0, 130.0, 40.0])):
        X = _numerical(x)
        h = [1.0, 1e-4, 1e-4, 1e-4]

        def A(dt=0.0, d=None, s=0.0):
            Xs = list(X)
            if d is not None:
                Xs[d] += s
            g = geom(tuple(Xs), 30.0 + dt, **kw)
            return g[32] * gcon(g)[:, 0]          # sqrt(-g) g^{mu t}

        dA0 = (A(h[0])[0] - A(-h[0])[0]) / (2 * h[0])
        terms = [dA0] + [(A(0.0, i, h[i + 1])[i + 1] - A(0.0, i, -h[i + 1])[i + 1]) / (2 * h[i + 1])
                         for i in range(3)]
        res = sum(terms)
        scale = abs(dA0)
        U = 0.4 / np.linalg.norm(x - y1) + 0.6 / np.linalg.norm(x - y2)
        assert scale > 0 and 1.5 * U + 0.01 < 0.1
        assert abs(res) <= (1.5 * U + 0.01) * scale, (terms, res)
