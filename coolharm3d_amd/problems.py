"""Initial data for the two problem families the hot path is exercised on.

These are problem set-ups (the role of the reference's ``init.*.c`` files), not
part of the per-substep advance.  They produce primitive arrays in the
reference's layout ``[i][j][k][var]`` over all cells including ghosts, with
``var`` = (rho, u, u~1, u~2, u~3, B1, B2, B3).  Ghost cells are filled with
the interior formulas; the boundary conditions and floors of the integrator
are applied afterwards by the caller (``Harm3D.fixup_and_bounds`` on the GPU,
``fixup(p); bounds(p,0)`` in the reference, ``src/init.torus.c:157-163``).

* :func:`balsara1` -- relativistic MHD shock tube, Balsara (2001) test 1, as set
  up by ``src/init.balsara1.c:44-136``.
* :func:`torus` -- Fishbone & Moncrief (1976) constant-angular-momentum torus in
  Kerr-Schild coordinates with a single poloidal loop, following the recipe of
  ``src/init.torus.c:337-535`` (Boyer-Lindquist u^phi, Kerr-Schild transform,
  A_phi proportional to max(rho/rho_max - 0.2, 0), normalised so that
  (Gamma-1) u_max / (b^2_max / 2) = beta) with a seeded +-2 % perturbation of u.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

NG = 3
NP = 8
RHO, UU, U1, U2, U3, B1, B2, B3 = range(8)


@dataclass
class GridSpec:
    """Grid and coordinate parameters shared by the GPU library and the checkers."""
    n1: int
    n2: int
    n3: int
    startx: tuple   # numerical-coordinate origin of cell (0,0,0) corner (ghosts included)
    dx: tuple       # (dt, dx1, dx2, dx3)
    metric: str     # "minkowski" or "kerr_schild"
    coords: str     # "identity", "mks" (COORD_DIAGONAL) or "diag3" (COORD_DIAGONAL3)
    gam: float
    cour: float
    t_end: float
    a: float = 0.0
    R0: float = 0.0
    h_slope: float = 0.0
    th_beg: float = 0.0
    th_length: float = math.pi
    diag3_exponent: int = 9
    diag3_factor: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def shape(self):
        return (self.n1 + 2 * NG, self.n2 + 2 * NG, self.n3 + 2 * NG)


def balsara1_grid(n1=1024, cour=0.4, dt0=1.0e-5):
    """Grid of ``src/init.balsara1.c:44-75`` (x in [-0.5,0.5), Gamma = 2)."""
    dx1 = 1.0 / n1
    dx2 = dx3 = 1.0
    startx = (0.0, -0.5 - NG * dx1, -0.5 - NG * dx2, -0.5 - NG * dx3)
    return GridSpec(n1, 1, 1, startx, (dt0, dx1, dx2, dx3), "minkowski", "identity",
                    gam=2.0, cour=cour, t_end=0.4)


def balsara1(grid: GridSpec):
    """Left/right states of Balsara test 1 (``src/init.balsara1.c:93-130``)."""
    N1T, N2T, N3T = grid.shape
    p = np.zeros((N1T, N2T, N3T, NP))
    i = np.arange(N1T)
    x1 = grid.startx[1] + (i + 0.5) * grid.dx[1]
    left = x1 < 0.0
    gam = grid.gam
    p[:, :, :, RHO] = np.where(left, 1.0, 0.125)[:, None, None]
    p[:, :, :, UU] = np.where(left, 1.0 / (gam - 1.0), 0.1 / (gam - 1.0))[:, None, None]
    p[:, :, :, B1] = 0.5
    p[:, :, :, B2] = np.where(left, 1.0, -1.0)[:, None, None]
    return p


def torus_grid(n1=192, n2=192, n3=128, a=0.9, Rin=1.3, Rout=240.0, hslope_harm=0.3,
               gam=5.0 / 3.0, cour=0.8, x3_extent=0.5 * math.pi, dt0=1.0e-5, t_end=1.0e4, coords="mks",
               diag3_h_slope=0.0729419909586, diag3_exponent=9):
    """Kerr-Schild grid of ``src/init.torus.c:54-118`` with R0 = 0.

    coords "mks" (COORD_DIAGONAL): ``hslope_harm`` is HARM's h; the
    reference's sin coefficient is (1-h)/2 (``src/coord.c:252``), so h = 0.3
    gives h_slope = 0.35.  coords "diag3" (COORD_DIAGONAL3, ``src/coord.c:270-276``):
    theta = pi/2 (1 + h_slope f + diag3_factor f^n), f = 2 x2 - 1, with the
    values of the reference's binary-disk set-up (``src/init.nz_pn_bbh_disk.c:332-360``):
    h_slope = 0.0729419909586 (the polar-to-equatorial theta spacing ratio),
    n = 9, diag3_factor = 1 - 2 th_cutout/pi - h_slope.
    """
    th_cutout = 1.0e-15 * math.pi
    th_beg = th_cutout
    th_length = (math.pi - th_cutout) - th_beg
    if coords == "diag3":
        h_slope = diag3_h_slope
        factor = 1.0 - 2 * th_cutout / math.pi - h_slope
    else:
        h_slope = 0.5 * (1.0 - hslope_harm)
        factor = 0.0
    L1 = math.log(Rout / Rin)
    dx1, dx2, dx3 = L1 / n1, 1.0 / n2, x3_extent / n3
    startx = (0.0, math.log(Rin) - NG * dx1, -NG * dx2, -NG * dx3)
    return GridSpec(n1, n2, n3, startx, (dt0, dx1, dx2, dx3), "kerr_schild", coords,
                    gam=gam, cour=cour, t_end=t_end, a=a, R0=0.0, h_slope=h_slope,
                    th_beg=th_beg, th_length=th_length, diag3_exponent=diag3_exponent, diag3_factor=factor,
                    extra=dict(Rin=Rin, Rout=Rout, x3_extent=x3_extent))


def _seesaw01(x):
    """Reflection of x into [0,1] with period 2 (``src/coord.c:4316``)."""
    x = np.asarray(x, dtype=np.float64)
    shift = np.where(x < 0.0, 1 + np.floor(np.abs(x) / 2.0), -np.floor(np.abs(x) / 2.0))
    x1 = x + 2.0 * shift
    return np.where(x1 < 1.0, x1, 2.0 - x1)


def theta_map(grid: GridSpec, xp2):
    """theta(x2) and dtheta/dx2 of the grid's spherical map, as x_of_xp() and
    dx_dxp_calc() form them (``src/coord.c:245-276``, ``:447-472``)."""
    xp2 = np.asarray(xp2, dtype=np.float64)
    x2 = _seesaw01(xp2)
    if grid.coords == "diag3":
        n, c = grid.diag3_exponent, grid.diag3_factor
        f = 2.0 * x2 - 1.0
        th = 0.5 * math.pi * (1.0 + grid.h_slope * f + c * np.power(f, float(n)))
        dth = math.pi * (grid.h_slope + n * c * np.power(f, float(n - 1)))
    else:
        th = grid.th_length * x2 + grid.th_beg + grid.h_slope * np.sin(2.0 * math.pi * x2)
        dth = grid.th_length + 2.0 * math.pi * grid.h_slope * np.cos(2.0 * math.pi * xp2)
    return th, dth


def mks_r_th(grid: GridSpec, xp1, xp2):
    r = grid.R0 + np.exp(xp1)
    th, _ = theta_map(grid, xp2)
    return r, th


def torus(grid: GridSpec, rin=6.0, rmax=12.0, beta=100.0, kappa=1.0e-3, randfact=0.04,
          seed=1234, rhomin=1.0e-7, uumin=1.0e-9, rhopow=-1.5, uupow=-2.5):
    """Fishbone-Moncrief torus with a single poloidal loop (see module doc)."""
    N1T, N2T, N3T = grid.shape
    a, gam = grid.a, grid.gam
    ig = np.arange(N1T)
    jg = np.arange(N2T)
    xp1 = grid.startx[1] + (ig + 0.5) * grid.dx[1]
    xp2 = grid.startx[2] + (jg + 0.5) * grid.dx[2]
    r1, _ = mks_r_th(grid, xp1, np.full_like(xp1, 0.5))
    _, th1 = mks_r_th(grid, np.zeros_like(xp2), xp2)
    r = r1[:, None] * np.ones((1, N2T))
    th = np.ones((N1T, 1)) * th1[None, :]
    sth, cth = np.sin(th), np.cos(th)

    # angular momentum at the pressure maximum (prograde), src/init.torus.c:586-597
    sr = math.sqrt(rmax)
    l = ((rmax * rmax * (rmax * rmax + a * a) - 2 * rmax * a * a - a * sr * (rmax * rmax - a * a))
         / (sr * rmax * (rmax * rmax - 3.0 * rmax + 2.0 * a * sr)))

    DD = r * r - 2.0 * r + a * a
    AA = (r * r + a * a) ** 2 - DD * a * a * sth * sth
    SS = r * r + a * a * cth * cth
    DDin = rin * rin - 2.0 * rin + a * a
    AAin = (rin * rin + a * a) ** 2 - DDin * a * a
    SSin = rin * rin
    with np.errstate(invalid="ignore", divide="ignore"):
        lnh = (0.5 * np.log((1.0 + np.sqrt(1.0 + 4.0 * (l * l * SS * SS) * DD / (AA * sth * AA * sth)))
                            / (SS * DD / AA))
               - 0.5 * np.sqrt(1.0 + 4.0 * (l * l * SS * SS) * DD / (AA * AA * sth * sth))
               - 2.0 * a * r * l / AA
               - (0.5 * math.log((1.0 + math.sqrt(1.0 + 4.0 * (l * l * SSin * SSin) * DDin / (AAin * AAin)))
                                 / (SSin * DDin / AAin))
                  - 0.5 * math.sqrt(1.0 + 4.0 * (l * l * SSin * SSin) * DDin / (AAin * AAin))
                  - 2.0 * a * rin * l / AAin))
    inside = (r >= rin) & np.isfinite(lnh) & (lnh >= 0.0) & (DD > 0)
    lnh = np.where(inside, lnh, 0.0)
    hm1 = np.exp(lnh) - 1.0
    rho = np.where(inside, np.power(np.maximum(hm1, 0.0) * (gam - 1.0) / (kappa * gam), 1.0 / (gam - 1.0)), 0.0)
    u = kappa * np.power(rho, gam) / (gam - 1.0)

    # u^phi in BL, src/init.torus.c:459-462
    with np.errstate(invalid="ignore", divide="ignore"):
        expm2chi = SS * SS * DD / (AA * AA * sth * sth)
        up1 = np.sqrt((-1.0 + np.sqrt(1.0 + 4.0 * l * l * expm2chi)) / 2.0)
        up = 2.0 * a * r * np.sqrt(1.0 + up1 * up1) / np.sqrt(AA * SS * DD) + np.sqrt(SS / AA) * up1 / sth
        # BL metric components for u^t (setutcon, src/metric.c:78)
        gtt = -(1.0 - 2.0 * r / SS)
        gtp = -2.0 * a * r * sth * sth / SS
        gpp = AA * sth * sth / SS
        bq = gtp * up
        cq = gpp * up * up + 1.0
        ut = (-bq - np.sqrt(bq * bq - gtt * cq)) / gtt
        # u^r = 0, so the BL->KS transform leaves u^t, u^phi unchanged
        # (src/metric.c:1940); MKS u'^1 = u'^2 = 0.  ucon2pr with KS/MKS gcon:
        gcon00 = -1.0 - 2.0 * r / SS
        gcon01 = (2.0 * r / SS) / (r - grid.R0)      # transformed by dx1/dr = 1/(r - R0)
        v1 = -gcon01 * ut / gcon00
        v3 = up.copy()
    v1 = np.where(inside, v1, 0.0)
    v3 = np.where(inside, v3, 0.0)

    rhomax = rho[NG:-NG, NG:-NG].max()
    umax = u[NG:-NG, NG:-NG][(r[NG:-NG, NG:-NG] > rin)].max()

    p = np.zeros((N1T, N2T, N3T, NP))
    rho_n = rho / rhomax
    u_n = u / rhomax
    umax /= rhomax
    rng = np.random.default_rng(seed)
    pert = 1.0 + randfact * (rng.random((N1T, N2T, N3T)) - 0.5)
    p[..., RHO] = rho_n[:, :, None]
    p[..., UU] = u_n[:, :, None] * pert
    p[..., U1] = v1[:, :, None]
    p[..., U3] = v3[:, :, None]

    # floors (src/fixup.c:110, src/coord.c:3791-3797) on the atmosphere
    rhoflr = rhomin * np.power(r, rhopow)
    uuflr = uumin * np.power(r, uupow)
    p[..., RHO] = np.maximum(p[..., RHO], rhoflr[:, :, None])
    p[..., UU] = np.maximum(p[..., UU], uuflr[:, :, None])

    # vector potential at corners (i,j) from the 4 surrounding cell densities
    rho2 = p[:, :, NG, RHO]
    A = np.zeros((N1T + 1, N2T + 1))
    ravg = 0.25 * (rho2[1:, 1:] + rho2[:-1, 1:] + rho2[1:, :-1] + rho2[:-1, :-1])
    A[1:-1, 1:-1] = np.maximum(ravg / 1.0 - 0.2, 0.0)
    # gdet of MKS at cell centres: Sigma |sin th| * (r - R0) * dth/dx2
    dthdx2 = theta_map(grid, xp2)[1]
    gdet = SS * np.abs(sth) * (r - grid.R0) * dthdx2[None, :]
    B1v = -(A[:-1, :-1] - A[:-1, 1:] + A[1:, :-1] - A[1:, 1:]) / (2.0 * grid.dx[2] * gdet)
    B2v = (A[:-1, :-1] + A[:-1, 1:] - A[1:, :-1] - A[1:, 1:]) / (2.0 * grid.dx[1] * gdet)

    # b^2 from the KS metric (b^2 = (B^2 + (B.u)^2)/ (u^t alpha)^2 ... computed
    # directly: b^t = B^i u_i, b^i = (B^i + b^t u^i)/u^t, b^2 = b.b)
    gc = _ks_gcov_mks(grid, r, th)
    ucon = _ucon_from_prims(gc, r, th, grid, v1, np.zeros_like(v1), v3)
    ucov = np.einsum("...ab,...b->...a", gc, ucon)
    Bv = np.stack([np.zeros_like(B1v), B1v, B2v, np.zeros_like(B1v)], axis=-1)
    bt = (Bv[..., 1] * ucov[..., 1] + Bv[..., 2] * ucov[..., 2] + Bv[..., 3] * ucov[..., 3])
    bcon = np.stack([bt,
                     (Bv[..., 1] + bt * ucon[..., 1]) / ucon[..., 0],
                     (Bv[..., 2] + bt * ucon[..., 2]) / ucon[..., 0],
                     (Bv[..., 3] + bt * ucon[..., 3]) / ucon[..., 0]], axis=-1)
    bsq = np.einsum("...a,...ab,...b->...", bcon, gc, bcon)
    bsq_max = bsq[NG:-NG, NG:-NG].max()
    beta_act = (gam - 1.0) * umax / (0.5 * bsq_max)
    norm = math.sqrt(beta_act / beta)
    p[..., B1] = (B1v * norm)[:, :, None]
    p[..., B2] = (B2v * norm)[:, :, None]
    return p


def _ks_gcov_mks(grid, r, th):
    """Kerr-Schild g_{mu nu} (``src/metric.c:1168``) transformed to MKS."""
    a = grid.a
    cth, sth = np.cos(th), np.sin(th)
    s2 = sth * sth
    SS = r * r + a * a * cth * cth
    t9 = r / SS
    g = np.zeros(r.shape + (4, 4))
    g[..., 0, 0] = 2.0 * t9 - 1.0
    g[..., 0, 1] = g[..., 1, 0] = 2.0 * t9
    g[..., 0, 3] = g[..., 3, 0] = -2.0 * t9 * a * s2
    g[..., 1, 1] = 1.0 + 2.0 * t9
    g[..., 1, 3] = g[..., 3, 1] = -a * (1.0 + 2.0 * t9) * s2
    g[..., 2, 2] = SS
    g[..., 3, 3] = s2 * (r * r + a * a * (1.0 + 2.0 * t9 * s2))
    N2T = r.shape[1]
    xp2 = grid.startx[2] + (np.arange(N2T) + 0.5) * grid.dx[2]
    J = np.ones(r.shape + (4,))
    J[..., 1] = r - grid.R0
    J[..., 2] = theta_map(grid, xp2)[1][None, :]
    return g * J[..., :, None] * J[..., None, :]


def _ucon_from_prims(gcov, r, th, grid, v1, v2, v3):
    """u^mu from u~^i via the normal observer (``src/phys.c:216-241``)."""
    gcon = np.linalg.inv(gcov)
    qsq = (gcov[..., 1, 1] * v1 * v1 + gcov[..., 2, 2] * v2 * v2 + gcov[..., 3, 3] * v3 * v3
           + 2 * (gcov[..., 1, 2] * v1 * v2 + gcov[..., 1, 3] * v1 * v3 + gcov[..., 2, 3] * v2 * v3))
    gamma = np.sqrt(1.0 + np.maximum(qsq, 0.0))
    alpha = 1.0 / np.sqrt(-gcon[..., 0, 0])
    ut = gamma / alpha
    beta = [gcon[..., 0, i] * alpha * alpha for i in (1, 2, 3)]
    return np.stack([ut, v1 - ut * beta[0], v2 - ut * beta[1], v3 - ut * beta[2]], axis=-1)


# ---------------------------------------------------------------------------
# Circumbinary disk around an inspiralling black-hole binary

def bbh_grid(n1=300, n2=160, n3=400, sep=20.0, r_in=1.5, r_out=13.0, hslope_harm=0.3, m1=0.5, m2=0.5,
             inspiral=True, gam=5.0 / 3.0, cour=0.8, dt0=1.0e-3, t_end=1.0e6, coords="mks",
             diag3_h_slope=0.0729419909586, diag3_exponent=9):
    """Spherical grid about the binary's centre of mass: r in [r_in, r_out]
    separations (log-radial X1), full 2 pi in phi; M = m1 + m2 = 1 and the
    separation `sep` in units of M.  Theta map: coords "mks", the modified
    Kerr-Schild map with HARM's h = hslope_harm, or "diag3", the reference's
    binary-disk map (COORD_DIAGONAL3 with its h_slope and exponent,
    ``src/init.nz_pn_bbh_disk.c:332-360``)."""
    th_cutout = 1.0e-15 * math.pi
    th_beg, th_length = th_cutout, math.pi * (1.0 - 2.0e-15)
    if coords == "diag3":
        h_slope, factor = diag3_h_slope, 1.0 - 2 * th_cutout / math.pi - diag3_h_slope
    else:
        h_slope, factor = 0.5 * (1.0 - hslope_harm), 0.0
    Rin, Rout = r_in * sep, r_out * sep
    dx1, dx2, dx3 = math.log(Rout / Rin) / n1, 1.0 / n2, 2.0 * math.pi / n3
    startx = (0.0, math.log(Rin) - NG * dx1, -NG * dx2, -NG * dx3)
    return GridSpec(n1, n2, n3, startx, (dt0, dx1, dx2, dx3), "pn_binary", coords, gam=gam, cour=cour,
                    t_end=t_end, a=0.0, R0=0.0, h_slope=h_slope, th_beg=th_beg, th_length=th_length,
                    diag3_exponent=diag3_exponent, diag3_factor=factor,
                    extra=dict(Rin=Rin, Rout=Rout, sep=sep, m1=m1, m2=m2, inspiral=inspiral))


def bbh_gdet(grid: GridSpec, metric_cls, hostcheck=False, t=0.0, chunk=16):
    """sqrt(-g) of the binary metric at the zone centres [N1T][N2T][N3T], computed
    by the library (metric_cls: coolharm3d_amd.pn.PnMetric) in x1 slabs."""
    from .pn import bbh_block, CENT
    N1T, N2T, N3T = grid.n1 + 2 * NG, grid.n2 + 2 * NG, grid.n3 + 2 * NG
    out = np.empty((N1T, N2T, N3T))
    e = grid.extra
    for i0 in range(0, N1T, chunk):
        m = min(chunk, N1T - i0)
        c = bbh_block(m, N2T, N3T, sep=e["sep"], m1=e["m1"], m2=e["m2"], inspiral=e["inspiral"], ghosts=0,
                      h_slope=0.0)
        c.h_slope, c.th_beg, c.th_length = grid.h_slope, grid.th_beg, grid.th_length
        c.coords = 4 if grid.coords == "diag3" else 1
        c.diag3_exponent, c.diag3_factor = grid.diag3_exponent, grid.diag3_factor
        for l in range(1, 4):
            c.startx[l] = grid.startx[l] + (i0 * grid.dx[1] if l == 1 else 0.0)
            c.dx[l] = grid.dx[l]
        out[i0:i0 + m] = metric_cls(c, hostcheck=hostcheck).geometry(t, CENT)[..., 32]
    return out


def bbh_disk(grid: GridSpec, gdet, h_o_r=0.1, beta=100.0, noise=0.01, seed=2718, r_edge=2.0,
             rhomin=1.0e-7, uumin=1.0e-9, rhopow=-1.5, uupow=-2.5):
    """Circumbinary disk: rho ~ r^-1.5 with a Gaussian vertical profile of
    scale height H = h_o_r r, tapered inside r_edge separations; pressure of a
    disk of that thickness in vertical balance, p = rho (H/r)^2 M/r; Keplerian
    rotation u~^phi = sqrt(M/r^3); seeded +-noise relative density noise;
    poloidal loops A_phi ~ max(rho/rho_max - 0.2, 0) of the azimuthally averaged
    density, normalised to a minimum plasma beta; atmosphere floors
    rhomin r^rhopow, uumin r^uupow.  gdet: sqrt(-g) at the zone centres (the
    field is sqrt(-g) B^i = curl A, so its constrained-transport divergence is
    zero)."""
    N1T, N2T, N3T = grid.n1 + 2 * NG, grid.n2 + 2 * NG, grid.n3 + 2 * NG
    gam, M = grid.gam, 1.0
    sep = grid.extra["sep"]
    xp1 = grid.startx[1] + (np.arange(N1T) + 0.5) * grid.dx[1]
    xp2 = grid.startx[2] + (np.arange(N2T) + 0.5) * grid.dx[2]
    r1, th1 = mks_r_th(grid, xp1, xp2)
    r = r1[:, None] * np.ones((1, N2T))
    th = np.ones((N1T, 1)) * th1[None, :]
    z_over_r = np.cos(th)
    taper = 1.0 / (1.0 + np.exp(-(r - r_edge * sep) / (0.1 * sep)))
    rho = r ** -1.5 * np.exp(-0.5 * (z_over_r / h_o_r) ** 2) * taper
    rho /= rho[NG:-NG, NG:-NG].max()
    pres = rho * h_o_r ** 2 * M / r
    u = pres / (gam - 1.0)
    rng = np.random.default_rng(seed)
    pert = 1.0 + noise * (2.0 * rng.random((N1T, N2T, N3T)) - 1.0)
    p = np.zeros((N1T, N2T, N3T, NP))
    p[..., RHO] = rho[:, :, None] * pert
    p[..., UU] = u[:, :, None]
    p[..., U3] = (np.sqrt(M / r ** 3))[:, :, None]
    rhoflr = rhomin * np.power(r, rhopow)
    uuflr = uumin * np.power(r, uupow)
    p[..., RHO] = np.maximum(p[..., RHO], rhoflr[:, :, None])
    p[..., UU] = np.maximum(p[..., UU], uuflr[:, :, None])
    # loops from corner-averaged azimuthal-mean density
    A = np.zeros((N1T + 1, N2T + 1))
    ravg = 0.25 * (rho[1:, 1:] + rho[:-1, 1:] + rho[1:, :-1] + rho[:-1, :-1])
    A[1:-1, 1:-1] = np.maximum(ravg - 0.2, 0.0)
    g1 = -(A[:-1, :-1] - A[:-1, 1:] + A[1:, :-1] - A[1:, 1:]) / (2.0 * grid.dx[2])
    g2 = (A[:-1, :-1] + A[:-1, 1:] - A[1:, :-1] - A[1:, 1:]) / (2.0 * grid.dx[1])
    B1v = g1[:, :, None] / gdet
    B2v = g2[:, :, None] / gdet
    # b^2 ~ gamma_ij B^i B^j in the weak field of the grid (r >= 1.5 separations)
    dthdx2 = theta_map(grid, xp2)[1]
    bsq = (r[:, :, None] * B1v) ** 2 + (r[:, :, None] * dthdx2[None, :, None] * B2v) ** 2
    bsq_max = bsq[NG:-NG, NG:-NG, NG:-NG].max()
    pmax = (gam - 1.0) * p[NG:-NG, NG:-NG, NG:-NG, UU].max()
    norm = math.sqrt(pmax / (0.5 * bsq_max) / beta) if bsq_max > 0 else 0.0
    p[..., B1] = B1v * norm
    p[..., B2] = B2v * norm
    return p
