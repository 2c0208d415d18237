"""ctypes wrapper of the CPU restatement oracle/_build/liboracle.so (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module; the GPU library never does.  :class:`OracleRun` mirrors the
per-step interface of the reference (step_ch(), fixup(); bounds(), the
pflag/nfail maps of fail()) on one grid held in host memory in the reference's
layout ``[i][j][k][var]`` with ``NG = 3`` ghost zones.

The configuration is derived from a :class:`coolharm3d_amd.problems.GridSpec`
with the same defaults the two problem families use in the reference's
options (src/options.h): Minkowski/identity -> MC, Toth CT, Cartesian outflow;
Kerr-Schild/MKS -> PPM+MC, PPM EMFs, spherical outflow, h/r cooling, radial
rescaling.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
N_NFAIL = 15
NG = 3


class OrcConfig(ctypes.Structure):
    _fields_ = [("n1", ctypes.c_int), ("n2", ctypes.c_int), ("n3", ctypes.c_int),
                ("metric", ctypes.c_int), ("coords", ctypes.c_int), ("recon", ctypes.c_int),
                ("flux", ctypes.c_int), ("ct_para", ctypes.c_int), ("cooling", ctypes.c_int),
                ("rescale_r", ctypes.c_int), ("bc", ctypes.c_int),
                ("startx", ctypes.c_double * 4), ("dx", ctypes.c_double * 4),
                ("gam", ctypes.c_double), ("cour", ctypes.c_double), ("t_final", ctypes.c_double),
                ("a", ctypes.c_double), ("R0", ctypes.c_double), ("h_slope", ctypes.c_double),
                ("th_beg", ctypes.c_double), ("th_length", ctypes.c_double), ("r_isco", ctypes.c_double),
                ("h_o_r", ctypes.c_double), ("rhomin", ctypes.c_double), ("uumin", ctypes.c_double),
                ("rhopower", ctypes.c_double), ("uupower", ctypes.c_double),
                ("diag3_exponent", ctypes.c_int), ("diag3_factor", ctypes.c_double)]


_LIB = None


def have_oracle() -> bool:
    return os.path.exists(LIB_PATH)


def build() -> None:
    subprocess.run(["make", "-C", HERE, "oracle"], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _LIB
    if _LIB is None:
        if not have_oracle():
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(OrcConfig)]
        for nm in ("orc_destroy", "orc_fixup_bounds", "orc_reset_nfail"):
            getattr(L, nm).argtypes = [ctypes.c_void_p]
        L.orc_dims.argtypes = [ctypes.c_void_p, ip]
        L.orc_set_prims.argtypes = [ctypes.c_void_p, dp]
        L.orc_get_prims.argtypes = [ctypes.c_void_p, dp]
        L.orc_set_time.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.orc_get_time.argtypes = [ctypes.c_void_p, dp]
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_get_pflag.argtypes = [ctypes.c_void_p, ip]
        L.orc_get_gamma.argtypes = [ctypes.c_void_p, dp]
        L.orc_get_nfail.argtypes = [ctypes.c_void_p, ip]
        L.orc_get_nfailtot.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)]
        L.orc_get_vmax.argtypes = [ctypes.c_void_p, dp]
        L.orc_get_geom.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.orc_get_conn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, dp]
        L.orc_divb_max.argtypes = [ctypes.c_void_p]
        L.orc_divb_max.restype = ctypes.c_double
        L.orc_prim_to_flux.argtypes = [dp, ctypes.c_double, dp, ctypes.c_int, dp]
        L.orc_vchar.argtypes = [dp, ctypes.c_double, dp, ctypes.c_int, dp, dp]
        L.orc_utoprim_2d.argtypes = [dp, ctypes.c_double, dp, dp, dp, dp, ip]
        L.orc_utoprim_1d.argtypes = [dp, ctypes.c_double, dp, dp, dp, dp]
        L.orc_face_flux.argtypes = [dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp, dp, dp, dp]
        L.orc_recon_line.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp]
        L.orc_ks_geom_rth.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def risco_prograde(a: float) -> float:
    """risco_calc(1), src/metric.c:458-472, with asq = a^2."""
    t1, t2 = (1.0 + a) ** (1.0 / 3.0), (1.0 - a) ** (1.0 / 3.0)
    z1 = 1.0 + t1 * t2 * (t1 + t2)
    z2 = math.sqrt(3.0 * a * a + z1 * z1)
    return 3.0 + z2 - math.sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2))


def config_for(grid, **kw) -> OrcConfig:
    c = OrcConfig()
    c.n1, c.n2, c.n3 = grid.n1, grid.n2, grid.n3
    ks = grid.metric == "kerr_schild"
    c.metric = 2 if ks else 0
    c.coords = {"identity": 0, "mks": 1, "diag3": 4}[grid.coords]
    c.diag3_exponent, c.diag3_factor = grid.diag3_exponent, grid.diag3_factor
    c.recon = kw.get("recon", 5 if ks else 2)
    c.flux = kw.get("flux", 1)
    c.ct_para = kw.get("ct_para", 1 if ks else 0)
    c.cooling = kw.get("cooling", 1 if ks else 0)
    c.rescale_r = kw.get("rescale_r", 1 if ks else 0)
    c.bc = kw.get("bc", 10 if ks else 0)
    for l in range(4):
        c.startx[l] = grid.startx[l]
        c.dx[l] = grid.dx[l]
    c.gam, c.cour = grid.gam, grid.cour
    c.t_final = grid.t_end + grid.startx[0]
    c.a, c.R0, c.h_slope = grid.a, grid.R0, grid.h_slope
    c.th_beg, c.th_length = grid.th_beg, grid.th_length
    c.r_isco = kw.get("r_isco", risco_prograde(grid.a) if ks else 0.0)
    c.h_o_r = kw.get("h_o_r", 0.1 if ks else 0.05)
    if ks:
        c.rhomin, c.uumin = kw.get("rhomin", 1.0e-7), kw.get("uumin", 1.0e-9)
    else:
        c.rhomin, c.uumin = kw.get("rhomin", 2.0e-10), kw.get("uumin", 2.0e-12)
    c.rhopower, c.uupower = kw.get("rhopower", -1.5), kw.get("uupower", -2.5)
    return c


class OracleRun:
    """One grid advanced by the CPU restatement."""

    def __init__(self, grid, **kw):
        self.L = lib()
        self.cfg = config_for(grid, **kw)
        self.h = self.L.orc_create(ctypes.byref(self.cfg))
        if not self.h:
            raise ValueError("oracle rejected the configuration")
        d = np.zeros(6, np.int32)
        self.L.orc_dims(self.h, _ip(d))
        self.shape = (int(d[3]), int(d[4]), int(d[5]))
        self.nt = self.shape[0] * self.shape[1] * self.shape[2]

    def close(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_prims(self, p):
        a = np.ascontiguousarray(p, dtype=np.float64)
        assert a.shape == self.shape + (8,)
        self.L.orc_set_prims(self.h, _dp(a))

    def prims(self):
        a = np.empty(self.shape + (8,))
        self.L.orc_get_prims(self.h, _dp(a))
        return a

    def set_time(self, t, dt):
        self.L.orc_set_time(self.h, t, dt)

    def time(self):
        a = np.zeros(4)
        self.L.orc_get_time(self.h, _dp(a))
        return dict(t=a[0], dt=a[1], t_old=a[2], dt_old=a[3])

    def fixup_bounds(self):
        self.L.orc_fixup_bounds(self.h)

    def step(self, n=1):
        self.L.orc_step(self.h, n)

    def pflag(self):
        a = np.empty(self.shape, np.int32)
        self.L.orc_get_pflag(self.h, _ip(a))
        return a

    def gamma(self):
        a = np.empty(self.shape)
        self.L.orc_get_gamma(self.h, _dp(a))
        return a

    def nfail(self):
        a = np.empty((N_NFAIL,) + self.shape, np.int32)
        self.L.orc_get_nfail(self.h, _ip(a))
        return a

    def nfailtot(self):
        a = np.zeros(N_NFAIL, np.int64)
        self.L.orc_get_nfailtot(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
        return a

    def reset_nfail(self):
        self.L.orc_reset_nfail(self.h)

    def vmax(self):
        a = np.zeros(4)
        self.L.orc_get_vmax(self.h, _dp(a))
        return a

    def geom(self, i, j, pos):
        a = np.zeros(42)
        self.L.orc_get_geom(self.h, i, j, pos, _dp(a))
        return a

    def conn(self, i, j):
        a = np.zeros(64)
        self.L.orc_get_conn(self.h, i, j, _dp(a))
        return a.reshape(4, 4, 4)

    def divb_max(self):
        return self.L.orc_divb_max(self.h)


# ---- per-zone functions -----------------------------------------------------
def prim_to_flux(geom42, gam, prim, direction):
    out = np.zeros(8)
    lib().orc_prim_to_flux(_dp(np.ascontiguousarray(geom42, np.float64)), gam,
                           _dp(np.ascontiguousarray(prim, np.float64)), direction, _dp(out))
    return out


def vchar(geom42, gam, prim, direction):
    cmax, cmin = np.zeros(1), np.zeros(1)
    lib().orc_vchar(_dp(np.ascontiguousarray(geom42, np.float64)), gam,
                    _dp(np.ascontiguousarray(prim, np.float64)), direction, _dp(cmax), _dp(cmin))
    return cmax[0], cmin[0]


def utoprim_2d(geom42, gam, U, guess):
    prim = np.array(guess, np.float64)
    gamma, bsq = np.zeros(1), np.zeros(1)
    it = np.zeros(1, np.int32)
    ret = lib().orc_utoprim_2d(_dp(np.ascontiguousarray(geom42, np.float64)), gam,
                               _dp(np.ascontiguousarray(U, np.float64)), _dp(prim), _dp(gamma), _dp(bsq), _ip(it))
    return ret, prim, gamma[0], bsq[0], int(it[0])


def ks_geom_rth(a, r, th):
    """Kerr-Schild record (struct of_geom layout, 42 doubles) at (r, th)."""
    out = np.zeros(42)
    lib().orc_ks_geom_rth(a, r, th, _dp(out))
    return out


def utoprim_1d(geom42, gam, U, guess):
    prim = np.array(guess, np.float64)
    gamma, bsq = np.zeros(1), np.zeros(1)
    ret = lib().orc_utoprim_1d(_dp(np.ascontiguousarray(geom42, np.float64)), gam,
                               _dp(np.ascontiguousarray(U, np.float64)), _dp(prim), _dp(gamma), _dp(bsq))
    return ret, prim, gamma[0], bsq[0]


def face_flux(geom42, gam, flux, direction, pl, pr):
    out, ctop = np.zeros(8), np.zeros(1)
    lib().orc_face_flux(_dp(np.ascontiguousarray(geom42, np.float64)), gam, flux, direction,
                        _dp(np.ascontiguousarray(pl, np.float64)), _dp(np.ascontiguousarray(pr, np.float64)),
                        _dp(out), _dp(ctop))
    return out, ctop[0]


def recon_line(recon, y):
    y = np.ascontiguousarray(y, np.float64)
    left, right = np.full(y.size, np.nan), np.full(y.size, np.nan)
    lib().orc_recon_line(recon, y.size, _dp(y), _dp(left), _dp(right))
    return left, right
