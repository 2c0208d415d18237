"""ctypes wrapper of the reference build oracle/_ref/<cfg>/libref.so (test infrastructure).

Only tests/, tests/golden/make_golden.py and bench.py's cpu_baseline leg use
this.  Each configuration is a separate shared library because the reference
is configured at compile time (src/options.h) and keeps its state in process
globals; one process can set up one grid per library (the reference allocates
its arrays once, src/decomp.c:140).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
N_NFAIL = 15


def ref_lib_path(cfg: str) -> str:
    return os.path.join(HERE, "_ref", cfg, "libref.so")


def have_ref(cfg: str) -> bool:
    return os.path.exists(ref_lib_path(cfg))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


_C2P_LIB = {}


def c2p_batch(a, gam, U, rth, prim, cfg="c1"):
    """The reference's Utoprim_2d_fast() over independent Kerr-Schild zones
    (ref_c2p_batch): U[8][n], rth[2][n], prim[8][n] guess in / result out.
    Returns the codes.  Uses the Kerr-Schild build (c1)."""
    if cfg not in _C2P_LIB:
        L = ctypes.CDLL(ref_lib_path(cfg), mode=os.RTLD_LOCAL | os.RTLD_NOW)
        L.ref_c2p_batch.argtypes = [ctypes.c_long, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _C2P_LIB[cfg] = L
    n = U.shape[1]
    for arr in (U, rth, prim):
        assert arr.dtype == np.float64 and arr.flags["C_CONTIGUOUS"]
    code = np.zeros(n, np.int32)
    _C2P_LIB[cfg].ref_c2p_batch(n, a, gam, _p(U), _p(rth), _p(prim), _p(code))
    return code


class RefRun:
    """One reference process state (load the .so under a private name so two
    configurations can coexist in one interpreter)."""

    def __init__(self, cfg: str, n1: int, n2: int, n3: int, params=None, restart=False):
        path = ref_lib_path(cfg)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        # dlopen with RTLD_LOCAL so each config keeps its own globals
        self.lib = ctypes.CDLL(path, mode=os.RTLD_LOCAL | os.RTLD_NOW)
        L = self.lib
        L.ref_set_time.argtypes = [ctypes.c_double] * 3
        L.ref_set_end_time.argtypes = [ctypes.c_double]
        if params is not None:
            arr = (ctypes.c_double * 64).in_dll(L, "ref_params")
            for i, v in enumerate(params):
                arr[i] = float(v)
        if L.ref_setup(n1, n2, n3) != 0:
            raise RuntimeError("reference already set up in this process")
        if restart:
            # the reference's restart_init() reads the checkpoint in the working directory
            if L.ref_restart_init() != 0:
                raise FileNotFoundError("no restart file in the working directory")
        else:
            L.ref_init()
        d = np.zeros(7, np.int32)
        L.ref_dims(_p(d))
        self.N = tuple(int(x) for x in d[:3])
        self.NT = tuple(int(x) for x in d[3:6])
        self.ncell = self.NT[0] * self.NT[1] * self.NT[2]

    # --- state -----------------------------------------------------------
    def scalars(self):
        s = np.zeros(20)
        self.lib.ref_get_scalars(_p(s))
        return dict(t=s[0], dt=s[1], cour=s[2], gam=s[3], dx1=s[4], dx2=s[5], dx3=s[6],
                    startx1=s[7], startx2=s[8], startx3=s[9], a=s[10], R0=s[11],
                    h_slope=s[12], th_length=s[13], th_beg=s[14], r_isco=s[15], M=s[16],
                    t_end_len=s[17], startx0=s[18], dt_old=s[19])

    def set_time(self, t, dt, cour):
        self.lib.ref_set_time(t, dt, cour)

    def set_end_time(self, tend):
        self.lib.ref_set_end_time(tend)

    def prims(self):
        a = np.zeros(self.ncell * 8)
        self.lib.ref_get_prims(_p(a))
        return a.reshape(self.NT + (8,))

    def ph(self):
        a = np.zeros(self.ncell * 8)
        self.lib.ref_get_ph(_p(a))
        return a.reshape(self.NT + (8,))

    def set_prims(self, p):
        a = np.ascontiguousarray(p, dtype=np.float64).reshape(-1)
        assert a.size == self.ncell * 8
        self.lib.ref_set_prims(_p(a))

    def pflag(self):
        a = np.zeros(self.ncell, np.int32)
        self.lib.ref_get_pflag(_p(a))
        return a.reshape(self.NT)

    def nfail(self):
        a = np.zeros(self.ncell * N_NFAIL, np.int32)
        self.lib.ref_get_nfail(_p(a))
        return a.reshape((N_NFAIL,) + self.NT)

    def reset_nfail(self):
        self.lib.ref_reset_nfail()

    def gamma(self):
        a = np.zeros(self.ncell)
        self.lib.ref_get_gamma(_p(a))
        return a.reshape(self.NT)

    def U(self, level):
        a = np.zeros(self.ncell * 8)
        self.lib.ref_get_U(level, _p(a))
        return a.reshape(self.NT + (8,))

    def F(self):
        a = np.zeros(self.ncell * 24)
        self.lib.ref_get_F(_p(a))
        return a.reshape(self.NT + (3, 8))

    def geom(self, i, j, k, pos):
        a = np.zeros(42)
        self.lib.ref_get_geom(i, j, k, pos, _p(a))
        return a

    def conn(self, i, j, k):
        a = np.zeros(64)
        self.lib.ref_get_conn(i, j, k, _p(a))
        return a.reshape(4, 4, 4)

    def coord(self, i, j, k, pos):
        a = np.zeros(11)
        self.lib.ref_get_coord(i, j, k, pos, _p(a))
        return a

    def restart_write(self):
        self.lib.ref_restart_write()

    def set_n_restart(self, n):
        self.lib.ref_set_n_restart(int(n))

    def nstep(self):
        return int(self.lib.ref_get_nstep())

    def fixup_bounds(self):
        self.lib.ref_fixup_bounds()

    def step(self, n=1):
        self.lib.ref_step(n)

    def check_for_crash(self):
        """The reference's check_for_crash(); exits the process on a crash."""
        self.lib.ref_check_for_crash()

    BRANCH_NAMES = ("Utoprim_2d_fast", "Utoprim_1d", "Utoprim_1d_ee", "Utoprim_1d_ee2",
                    "fixup_interp_v", "fixup_interp_prim")

    def reset_branch(self):
        self.lib.ref_reset_branch()

    def branch(self):
        """Calls and successful returns of the recovery tree's solvers and
        fixups since the last reset_branch() (ref_harness.c, linked with
        --wrap): {name: (calls, returned 0, {code: count})}."""
        calls = (ctypes.c_long * 12).in_dll(self.lib, "ref_branch")
        codes = (ctypes.c_long * (6 * 32 * 2)).in_dll(self.lib, "ref_codes")
        out = {}
        for s, name in enumerate(self.BRANCH_NAMES):
            hist = {}
            for c in range(32):
                code, count = codes[(s * 32 + c) * 2], codes[(s * 32 + c) * 2 + 1]
                if count:
                    hist[int(code)] = int(count)
            out[name] = (int(calls[2 * s]), int(calls[2 * s + 1]), hist)
        return out

    def diag(self):
        a = np.zeros(3)
        self.lib.ref_get_diag(_p(a))
        return dict(M_tot=a[0], failure_exists=a[1], nstep=a[2])


def torus_params(grid, dt0=1.0e-5):
    """ref_params[] layout of oracle/ref_init_kerr.c for a torus GridSpec."""
    ex = grid.extra
    return [grid.a, ex["Rin"], ex["Rout"], grid.h_slope, grid.gam, grid.cour,
            ex["x3_extent"], grid.t_end, dt0, grid.diag3_exponent]
