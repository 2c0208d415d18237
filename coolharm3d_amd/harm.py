"""Host-side mirror of coolharm3d's step interface for the GPU library.

:class:`Harm3D` owns one grid resident in HBM and exposes the reference's
per-step operations under the reference's names:

* ``step_ch()`` -- one full midpoint-RK2 step (``src/step_ch.c:50-234``);
* ``advance(stage)`` -- one ``advance()`` stage with boundaries and fixups;
* ``fixup_bounds()`` -- ``fixup(p); bounds(p,0)`` on initial data;
* ``pflag()``, ``nfail()``, ``nfailtot()`` -- the failure maps ``fail()`` writes
  (``src/diag.c:159-492``).

Primitive arrays cross the boundary in the reference's layout
``[i][j][k][var]`` over all zones including ``NG = 3`` ghosts.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from .problems import GridSpec, NG

RECON_MC, RECON_PPM_MC = 2, 5
FLUX_LF, FLUX_HLLE = 0, 1
BC_CARTESIAN_OUTFLOW, BC_SPHERICAL_OUTFLOW3 = 0, 10
N_NFAIL = 15
NFAIL_NAMES = ["FLOOR", "TMAX", "GAMMA_MAX", "GAMMA_CALC", "USE_FULL_INV", "USE_EE", "UTOPRIM",
               "UTOPRIM_EE", "USE_INTERP_PRIM", "USE_INTERP_V", "INTERP_PRIM", "INTERP_V",
               "CUTOUT_PRESSURE", "CUTOUT_GAMMA", "CUTOUT_BC"]


def risco_prograde(a: float) -> float:
    """risco_calc(1), src/metric.c:458-472, with asq = a^2."""
    t1, t2 = (1.0 + a) ** (1.0 / 3.0), (1.0 - a) ** (1.0 / 3.0)
    z1 = 1.0 + t1 * t2 * (t1 + t2)
    z2 = math.sqrt(3.0 * a * a + z1 * z1)
    return 3.0 + z2 - math.sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2))


def config_for(grid: GridSpec, **kw) -> _lib.HgpuConfig:
    """hgpu_config for a GridSpec.  Defaults follow the two problem families:
    Minkowski/identity -> MC, Toth CT, Cartesian outflow, no cooling;
    Kerr-Schild/MKS -> PPM+MC, PPM EMFs, spherical outflow, h/r cooling."""
    c = _lib.HgpuConfig()
    c.n1, c.n2, c.n3 = grid.n1, grid.n2, grid.n3
    bbh = grid.metric == "pn_binary"
    ks = grid.metric == "kerr_schild" or bbh
    c.metric = 3 if bbh else (2 if ks else 0)
    c.coords = {"identity": 0, "mks": 1, "diag3": 4}[grid.coords]
    c.diag3_exponent, c.diag3_factor = grid.diag3_exponent, grid.diag3_factor
    c.recon = kw.get("recon", RECON_PPM_MC if ks else RECON_MC)
    c.flux = kw.get("flux", FLUX_HLLE)
    c.ct_para = kw.get("ct_para", 1 if ks else 0)
    c.cooling = kw.get("cooling", 1 if ks else 0)
    c.rescale_r = kw.get("rescale_r", 1 if ks else 0)
    c.bc = kw.get("bc", BC_SPHERICAL_OUTFLOW3 if ks else BC_CARTESIAN_OUTFLOW)
    c.track_nfail = kw.get("track_nfail", 1)
    for l in range(4):
        c.startx[l] = grid.startx[l]
        c.dx[l] = grid.dx[l]
    c.gam, c.cour = grid.gam, grid.cour
    c.t_final = grid.t_end + grid.startx[0]
    c.a, c.R0, c.h_slope = grid.a, grid.R0, grid.h_slope
    c.th_beg, c.th_length = grid.th_beg, grid.th_length
    c.r_isco = kw.get("r_isco", risco_prograde(grid.a) if ks else 0.0)
    if bbh:
        e = grid.extra
        c.bbh_m1, c.bbh_m2, c.bbh_sep = e["m1"], e["m2"], e["sep"]
        c.bbh_phi0, c.bbh_inspiral = e.get("phi0", 0.0), 1 if e.get("inspiral", True) else 0
    c.h_o_r = kw.get("h_o_r", 0.1 if ks else 0.05)
    if ks:
        c.rhomin, c.uumin = kw.get("rhomin", 1.0e-7), kw.get("uumin", 1.0e-9)
    else:
        c.rhomin, c.uumin = kw.get("rhomin", 2.0e-10), kw.get("uumin", 2.0e-12)
    c.rhopower, c.uupower = kw.get("rhopower", -1.5), kw.get("uupower", -2.5)
    c.precision = {"fp64": 0, "mixed": 1}[kw.get("precision", "fp64")]
    c.bc_nbr = kw.get("bc_nbr", 0)
    c.no_confirm_iters = 1 if kw.get("no_confirm_iters", False) else 0
    for d, g in enumerate(kw.get("gpos", (0, 0, 0))):
        c.gpos[d] = g
    return c


def _dp(a):
    return a.ctypes.data_as(_lib.c_double_p)


def _ip(a):
    return a.ctypes.data_as(_lib.c_int_p)


class Harm3D:
    """One grid of the per-substep GRMHD advance resident on the GPU."""

    def __init__(self, grid: GridSpec, hostcheck: bool = False, **kw):
        self.grid = grid
        self.lib = _lib.load(hostcheck=hostcheck)
        self.cfg = config_for(grid, **kw)
        h = ctypes.c_void_p()
        _lib.check(self.lib, self.lib.hgpu_create(ctypes.byref(self.cfg), ctypes.byref(h)), "hgpu_create")
        self.h = h
        d = np.zeros(7, np.int32)
        self.lib.hgpu_dims(self.h, _ip(d))
        self.shape = (int(d[3]), int(d[4]), int(d[5]))
        self.nt = self.shape[0] * self.shape[1] * self.shape[2]
        self.nstep = 0

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.hgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        _lib.check(self.lib, rc, what)

    # -- state ---------------------------------------------------------------
    def set_prims(self, p, which=0):
        a = np.ascontiguousarray(p, dtype=np.float64)
        if a.shape != self.shape + (8,):
            raise ValueError(f"prims must have shape {self.shape + (8,)}, got {a.shape}")
        self._chk(self.lib.hgpu_upload_prims(self.h, which, _dp(a)), "hgpu_upload_prims")

    def prims(self, which=0):
        a = np.empty(self.shape + (8,))
        self._chk(self.lib.hgpu_download_prims(self.h, which, _dp(a)), "hgpu_download_prims")
        return a

    def set_time(self, t, dt):
        self._chk(self.lib.hgpu_set_time(self.h, t, dt), "hgpu_set_time")

    def time(self):
        t, dt, dto = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._chk(self.lib.hgpu_get_time(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(dto)),
                  "hgpu_get_time")
        return t.value, dt.value, dto.value

    def fixup_bounds(self):
        self._chk(self.lib.hgpu_fixup_bounds(self.h), "hgpu_fixup_bounds")

    def step_ch(self, nsteps=1):
        self._chk(self.lib.hgpu_step(self.h, nsteps), "hgpu_step")
        self.nstep += nsteps

    def set_graph(self, on=True):
        """Replay steps from a captured HIP graph (hgpu_set_graph)."""
        self._chk(self.lib.hgpu_set_graph(self.h, 1 if on else 0), "hgpu_set_graph")

    def set_metric_eval(self, in_kernels=True):
        """Binary metric: evaluate face and update records inside the kernels
        (True) or fill and read stored levels (False) -- hgpu_set_metric_eval."""
        self._chk(self.lib.hgpu_set_metric_eval(self.h, 1 if in_kernels else 0), "hgpu_set_metric_eval")

    def metric_eval(self):
        v = np.zeros(1, np.int32)
        self._chk(self.lib.hgpu_get_metric_eval(self.h, _ip(v)), "hgpu_get_metric_eval")
        return bool(v[0])

    def m_tot(self):
        """M_tot of the last stage: the sum of U[RHO] over the zones (hgpu_get_m_tot)."""
        v = ctypes.c_double()
        self._chk(self.lib.hgpu_get_m_tot(self.h, ctypes.byref(v)), "hgpu_get_m_tot")
        return float(v.value)

    def device_bytes(self):
        """Device memory the context holds (hgpu_device_bytes)."""
        v = ctypes.c_longlong()
        self._chk(self.lib.hgpu_device_bytes(self.h, ctypes.byref(v)), "hgpu_device_bytes")
        return int(v.value)

    FAMILIES = ("flux1", "flux2", "flux3", "emf", "conserved", "invert")

    def autotune(self, reps=2):
        """Times the launch-shape candidates of the stencil kernels on the
        current state and keeps the fastest (hgpu_autotune).  Returns
        {family: (bs, kc, flags, ms)}."""
        rep = np.zeros(20)
        self._chk(self.lib.hgpu_autotune(self.h, reps, _dp(rep)), "hgpu_autotune")
        return {f: (int(rep[4 * i]), int(rep[4 * i + 1]), int(rep[4 * i + 2]), float(rep[4 * i + 3]))
                for i, f in enumerate(self.FAMILIES[:5])}

    def set_launch_shape(self, family, bs=256, kc=0, flags=1):
        """flags bit 0: XCD-contiguous tile order; bit 1: x1-fastest tile order;
        bit 2: register-capped kernel variant; bit 3 (conserved update only):
        update and inversion fused in one kernel; bits 4-5 (face fluxes, PPM):
        0/1/2 for 1/2/4 faces per thread."""
        fam = self.FAMILIES.index(family) if isinstance(family, str) else int(family)
        self._chk(self.lib.hgpu_set_launch_shape(self.h, fam, bs, kc, flags), "hgpu_set_launch_shape")

    def launch_shape(self, family):
        fam = self.FAMILIES.index(family) if isinstance(family, str) else int(family)
        v = np.zeros(3, np.int32)
        self._chk(self.lib.hgpu_get_launch_shape(self.h, fam, _ip(v[0:1]), _ip(v[1:2]), _ip(v[2:3])),
                  "hgpu_get_launch_shape")
        return tuple(int(x) for x in v)

    def advance(self, stage):
        """One advance() stage with its boundaries and fixups (hgpu_substep)."""
        self._chk(self.lib.hgpu_substep(self.h, stage), "hgpu_substep")

    # -- a stage in the parts a decomposed run separates by its halo exchange
    def advance_only(self, stage):
        """advance() and the physical boundary conditions of bounds(prim,0)
        (hgpu_advance)."""
        self._chk(self.lib.hgpu_advance(self.h, stage), "hgpu_advance")

    def boundary_flags(self):
        """(failure_exists, boundary_phys_pflag, boundary_mpi_pflag) of the stage."""
        v = np.zeros(3, np.int32)
        self._chk(self.lib.hgpu_get_boundary_flags(self.h, _ip(v[0:1]), _ip(v[1:2]), _ip(v[2:3])),
                  "hgpu_get_boundary_flags")
        return int(v[0]), int(v[1]), int(v[2])

    def fixup_stage(self, stage, phys, mpi):
        """The stage's interpolating fixups and bounds(prim,1) with the given
        (globally reduced) boundary flags (hgpu_fixup_stage)."""
        self._chk(self.lib.hgpu_fixup_stage(self.h, stage, int(phys), int(mpi)), "hgpu_fixup_stage")

    def end_step(self, vmax4=None):
        """t += dt and the next dt from max_char_speed (given, or the block's own)."""
        if vmax4 is None:
            self._chk(self.lib.hgpu_end_step(self.h, None), "hgpu_end_step")
        else:
            v = np.ascontiguousarray(vmax4, np.float64)
            assert v.shape == (4,)
            self._chk(self.lib.hgpu_end_step(self.h, _dp(v)), "hgpu_end_step")
        self.nstep += 1

    def set_dt(self, dt):
        self._chk(self.lib.hgpu_set_dt(self.h, float(dt)), "hgpu_set_dt")

    def download_box(self, which, lo, hi):
        """(prims [ni][nj][nk][8], pflag [ni][nj][nk]) of the zones lo <= idx <= hi."""
        lo_a, hi_a = np.array(lo, np.int32), np.array(hi, np.int32)
        n = tuple(int(h - l + 1) for l, h in zip(lo, hi))
        pr = np.empty(n + (8,))
        pf = np.empty(n, np.int32)
        self._chk(self.lib.hgpu_download_box(self.h, which, _ip(lo_a), _ip(hi_a), _dp(pr), _ip(pf)),
                  "hgpu_download_box")
        return pr, pf

    def upload_box(self, which, lo, hi, prims, pflag):
        lo_a, hi_a = np.array(lo, np.int32), np.array(hi, np.int32)
        pr = np.ascontiguousarray(prims, np.float64)
        pf = np.ascontiguousarray(pflag, np.int32)
        self._chk(self.lib.hgpu_upload_box(self.h, which, _ip(lo_a), _ip(hi_a), _dp(pr), _ip(pf)),
                  "hgpu_upload_box")

    def sync(self):
        self._chk(self.lib.hgpu_sync(self.h), "hgpu_sync")

    # -- failure bookkeeping -------------------------------------------------
    def pflag(self):
        a = np.empty(self.shape, np.int32)
        self._chk(self.lib.hgpu_download_pflag(self.h, _ip(a)), "hgpu_download_pflag")
        return a

    def gamma(self):
        a = np.empty(self.shape)
        self._chk(self.lib.hgpu_download_gamma(self.h, _dp(a)), "hgpu_download_gamma")
        return a

    def nfail(self):
        a = np.empty((N_NFAIL,) + self.shape, np.int32)
        self._chk(self.lib.hgpu_download_nfail(self.h, _ip(a)), "hgpu_download_nfail")
        return a

    def nfailtot(self):
        a = np.zeros(N_NFAIL, np.int64)
        self._chk(self.lib.hgpu_get_nfailtot(self.h, a.ctypes.data_as(_lib.c_ll_p)), "hgpu_get_nfailtot")
        return a

    def reset_nfail(self):
        self._chk(self.lib.hgpu_reset_nfail(self.h), "hgpu_reset_nfail")

    def status(self):
        fe = ctypes.c_int()
        v = np.zeros(4)
        self._chk(self.lib.hgpu_get_status(self.h, ctypes.byref(fe), _dp(v)), "hgpu_get_status")
        return fe.value, v


    def crash(self):
        """(crash_step, n_bad) of the device's check_for_crash(): the first
        step (counted from 1) whose new failures exceed half of 2 N1 N2 N3,
        0 when none has."""
        st, nb = ctypes.c_longlong(), ctypes.c_longlong()
        self._chk(self.lib.hgpu_get_crash(self.h, ctypes.byref(st), ctypes.byref(nb)), "hgpu_get_crash")
        return int(st.value), int(nb.value)
    def fluxes(self):
        a = np.empty((3, 8) + self.shape)
        self._chk(self.lib.hgpu_download_fluxes(self.h, _dp(a)), "hgpu_download_fluxes")
        return a

    def divb_max(self):
        v = ctypes.c_double()
        self._chk(self.lib.hgpu_divb_max(self.h, ctypes.byref(v)), "hgpu_divb_max")
        return v.value

    # -- geometry --------------------------------------------------------------
    def geometry_counts(self):
        ng, nc = ctypes.c_longlong(), ctypes.c_longlong()
        self._chk(self.lib.hgpu_geometry_counts(self.h, ctypes.byref(ng), ctypes.byref(nc)), "counts")
        return ng.value, nc.value

    def geometry(self):
        """Geometry records [n_geom][42] and connection [n_conn][4][4][4]; for
        the binary metric the records of the state's level and the metric
        derivatives [28][n_conn] of the last stage's level."""
        ng, nc = self.geometry_counts()
        g = np.empty((ng, 42))
        c = np.empty((28, nc)) if self.cfg.metric == 3 else np.empty((nc, 4, 4, 4))
        self._chk(self.lib.hgpu_download_geometry(self.h, _dp(g), _dp(c)), "hgpu_download_geometry")
        return g, c

    def set_geometry(self, geom=None, conn=None):
        ng, nc = self.geometry_counts()
        gp = cp = None
        if geom is not None:
            geom = np.ascontiguousarray(geom, np.float64)
            assert geom.size == ng * 42
            gp = _dp(geom)
        if conn is not None:
            conn = np.ascontiguousarray(conn, np.float64)
            assert conn.size == nc * 64
            cp = _dp(conn)
        self._chk(self.lib.hgpu_upload_geometry(self.h, gp, cp), "hgpu_upload_geometry")

    # -- timing ----------------------------------------------------------------
    def profile_enable(self, on=True):
        self._chk(self.lib.hgpu_profile_enable(self.h, 1 if on else 0), "hgpu_profile_enable")

    def profile_read(self, name):
        ms, n = ctypes.c_double(), ctypes.c_longlong()
        self._chk(self.lib.hgpu_profile_read(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(n)),
                  "hgpu_profile_read")
        return ms.value, n.value
