"""Time-dependent metric of an inspiralling black-hole binary (the
circumbinary-disk configuration): Python side of ``hgpu_pn_*``.

The reference declares its post-Newtonian binary metrics
(``dyn_35pn_nz_gcov_func``, src/metric.h:104-171) without bodies; the library
evaluates the 1PN near-zone metric of two point masses in harmonic coordinates
(Blanchet, Faye & Ponsot 1998) on a quasi-circular orbit decaying by
leading-order radiation reaction (Peters 1964), in spherical numerical
coordinates about the centre of mass.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from .c2p import DevArray

FACE1, FACE2, FACE3, CENT, CORN = range(5)


def bbh_block(n1=300, n2=160, n3=400, sep=20.0, r_in_over_sep=1.5, r_out_over_sep=13.0, h_slope=0.3,
              ghosts=3, m1=0.5, m2=0.5, phi0=0.0, inspiral=True, coords="mks", diag3_exponent=9):
    """Configuration of the circumbinary block: r in [1.5, 13] separations,
    log-radial X1, full 2 pi in phi; the block includes `ghosts` ghost layers
    on every side.  Theta map: "mks", the modified Kerr-Schild map with sin
    coefficient h_slope, or "diag3", COORD_DIAGONAL3 with h_slope its linear
    coefficient and the reference's exponent and factor 1 - h_slope."""
    c = _lib.HgpuPnConfig()
    c.m1, c.m2, c.b0, c.phi0, c.inspiral = m1, m2, sep, phi0, 1 if inspiral else 0
    c.R0, c.th_beg, c.th_length, c.h_slope = 0.0, 0.0, math.pi, h_slope
    c.coords = 4 if coords == "diag3" else 1
    c.diag3_exponent, c.diag3_factor = diag3_exponent, (1.0 - h_slope) if coords == "diag3" else 0.0
    dx1 = (math.log(r_out_over_sep * sep) - math.log(r_in_over_sep * sep)) / n1
    dx = [0.0, dx1, 1.0 / n2, 2 * math.pi / n3]
    start = [0.0, math.log(r_in_over_sep * sep), 0.0, 0.0]
    c.n1, c.n2, c.n3 = n1 + 2 * ghosts, n2 + 2 * ghosts, n3 + 2 * ghosts
    for l in range(4):
        c.startx[l] = start[l] - (ghosts * dx[l] if l > 0 else 0.0)
        c.dx[l] = dx[l]
    return c


class PnMetric:
    def __init__(self, cfg, hostcheck: bool = False):
        self.cfg = cfg
        self.host = hostcheck
        self.lib = _lib.load(hostcheck=hostcheck)
        if not hostcheck:
            n = ctypes.c_int()
            _lib.check(self.lib, self.lib.hgpu_dev_count(ctypes.byref(n)), "hgpu_dev_count")
            if n.value < 1:
                raise _lib.HgpuError("no HIP device available (the hgpu library runs only on a GPU)")
        self.n = cfg.n1 * cfg.n2 * cfg.n3

    def geometry_dev(self, t, pos, out=None):
        out = out if out is not None else DevArray(self.lib, (self.n, 42), np.float64)
        _lib.check(self.lib, self.lib.hgpu_pn_geometry(None, ctypes.byref(self.cfg), t, pos, out.ptr),
                   "hgpu_pn_geometry")
        return out

    def connection_dev(self, t, out=None):
        out = out if out is not None else DevArray(self.lib, (self.n, 64), np.float64)
        _lib.check(self.lib, self.lib.hgpu_pn_connection(None, ctypes.byref(self.cfg), t, out.ptr),
                   "hgpu_pn_connection")
        return out

    def geometry(self, t, pos=CENT):
        """numpy [n1][n2][n3][42] records."""
        return self.geometry_dev(t, pos).numpy().reshape(self.cfg.n1, self.cfg.n2, self.cfg.n3, 42)

    def connection(self, t):
        """numpy [n1][n2][n3][4][4][4], Gamma^d_{ab} at [..][d][a][b]."""
        return self.connection_dev(t).numpy().reshape(self.cfg.n1, self.cfg.n2, self.cfg.n3, 4, 4, 4)
