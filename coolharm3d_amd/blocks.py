"""A decomposed grid on the host side of the C ABI: one context per block,
stepped the way a domain-decomposed step_ch() runs (src/step_ch.c:50-234 with
USEMPI), with the halo exchange done through host copies.

The reference cuts the global grid into ncpux[1] x ncpux[2] x ncpux[3] blocks
of equal size, one per MPI rank (src/harm_mpi.c:1445-1600), and inside
bounds(prim, 0) exchanges the NG ghost layers of every face that borders
another block, direction after direction so that edges and corners arrive
through the faces (bounds_mpi(), bounds_pflag_mpi(), src/harm_mpi.c:1631-1744;
the send/receive ranges of src/harm_mpi.c:470-560: NG layers along the face,
every zone across it).  The boundary flags are reduced over the ranks
(mpi_global_imax2(), src/bounds.c:2164-2166), the wave speeds and the time
step too (mpi_global_vmax(), mpi_global_min(), src/step_ch.c:561, :618).

:class:`BlockGrid` does exactly that with hgpu_advance(), hgpu_download_box(),
hgpu_upload_box(), hgpu_get_boundary_flags(), hgpu_fixup_stage() and
hgpu_end_step(): what a rank's step_ch() does between its calls into the
library is what the reference's MPI layer does between the same stages.
Topology as the reference's (src/harm_mpi.c:1560-1600): outer faces physical,
except the phi direction of a spherical grid split along phi, whose outer
faces wrap around to the other end.
"""
from __future__ import annotations

import itertools

import numpy as np

from .harm import Harm3D, BC_SPHERICAL_OUTFLOW3, config_for
from .problems import GridSpec, NG


class BlockGrid:
    """A GridSpec cut into ncpux = (c1, c2, c3) equal blocks, each a Harm3D."""

    def __init__(self, grid: GridSpec, ncpux, hostcheck: bool = False, **kw):
        self.grid = grid
        self.ncpux = tuple(int(c) for c in ncpux)
        N = (grid.n1, grid.n2, grid.n3)
        for n, c in zip(N, self.ncpux):
            if c < 1 or n % c:
                raise ValueError(f"{n} zones do not split into {c} equal blocks")
        self.nb = tuple(n // c for n, c in zip(N, self.ncpux))
        bc = config_for(grid, **kw).bc
        self.periodic3 = (bc == BC_SPHERICAL_OUTFLOW3) and self.ncpux[2] > 1
        sub = GridSpec(**{**grid.__dict__, "n1": self.nb[0], "n2": self.nb[1], "n3": self.nb[2]})
        self.blocks = {}
        for pos in itertools.product(*(range(c) for c in self.ncpux)):
            nbr = 0
            for d in range(3):
                for side in (0, 1):
                    if self.neighbour(pos, d, side) is not None:
                        nbr |= 1 << (2 * d + side)
            gpos = tuple(p * n for p, n in zip(pos, self.nb))
            self.blocks[pos] = Harm3D(sub, hostcheck=hostcheck, bc_nbr=nbr, gpos=gpos, **kw)

    # -- topology ------------------------------------------------------------
    def neighbour(self, pos, d, side):
        """Block across face (d, side) of block pos, or None for a physical face."""
        q = list(pos)
        q[d] += 1 if side else -1
        if 0 <= q[d] < self.ncpux[d]:
            return tuple(q)
        if d == 2 and self.periodic3:
            q[d] %= self.ncpux[d]
            return tuple(q)
        return None

    def _slab(self, d, side, recv):
        """(lo, hi) of the NG layers a block sends through face (d, side)
        (recv = False) or receives into its ghosts (recv = True)."""
        n = self.nb[d]
        lo = [0, 0, 0]
        hi = [self.nb[e] + 2 * NG - 1 for e in range(3)]
        if recv:
            lo[d], hi[d] = (0, NG - 1) if side == 0 else (n + NG, n + 2 * NG - 1)
        else:
            lo[d], hi[d] = (NG, 2 * NG - 1) if side == 0 else (n, n + NG - 1)
        return lo, hi

    def exchange(self, which):
        """bounds_mpi() + bounds_pflag_mpi() of state which (0 p, 1 ph): per
        direction, every block's send slabs first, then every receive."""
        for d in range(3):
            sent = {}
            for pos, B in self.blocks.items():
                for side in (0, 1):
                    if self.neighbour(pos, d, side) is not None:
                        sent[(pos, side)] = B.download_box(which, *self._slab(d, side, recv=False))
            for pos, B in self.blocks.items():
                for side in (0, 1):
                    q = self.neighbour(pos, d, side)
                    if q is not None:
                        pr, pf = sent[(q, 1 - side)]
                        B.upload_box(which, *self._slab(d, side, recv=True), pr, pf)

    # -- state -----------------------------------------------------------------
    def _split(self, pos):
        return tuple(slice(p * n, p * n + n + 2 * NG) for p, n in zip(pos, self.nb))

    def set_prims(self, p):
        """Global state [N1TOT][N2TOT][N3TOT][8]; every block gets its part with ghosts."""
        for pos, B in self.blocks.items():
            B.set_prims(np.ascontiguousarray(p[self._split(pos)]))

    def _gather(self, get, shape_tail, dtype):
        N = tuple(n * c for n, c in zip(self.nb, self.ncpux))
        out = np.zeros(tuple(n + 2 * NG for n in N) + shape_tail, dtype)
        for pos, B in self.blocks.items():
            a = get(B)
            s = tuple(slice(NG + p * n, NG + p * n + n) for p, n in zip(pos, self.nb))
            out[s] = a[NG:-NG, NG:-NG, NG:-NG]
        return out

    def prims(self):
        """Global state; interior zones only (ghosts are zero)."""
        return self._gather(lambda B: B.prims(), (8,), np.float64)

    def pflag(self):
        return self._gather(lambda B: B.pflag(), (), np.int32)

    def nfail(self):
        """nfail[15][...] of the interior zones."""
        a = self._gather(lambda B: np.moveaxis(B.nfail(), 0, -1), (15,), np.int32)
        return np.moveaxis(a, -1, 0)

    def nfailtot(self):
        return sum(B.nfailtot() for B in self.blocks.values())

    def set_time(self, t, dt):
        for B in self.blocks.values():
            B.set_time(t, dt)

    def time(self):
        ts = {B.time() for B in self.blocks.values()}
        if len(ts) != 1:
            raise RuntimeError(f"blocks disagree on the clock: {sorted(ts)}")
        return ts.pop()

    # -- stepping ----------------------------------------------------------------
    def step_ch(self, nsteps=1):
        for _ in range(nsteps):
            for stage in (0, 1):
                which = 1 if stage == 0 else 0
                for B in self.blocks.values():
                    B.advance_only(stage)
                self.exchange(which)                      # bounds(prim,0): MPI part
                flags = [B.boundary_flags() for B in self.blocks.values()]
                phys = max(f[1] for f in flags)           # mpi_global_imax2()
                mpi = max(f[2] for f in flags)
                for B in self.blocks.values():
                    B.fixup_stage(stage, phys, mpi)
                if mpi:
                    self.exchange(which)                  # bounds(prim,1): MPI part
            vmax = np.max([B.status()[1] for B in self.blocks.values()], axis=0)   # mpi_global_vmax()
            for B in self.blocks.values():
                B.end_step(vmax)
            dts = [B.time()[1] for B in self.blocks.values()]
            if min(dts) != max(dts):                      # mpi_global_min()
                for B in self.blocks.values():
                    B.set_dt(min(dts))
