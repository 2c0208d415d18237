"""The named configurations at their full sizes on the GPU.

* Precision variant of the torus at 256x256x192, 100 steps: the mixed step
  (FP32 reconstruction, face fluxes and EMF averaging; FP64 update,
  con-to-prim, metric and cooling) against the FP64 step of the same library,
  density and internal energy relative L1 <= 1e-6 after 100 steps (the
  configuration's stated bar), both with the reference's confirming Newton
  iterations and without them (the separate no_confirm_iters option).
* The mixed step against the reference itself: the 24x24x16 torus golden
  (reference FP64 build) after 1, 10 and 100 steps, density and internal
  energy relative L1 <= 1e-6.
* The circumbinary disk on the inspiralling binary metric at 300x160x400,
  mixed precision, graph-replayed steps: after 12 steps the state is finite,
  div B stays at round-off, density and internal energy are positive and at
  or above the floors, the clock advanced with a positive dt, and the per-zone
  failure counters sum to the totals.
"""
import math

import numpy as np
import pytest

from helpers import golden, interior, rel_l1

pytestmark = pytest.mark.gpu


def _torus_run(g, p0, precision, steps, **kw):
    from coolharm3d_amd import Harm3D
    H = Harm3D(g, track_nfail=0, precision=precision, **kw)
    H.set_prims(p0)
    H.fixup_bounds()
    H.set_time(0.0, 1.0e-5)
    out = {}
    done = 0
    for s in steps:
        if s - done > 20:
            H.set_graph(True)
        H.step_ch(s - done)
        done = s
        out[s] = interior(H.prims())
    divb = H.divb_max()
    H.close()
    return out, divb


def test_precision_variant_256x256x192_100_steps():
    from coolharm3d_amd import torus_grid, torus
    g = torus_grid(256, 256, 192)
    p0 = torus(g)
    steps = (1, 10, 100)
    print("[fullsize] torus 256x256x192 fp64, 100 steps", flush=True)
    ref, _ = _torus_run(g, p0, "fp64", steps)
    for no_confirm in (False, True):
        print(f"[fullsize] torus 256x256x192 mixed (no_confirm_iters={no_confirm}), 100 steps", flush=True)
        got, divb = _torus_run(g, p0, "mixed", steps, no_confirm_iters=no_confirm)
        for s in steps:
            e = rel_l1(ref[s], got[s])
            print(f"[fullsize] mixed no_confirm={no_confirm} rel L1 after {s}: {np.array2string(e, precision=2)}",
                  flush=True)
            assert np.isfinite(got[s]).all()
        e100 = rel_l1(ref[100], got[100])
        assert e100[0] <= 1e-6 and e100[1] <= 1e-6, e100
        bmax = np.abs(got[100][..., 5:8]).max()
        assert divb <= 1e-11 * bmax / min(g.dx[1:])


@pytest.mark.parametrize("no_confirm", [False, True])
def test_mixed_against_reference_golden(no_confirm):
    """The mixed step against the reference build's own FP64 run."""
    from coolharm3d_amd import Harm3D, torus_grid
    G = golden("c1_torus24x24x16.npz")
    n = tuple(int(x) for x in G["n"])
    H = Harm3D(torus_grid(*n), precision="mixed", r_isco=float(G["r_isco"]), no_confirm_iters=no_confirm)
    H.set_prims(G["p0"])
    H.set_time(0.0, float(G["dt0"]))
    done = 0
    for s in (1, 10, 100):
        H.step_ch(s - done)
        done = s
        e = rel_l1(G[f"s{s}_prims"], interior(H.prims()))
        print(f"[mixed vs reference] no_confirm={no_confirm} after {s}: {np.array2string(e, precision=2)}",
              flush=True)
        assert e[0] <= 1e-6 and e[1] <= 1e-6, e


def test_bbh_disk_300x160x400_properties():
    from coolharm3d_amd import Harm3D, bbh_grid, bbh_disk
    from coolharm3d_amd.problems import bbh_gdet, mks_r_th
    from coolharm3d_amd.pn import PnMetric
    g = bbh_grid(300, 160, 400)
    print("[fullsize] binary disk 300x160x400: initial data", flush=True)
    p0 = bbh_disk(g, bbh_gdet(g, PnMetric))
    H = Harm3D(g, precision="mixed", track_nfail=1)
    H.set_prims(p0)
    del p0
    H.fixup_bounds()
    H.set_time(0.0, 1.0e-3)
    d0 = H.divb_max()
    H.reset_nfail()
    print("[fullsize] binary disk 300x160x400: 12 steps (graph replay after 2)", flush=True)
    H.step_ch(2)
    H.set_graph(True)
    H.step_ch(10)
    p = interior(H.prims())
    assert np.isfinite(p).all()
    bmax = np.abs(p[..., 5:8]).max()
    divb = H.divb_max()
    print(f"[fullsize] binary disk: max|divB| {d0:.3e} -> {divb:.3e}, max|B| {bmax:.3e}", flush=True)
    assert divb <= max(10 * d0, 1e-11 * bmax / min(g.dx[1:]))
    t, dt, _ = H.time()
    assert t > 0 and dt > 0
    # floors rhomin r^-1.5, uumin r^-2.5 at the zone centres (fixup_floor(),
    # src/fixup.c:110-121); zones filled by the interpolating fixups carry
    # averages of floored neighbours, so allow the spread of the floor over
    # one zone in r
    xp1 = g.startx[1] + (np.arange(g.n1) + 3 + 0.5) * g.dx[1]
    r = mks_r_th(g, xp1, np.array([0.5]))[0]
    rhoflr = 1.0e-7 * r ** -1.5
    uuflr = 1.0e-9 * r ** -2.5
    spread = math.exp(2.5 * g.dx[1]) * 1.0000001
    assert (p[..., 0] > 0).all() and (p[..., 1] > 0).all()
    assert (p[..., 0] * spread >= rhoflr[:, None, None]).all()
    assert (p[..., 1] * spread >= uuflr[:, None, None]).all()
    nf = H.nfail().reshape(15, -1).sum(axis=1)
    assert np.array_equal(nf, H.nfailtot())
    assert nf[4] > 0   # the inversion ran (USE_FULL_INV counts every zone)
    H.close()
