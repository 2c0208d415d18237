"""Writes the golden fixtures of tests/golden/ from the reference build.

Runs the reference C (oracle/_ref/<cfg>/libref.so, built from the reference
tree by oracle/Makefile) on identical initial data and records the primitive
state, the pflag map and the nfail counters after selected steps.  The
reference has no test fixtures of its own for this path, so these vectors are
the reference's outputs on the configurations the parity tests use:

* c0_balsara1024.npz  -- Balsara (2001) test 1, 1024 zones, Minkowski, MC, HLL,
  Toth CT, cour 0.4 (src/init.balsara1.c); state after 1, 10, 100 steps.
* c1_torus24x24x16.npz -- Fishbone-Moncrief torus, Kerr a = 0.9 in MKS
  coordinates, PPM, HLL, PPM EMFs, h/r cooling, floors 1e-7 r^-1.5 and
  1e-9 r^-2.5 (coolharm3d_amd.problems.torus, seed 1234); after 1, 10, 100 steps.
* c1d3_torus24x24x16.npz -- the same torus on the theta map of the
  reference's binary-disk set-up, COORD_DIAGONAL3 (src/coord.c:270-276;
  h_slope 0.0729419909586, exponent 9, src/init.nz_pn_bbh_disk.c:332-360);
  after 1, 10, 100 steps.
* c0c_balsara1024_crash.npz -- the driver's crash test, check_for_crash()
  (src/step_ch.c:1244-1339) after every step: the shock tube with rho < 0 in
  the first 600, 520, 510 zones; the step at which the reference ends the
  run and its n_bad (0 when it survives 5 steps).
* c0_balsara1024_t04.npz -- the shock tube run as the reference's driver runs
  it, step_ch() while t < GridLength[0] = 0.4 (src/main.c:62-81,
  src/init.balsara1.c:56): the number of steps, the state, pflag, nfail, t
  and dt at the end.
* c0f_balsara1024_failures.npz, c1f_torus24x24x16_failures.npz -- the same two
  set-ups with a few zones of the initial state overwritten (FAILURES_C0,
  FAILURES_C1: negative densities and internal energies, Lorentz factors of
  ~200 and ~1e4, singly and in 3x3x3 clusters) so that step one runs every
  branch of recover_primitives() (src/step_ch.c:801-1024), the entropy fix
  with both of its solvers (src/fixup.c:313-454) and both interpolating fixups
  with and without enough good neighbours (src/fixup.c:499-761).  Besides the
  state, pflag and nfail after 1, 2, 10 steps, these record the reference's
  branch counters of step one (ref_harness.c: calls and successful returns of
  Utoprim_2d_fast, Utoprim_1d, Utoprim_1d_ee, Utoprim_1d_ee2, the fixups'
  calls, and each solver's return codes).

Usage:  python tests/golden/make_golden.py   (needs oracle/_ref built: make -C oracle ref)
Each configuration runs in its own process (the reference keeps its grid in
process globals).
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

STEPS = (1, 10, 100)


def _interior(a):
    return a[3:-3, 3:-3, 3:-3]


def record(R, out, tag):
    out[f"{tag}_prims"] = _interior(R.prims())
    out[f"{tag}_pflag"] = _interior(R.pflag()).astype(np.int8)
    nf = R.nfail()
    out[f"{tag}_nfail"] = nf.astype(np.int32)
    s = R.scalars()
    out[f"{tag}_t"] = np.float64(s["t"])
    out[f"{tag}_dt"] = np.float64(s["dt"])


# Overwritten zones (i, j, k including ghosts), applied to the reference's
# initial state after its init-time fixup and boundaries.  Kinds:
#   "rho" / "u": set the density / internal energy to the value;
#   "ut": scale the velocity u~^i (coordinate components) to Euclidean length
#         value, keeping its direction (x1 where it is zero);
# a kind ending in "3" applies to the 3x3x3 block centred on the zone.  Every
# zone is at least NG zones from the grid edges, so the ghosts stay valid.
FAILURES_C0 = [("ut", 200, 3, 3, 200.0), ("ut", 300, 3, 3, 1.0e5), ("rho", 400, 3, 3, -1.0e-3),
               ("u", 500, 3, 3, -1.0e-3), ("ut", 699, 3, 3, 1.0e5), ("ut", 700, 3, 3, 1.0e5),
               ("ut", 701, 3, 3, 1.0e5), ("rho", 799, 3, 3, -1.0e-3), ("rho", 800, 3, 3, -1.0e-3),
               ("rho", 801, 3, 3, -1.0e-3), ("ut", 899, 3, 3, 200.0), ("ut", 900, 3, 3, 200.0),
               ("ut", 901, 3, 3, 200.0)]
FAILURES_C1 = [("ut3", 14, 14, 8, 1.0e4), ("rho3", 20, 20, 12, -1.0e-3), ("ut", 12, 12, 15, 20.0),
               ("ut", 18, 10, 14, 1.0e4), ("u", 8, 8, 12, -1.0e-3), ("u", 6, 14, 8, 1.0),
               ("rho", 22, 12, 6, -1.0e-3)]
FAIL_STEPS = (1, 2, 10)


def corrupt(p, spec):
    """Copy of the primitive state p[i][j][k][8] with the zones of spec overwritten."""
    p = np.array(p, dtype=np.float64, copy=True)
    zones = []
    for kind, i, j, k, val in spec:
        if kind.endswith("3"):
            zones += [(kind[:-1], i + a, j + b, k + c, val) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
        else:
            zones.append((kind, i, j, k, val))
    for kind, i, j, k, val in zones:
        for idx, n in zip((i, j, k), p.shape[:3]):
            assert n == 7 or 3 <= idx < n - 3, "overwritten zones must be interior"
        z = p[i, j, k]
        if kind == "rho":
            z[0] = val
        elif kind == "u":
            z[1] = val
        elif kind == "ut":
            v = z[2:5].copy()
            n = np.sqrt((v * v).sum())
            z[2:5] = (v / n if n > 0 else np.array([1.0, 0.0, 0.0])) * val
        else:
            raise ValueError(kind)
    return p


def record_branch(R, out, tag):
    """The reference's branch counters since the last reset, flattened:
    {tag}_branch[6][2] (calls, returned 0) and {tag}_codes[6][32][2]."""
    b = R.branch()
    calls = np.zeros((6, 2), np.int64)
    codes = np.zeros((6, 32, 2), np.int64)
    for s, name in enumerate(R.BRANCH_NAMES):
        n, ok, hist = b[name]
        calls[s] = (n, ok)
        for c, (code, count) in enumerate(sorted(hist.items())):
            codes[s, c] = (code, count)
    out[f"{tag}_branch"] = calls
    out[f"{tag}_codes"] = codes


def run_failures(R, out, p0):
    R.set_prims(p0)
    R.reset_nfail()
    out["p0"] = R.prims()
    done = 0
    for s in FAIL_STEPS:
        R.reset_branch()
        R.step(s - done)
        if done == 0:
            record_branch(R, out, "s1")
        done = s
        record(R, out, f"s{s}")


def make_c0f():
    from refwrap import RefRun
    G = dict(np.load(os.path.join(HERE, "c0_balsara1024.npz"), allow_pickle=False))
    R = RefRun("c0", 1024, 1, 1)
    R.set_time(0.0, 1.0e-5, 0.4)
    R.set_end_time(0.4)
    out = {"n": np.array([1024, 1, 1]), "cour": np.float64(0.4), "dt0": np.float64(1.0e-5)}
    run_failures(R, out, corrupt(G["p0"], FAILURES_C0))
    np.savez_compressed(os.path.join(HERE, "c0f_balsara1024_failures.npz"), **out)


def make_c1f():
    from refwrap import RefRun, torus_params
    from coolharm3d_amd.problems import torus_grid
    G = dict(np.load(os.path.join(HERE, "c1_torus24x24x16.npz"), allow_pickle=False))
    n = tuple(int(x) for x in G["n"])
    g = torus_grid(*n)
    R = RefRun("c1", *n, params=torus_params(g))
    R.set_time(0.0, 1.0e-5, g.cour)
    out = {"n": np.array(n), "cour": np.float64(g.cour), "dt0": np.float64(1.0e-5), "r_isco": G["r_isco"]}
    run_failures(R, out, corrupt(G["p0"], FAILURES_C1))
    np.savez_compressed(os.path.join(HERE, "c1f_torus24x24x16_failures.npz"), **out)


def make_c0_full():
    from refwrap import RefRun
    G = dict(np.load(os.path.join(HERE, "c0_balsara1024.npz"), allow_pickle=False))
    R = RefRun("c0", 1024, 1, 1)
    R.set_time(0.0, 1.0e-5, 0.4)
    R.set_prims(G["p0"])
    R.reset_nfail()
    nstep = 0
    while R.scalars()["t"] < 0.4:
        R.step(1)
        nstep += 1
    out = {"p0": G["p0"], "cour": np.float64(0.4), "dt0": np.float64(1.0e-5), "t_end": np.float64(0.4),
           "nstep": np.int64(nstep)}
    record(R, out, "end")
    np.savez_compressed(os.path.join(HERE, "c0_balsara1024_t04.npz"), **out)


def make_c0():
    from refwrap import RefRun
    R = RefRun("c0", 1024, 1, 1)
    R.set_time(0.0, 1.0e-5, 0.4)
    out = {"p0": R.prims(), "n": np.array([1024, 1, 1]), "cour": np.float64(0.4), "dt0": np.float64(1.0e-5)}
    R.reset_nfail()
    done = 0
    for s in STEPS:
        R.step(s - done)
        done = s
        record(R, out, f"s{s}")
    np.savez_compressed(os.path.join(HERE, "c0_balsara1024.npz"), **out)


def make_c1(coords="mks"):
    from refwrap import RefRun, torus_params
    from coolharm3d_amd.problems import torus, torus_grid
    n = (24, 24, 16)
    g = torus_grid(*n, coords=coords)
    raw = torus(g)
    cfg, name = ("c1d3", "c1d3_torus24x24x16.npz") if coords == "diag3" else ("c1", "c1_torus24x24x16.npz")
    R = RefRun(cfg, *n, params=torus_params(g))
    R.set_prims(raw)
    R.fixup_bounds()
    R.reset_nfail()
    R.set_time(0.0, 1.0e-5, g.cour)
    out = {"p0": R.prims(), "n": np.array(n), "cour": np.float64(g.cour), "dt0": np.float64(1.0e-5),
           "r_isco": np.float64(R.scalars()["r_isco"])}
    done = 0
    for s in STEPS:
        R.step(s - done)
        done = s
        record(R, out, f"s{s}")
    np.savez_compressed(os.path.join(HERE, name), **out)


CRASH_CASES = ((0, 600), (0, 520), (0, 510))   # zones [lo, hi) of the shock tube given rho = -1e-3
CRASH_STEPS = 5


def run_crash_case(lo, hi):
    """One process: the shock tube with rho = -1e-3 in zones [lo, hi), stepped
    with the driver's check_for_crash() after every step (src/main.c:81-100);
    prints the step reached, the reference ends the process on a crash."""
    from refwrap import RefRun
    R = RefRun("c0", 1024, 1, 1)
    R.set_time(0.0, 1.0e-5, 0.4)
    p = R.prims()
    p[3 + lo:3 + hi, :, :, 0] = -1.0e-3
    R.set_prims(p)
    R.reset_nfail()
    for s in range(1, CRASH_STEPS + 1):
        R.step(1)
        R.check_for_crash()
        print("survived", s, flush=True)


def make_c0c():
    import re
    out = {"cases": np.array(CRASH_CASES), "steps": np.int64(CRASH_STEPS), "cour": np.float64(0.4),
           "dt0": np.float64(1.0e-5)}
    crash_step, n_bad = [], []
    for lo, hi in CRASH_CASES:
        r = subprocess.run([sys.executable, __file__, "c0c_run", str(lo), str(hi)], capture_output=True, text=True)
        txt = r.stdout + r.stderr
        m = re.search(r"Reached critical fraction.*?nstep\s*=\s*(\d+).*?n_bad\s*=\s*(\d+)", txt, re.S)
        if m:
            crash_step.append(int(m.group(1)))
            n_bad.append(int(m.group(2)))
        else:
            assert f"survived {CRASH_STEPS}" in r.stdout, txt[-2000:]
            crash_step.append(0)
            n_bad.append(0)
    out["crash_step"] = np.array(crash_step)
    out["n_bad"] = np.array(n_bad)
    np.savez_compressed(os.path.join(HERE, "c0c_balsara1024_crash.npz"), **out)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "c0c_run":
        run_crash_case(int(sys.argv[2]), int(sys.argv[3]))
    elif len(sys.argv) > 1:
        {"c0c": make_c0c, "c0": make_c0, "c1": make_c1, "c0f": make_c0f, "c1f": make_c1f, "c0t": make_c0_full,
         "c1d3": lambda: make_c1("diag3")}[sys.argv[1]]()
    else:
        for cfg in ("c0", "c1", "c0f", "c1f", "c0t", "c1d3", "c0c"):
            subprocess.run([sys.executable, __file__, cfg], check=True, stdout=subprocess.DEVNULL)
            print("wrote", cfg)
