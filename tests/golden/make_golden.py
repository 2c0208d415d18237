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


def make_c1():
    from refwrap import RefRun, torus_params
    from coolharm3d_amd.problems import torus, torus_grid
    n = (24, 24, 16)
    g = torus_grid(*n)
    raw = torus(g)
    R = RefRun("c1", *n, params=torus_params(g))
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
    np.savez_compressed(os.path.join(HERE, "c1_torus24x24x16.npz"), **out)


if __name__ == "__main__":
    if len(sys.argv) > 1:
        {"c0": make_c0, "c1": make_c1}[sys.argv[1]]()
    else:
        for cfg in ("c0", "c1"):
            subprocess.run([sys.executable, __file__, cfg], check=True, stdout=subprocess.DEVNULL)
            print("wrote", cfg)
