#!/usr/bin/env python3
"""Benchmark of the per-substep GRMHD advance on one MI355X.

Workload: the plain FP64 3D torus of 192x192x128 zones -- Kerr a = 0.9 in
modified Kerr-Schild coordinates, PPM reconstruction, HLL fluxes, PPM-EMF
constrained transport, optically thin cooling to H/r = 0.1, floors
1e-7 r^-1.5 / 1e-9 r^-2.5 -- one GPU's share of the domain.  A step is one
step_ch(): predictor and corrector advance() with boundaries, fixups and the
CFL time step.  value = interior zones x steps / wall time of the timed steps
(zone-cycles/s, the reference's unit, doc/CHANGES:1826), with the state
resident in HBM.

Also reported:
  roofline          -- the dominant kernel, the conserved update + inversion
                       (k_update): the state's algorithmic bytes, 160 B per zone
                       and launch (320 B per zone-cycle over the step's two
                       update launches), over its average launch time (HIP
                       events on the library stream), against 8 TB/s; the
                       measured HBM bytes per launch (counter summary in
                       profiles/); each kernel family's own byte model beside it;
  mixed_precision   -- the same grid and steps with FP32 reconstruction, face
                       fluxes and EMFs (FP64 update, con-to-prim, metric, cooling),
                       and its speed-up over the FP64 step;
  precision_variant -- per-primitive relative L1 of the mixed step against the
                       FP64 step on the 256x256x192 torus after 1, 10, 100 steps;
  cpu_baseline      -- the reference C (oracle/_ref/c1, built from the reference
                       sources) on the host cores: independent processes on x3
                       slabs of the same grid, timed around step_ch() only;
  con2prim          -- batched Utoprim_2d_fast() over 2^24 independent Kerr-Schild
                       zones (r log-uniform in [1.3, 100], theta uniform in
                       [0.05, pi - 0.05], a = 0.9, Lorentz factor log-uniform in
                       [1, 50], b^2/rho in [1e-4, 1e2], u/rho in [1e-4, 10],
                       isotropic directions, Gamma = 5/3, FP64): solves/s, mean
                       Newton iterations and failure rate from the cold guess and
                       from the learned guess (8-64-64-2 GELU network, FP32,
                       trained on the GPU on zones the library solved itself,
                       training time reported apart), one-pass and compacted
                       solves, with the reference's own Utoprim_2d_fast() timed
                       on the host cores beside it;
  bbh               -- circumbinary disk around an inspiralling equal-mass binary
                       (separation 20 M, 1PN near-zone metric recomputed on the
                       device at three time levels every step) on 300x160x400
                       zones, r in [1.5, 13] separations: mixed-precision step,
                       autotuned launch shapes, graph-replayed steps.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-cycles/s per RK2 step, 1 MI355X, mixed vs FP64; con2prim solves/s"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes of the roofline figure: the step's state traffic, 320 B per
# zone-cycle (the primitive state read three times and written twice, 8 FP64
# values per pass), carried by the two conserved-update launches of a step:
# 160 B per interior zone and update launch.  Faces, EMFs, U^n, flags, halos
# and geometry are intermediates and excluded.
STATE_BYTES_PER_ZONE_CYCLE = 320.0
UPDATE_STATE_BYTES = STATE_BYTES_PER_ZONE_CYCLE / 2

# Per-kernel byte models (FP64 state), reported beside it:
#   flux   : per face   -- 8 primitives of the zone read once (64 B) + 7 face fluxes written (56 B; the
#                          normal field's flux is never read and not stored)
#   emf    : per edge   -- 6 B-field face fluxes read (48 B) + 3 EMFs written (24 B)
#   update : per zone   -- p read (64 B); 15 non-magnetic face fluxes + 3 EMFs read (144 B); result
#                          (64 B) and gamma (8 B) written; the predictor writes U^n (64 B), the
#                          corrector reads U^n and ph (128 B): 344 / 408 B
ALGO_BYTES = {"flux": 120.0, "emf": 72.0, "update": 0.5 * (344.0 + 408.0)}
# mixed precision: face fluxes and EMFs are FP32 (flux writes 28 B/face, EMF 24+12 B/edge,
# the update reads 18 FP32 values), and the corrector forms U^n again from the p^n it
# reads anyway (no U^n written or read: 64 B less per launch on average)
ALGO_BYTES_MIXED = {"flux": 92.0, "emf": 36.0, "update": 0.5 * (344.0 + 408.0) - 72.0 - 64.0}


def measured_traffic(workload):
    """HBM bytes per launch of each timed scope (flux, emf, update), from the
    committed counter summary profiles/hbm_traffic_<workload>.md (two rocprofv3
    --pmc runs, FETCH_SIZE and WRITE_SIZE, calibrated on a plain 1 GiB device
    copy); None when absent."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    for ln in open(path):
        if ln.startswith("hbm_bytes_per_launch "):
            return json.loads(ln[len("hbm_bytes_per_launch "):])
    return None


def units_per_launch(n1, n2, n3):
    ng = 3
    n1t, n2t, n3t = n1 + 2 * ng, n2 + 2 * ng, n3 + 2 * ng
    faces = ((n1 + 1) * n2t * n3t + n1t * (n2 + 1) * n3t + n1t * n2t * (n3 + 1)) / 3.0
    return {"flux": faces, "emf": (n1 + 1) * (n2 + 1) * (n3 + 1), "update": n1 * n2 * n3}


def cpu_worker(args):
    """One CPU process on an x3 slab (the reference build, or the oracle when
    the reference build is absent): prints zone-cycles and seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from coolharm3d_amd.problems import torus_grid, torus
    n1, n2, n3 = args.n1, args.n2, args.slab
    g = torus_grid(n1, n2, n3, x3_extent=0.5 * math.pi * n3 / args.n3)
    if args.cpu_kind == "reference":
        from refwrap import RefRun, torus_params
        R = RefRun("c1", n1, n2, n3, params=torus_params(g))
        R.set_prims(torus(g, seed=1234 + args.worker))
        R.fixup_bounds()
        R.set_time(0.0, 1.0e-5, g.cour)
    else:
        from pyoracle import OracleRun
        R = OracleRun(g)
        R.set_prims(torus(g, seed=1234 + args.worker))
        R.fixup_bounds()
        R.set_time(0.0, 1.0e-5)
    R.step(1)  # untimed first step (the reference also builds its cooling tables here)
    t0 = time.perf_counter()
    R.step(args.slab_steps)
    el = time.perf_counter() - t0
    import resource
    rss = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1024.0
    print(json.dumps({"zc": n1 * n2 * n3 * args.slab_steps, "s": el, "rss_mb": rss}), flush=True)


CPU_SHARE = 16   # host cores one GPU's job may use on the benchmark machine


def host_cpu():
    """The host's CPU model, logical CPU count and the cores the baselines use:
    every core of the job's share (at most CPU_SHARE processes, whatever the
    machine's total)."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None and os.path.exists("/proc/cpuinfo"):
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "affinity": avail, "used": min(CPU_SHARE, avail)}


def cpu_baseline(args):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from refwrap import have_ref
    from pyoracle import have_oracle
    kind = "reference" if have_ref("c1") else "port"
    if kind == "port" and not have_oracle():
        return {"value": None, "unit": "zone-cycles/s", "cores": 0, "kind": kind,
                "sample": "neither oracle/_ref/c1 nor oracle/_build/liboracle.so is built on this machine"}
    host = host_cpu()
    ncores = host["used"]
    slab = 4
    procs = []
    for w in range(ncores):
        cmd = [sys.executable, os.path.abspath(__file__), "--_cpu-worker", "--worker", str(w),
               "--cpu-kind", kind,
               "--n1", str(args.n1), "--n2", str(args.n2), "--n3", str(args.n3),
               "--slab", str(slab), "--slab-steps", str(args.cpu_steps)]
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                                      env=dict(os.environ, OMP_NUM_THREADS="1")))
    rate, n_ok = 0.0, 0
    for p in procs:
        out, _ = p.communicate()
        for line in out.splitlines():
            if line.startswith("{"):
                r = json.loads(line)
                rate += r["zc"] / r["s"]
                n_ok += 1
    what = "reference step_ch() (gcc -O3)" if kind == "reference" else "oracle restatement of step_ch() (gcc -O3)"
    return {"value": rate if n_ok else None, "unit": "zone-cycles/s", "cores": n_ok, "kind": kind,
            "cpu": host,
            "sample": (f"{what}, {n_ok} independent processes (the job's host-core share), each a "
                       f"{args.n1}x{args.n2}x{slab} x3-slab of the {args.n1}x{args.n2}x{args.n3} torus, "
                       f"{args.cpu_steps} timed steps after 1 untimed; rates summed (no halo exchange: "
                       f"an upper bound on an MPI run)")}


C2P_BYTES = 152.0   # per solve: U (64 B) + (r, theta) (16 B) read, prims (64 B) + code, iterations (8 B) written


def c2p_cpu_worker(args):
    """One CPU process: the reference's Utoprim_2d_fast() (oracle/_ref/c1) on its
    share of the workload's zones from the cold guess; prints solves and seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import refwrap
    from coolharm3d_amd.c2p import Con2PrimBatch
    H = Con2PrimBatch(hostcheck=True)          # host build: inputs only
    n = args.c2p_cpu_zones
    d = H.samples(n, 1 + args.worker, 0.9, 5.0 / 3.0)
    U, rth = d["U"].numpy(), d["rth"].numpy()
    prim = np.zeros((8, n))
    a = 0.9
    # cold guess rho = D, u = 1e-3 D: D = alpha U[RHO] / sqrt(-g), Kerr-Schild lapse and determinant
    r, th = rth
    c, s = np.cos(th), np.sin(th)
    sig = r * r + a * a * c * c
    alpha = 1.0 / np.sqrt(1.0 + 2.0 * r / sig)
    gdet = r * r * np.abs(s) * (1.0 + 0.5 * (a * a / (r * r)) * (1.0 + (c - s) * (c + s)))
    D = U[0] * alpha / gdet
    prim[0], prim[1] = D, 1e-3 * D
    prim = np.ascontiguousarray(prim)
    t0 = time.perf_counter()
    code = refwrap.c2p_batch(0.9, 5.0 / 3.0, U, rth, prim)
    el = time.perf_counter() - t0
    print(json.dumps({"n": n, "s": el, "fail": int((code != 0).sum())}), flush=True)


def c2p_cpu_baseline(args):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from refwrap import have_ref
    if not have_ref("c1"):
        return {"value": None, "unit": "solves/s", "cores": 0, "kind": "reference",
                "sample": "oracle/_ref/c1 is not built on this machine"}
    host = host_cpu()
    ncores = host["used"]
    procs = []
    for w in range(ncores):
        cmd = [sys.executable, os.path.abspath(__file__), "--_c2p-cpu-worker", "--worker", str(w),
               "--c2p-cpu-zones", str(args.c2p_cpu_zones)]
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                                      env=dict(os.environ, OMP_NUM_THREADS="1")))
    rate, n_ok = 0.0, 0
    for p in procs:
        out, _ = p.communicate()
        for line in out.splitlines():
            if line.startswith("{"):
                r = json.loads(line)
                rate += r["n"] / r["s"]
                n_ok += 1
    return {"value": rate if n_ok else None, "unit": "solves/s", "cores": n_ok, "kind": "reference",
            "cpu": host,
            "sample": (f"reference Utoprim_2d_fast() (gcc -O3, oracle/_ref/c1) from the cold guess, {n_ok} "
                       f"independent processes, each {args.c2p_cpu_zones} zones of the same distribution "
                       f"(different seeds); rates summed")}


def con2prim_report(args):
    """Batched con-to-prim of 2^k independent zones: cold vs learned guess,
    one-pass vs compacted solves."""
    import numpy as np
    from coolharm3d_amd.c2p import Con2PrimBatch, stats
    a, gam = 0.9, 5.0 / 3.0
    B = Con2PrimBatch()
    n = 1 << args.c2p_log2
    t0 = time.perf_counter()
    w, hist = B.train_guess(1 << 18, 777, a, gam, steps=args.c2p_train_steps, batch=4096, lr=3e-3)
    B.sync()
    train_s = time.perf_counter() - t0
    d = B.samples(n, 1, a, gam)
    work = B.workspace(n, True)
    out = (B.empty((8, n), np.float64), B.empty((n,), np.int32), B.empty((n,), np.int32))
    T = B.timer()
    runs = {}
    for name, kw in (("cold_onepass", dict(compact=False)), ("cold_compact", dict(compact=True)),
                     ("learned_onepass", dict(compact=False, mlp=w)), ("learned_compact", dict(compact=True, mlp=w))):
        ms = []
        for rep in range(1 + args.c2p_reps):
            T.start()
            B.solve(d["U"], d["rth"], a, gam, plain_guess="cold", work=work, out=out, **kw)
            el = T.stop()
            if rep > 0:
                ms.append(el)
        st = stats(out[1], out[2])
        t = float(np.median(ms))
        runs[name] = {"ms": t, "solves_per_s": n / t * 1e3, "mean_iters": st["mean_iters"],
                      "max_iters": st["max_iters"], "failure_rate": st["failure_rate"],
                      "achieved_GBs": C2P_BYTES * n / (t * 1e-3) / 1e9}
    best = max(runs, key=lambda k: runs[k]["solves_per_s"])
    rep = {"zones": n, "value": runs[best]["solves_per_s"], "unit": "solves/s", "best": best, "runs": runs,
           "plain_guess": "cold (rho = D, u = 1e-3 D, u~ = 0)",
           "learned_guess_gain": {
               "mean_iters": runs["cold_compact"]["mean_iters"] / runs["learned_compact"]["mean_iters"],
               "solves_per_s": runs["learned_compact"]["solves_per_s"] / runs["cold_compact"]["solves_per_s"]},
           "training": {"zones": 1 << 18, "steps": args.c2p_train_steps, "batch": 4096, "seconds": train_s,
                        "loss_first": float(hist[:10].mean()), "loss_last": float(hist[-10:].mean())},
           "roofline": {"bound": "hbm", "algo_bytes_per_solve": C2P_BYTES, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "achieved": runs[best]["achieved_GBs"], "frac": runs[best]["achieved_GBs"] / HBM_PEAK_GBS}}
    if args.cpu_baseline == "auto":
        rep["cpu_baseline"] = c2p_cpu_baseline(args)
    return rep


def bbh_report(args):
    """Circumbinary disk on the time-dependent binary metric (one GPU's share)."""
    import numpy as np
    from coolharm3d_amd import Harm3D, bbh_grid, bbh_disk
    from coolharm3d_amd.problems import bbh_gdet
    from coolharm3d_amd.pn import PnMetric
    shape = tuple(int(x) for x in args.bbh_grid.split("x"))
    g = bbh_grid(*shape)
    p0 = bbh_disk(g, bbh_gdet(g, PnMetric))
    H = Harm3D(g, precision="mixed", track_nfail=0)
    H.set_prims(p0)
    del p0
    H.fixup_bounds()
    H.set_time(0.0, 1.0e-3)
    H.step_ch(2)
    H.profile_enable(True)
    H.step_ch(3)
    kern = {k: H.profile_read(k)[0] / 3 for k in ("metric", "flux", "emf", "update", "recover", "bounds", "fixup")}
    H.profile_enable(False)
    if args.tune == "on":
        shapes = H.autotune(reps=1)
    else:
        shapes = {f: H.launch_shape(f) for f in H.FAMILIES[:5]}
    H.set_graph(args.graph == "on")
    H.step_ch(args.warmup)
    H.sync()
    t0 = time.perf_counter()
    H.step_ch(args.bbh_steps)
    H.sync()
    el = time.perf_counter() - t0
    divb = H.divb_max()
    finite = bool(np.isfinite(H.prims()).all())
    dev_gb = H.device_bytes() / 1e9
    H.close()
    ncell = shape[0] * shape[1] * shape[2]
    return {"workload": f"bbh_disk_{shape[0]}x{shape[1]}x{shape[2]}_mixed", "value": ncell * args.bbh_steps / el,
            "unit": "zone-cycles/s", "ms_per_step": el / args.bbh_steps * 1e3, "steps": args.bbh_steps,
            "steps_total": 5 + args.warmup + args.bbh_steps,
            "kernel_ms_per_step_eager": kern, "launch_shapes": shapes, "graph": args.graph == "on", "finite": finite,
            "divb_max": divb, "device_gb": dev_gb,
            "metric": "1PN near-zone, quasi-circular inspiral (Peters), separation 20 M"}


def timed_run(args, g, p0, precision):
    """Warm-up, launch-shape tuning, a profiled pass (per-kernel HIP-event times)
    and the timed step_ch() calls (replayed from a captured graph unless
    --graph off); returns timings and per-kernel figures."""
    from coolharm3d_amd import Harm3D
    H = Harm3D(g, track_nfail=1, precision=precision)
    H.set_prims(p0)
    H.fixup_bounds()
    H.set_time(0.0, 1.0e-5)
    H.step_ch(args.warmup)
    H.sync()
    shapes = H.autotune(reps=2) if args.tune == "on" else None
    nprof = args.prof_steps
    H.profile_enable(True)
    H.sync()
    H.step_ch(nprof)
    H.sync()
    kern = {nm: H.profile_read(nm) for nm in ("flux", "emf", "update", "recover", "bounds", "fixup", "step")}
    H.profile_enable(False)
    if args.graph == "on":
        H.set_graph(True)
        H.step_ch(2)
    H.sync()
    t0 = time.perf_counter()
    H.step_ch(args.steps)
    H.sync()
    el = time.perf_counter() - t0
    import numpy as np
    finite = bool(np.isfinite(H.prims()).all())
    H.close()
    return el, kern, finite, shapes


def precision_report(shape, steps=(1, 10, 100)):
    """Mixed-precision variant vs the FP64 step on the same torus: per-primitive
    relative L1 (interior zones) after each of `steps`, and each precision's
    ms/step over the last 80 of the steps (replayed from a graph, default
    launch shapes; the 10 steps before them capture the graph)."""
    import numpy as np
    from coolharm3d_amd import Harm3D, torus_grid, torus
    g = torus_grid(*shape)
    p0 = torus(g)
    states, ms = {}, {}
    for prec in ("fp64", "mixed"):
        H = Harm3D(g, track_nfail=0, precision=prec)
        H.set_prims(p0)
        H.fixup_bounds()
        H.set_time(0.0, 1.0e-5)
        done, out = 0, {}
        for s in steps:
            if s == steps[-1] and s - done > 80:
                H.set_graph(True)
                H.step_ch(s - done - 80)
                H.sync()
                t0 = time.perf_counter()
                H.step_ch(80)
                H.sync()
                ms[prec] = (time.perf_counter() - t0) / 80 * 1e3
            else:
                H.step_ch(s - done)
            done = s
            out[s] = H.prims()[3:-3, 3:-3, 3:-3].reshape(-1, 8)
        states[prec] = out
        H.close()
    names = ["rho", "u", "u1", "u2", "u3", "B1", "B2", "B3"]
    rep = {}
    for s in steps:
        a, b = states["fp64"][s], states["mixed"][s]
        num = np.abs(b - a).sum(axis=0)
        den = np.abs(a).sum(axis=0)
        rep[str(s)] = {names[v]: float(num[v] / den[v]) if den[v] > 0 else float(num[v]) for v in range(8)}
    timing = None
    if len(ms) == 2:
        ncell = shape[0] * shape[1] * shape[2]
        timing = {"ms_per_step": ms, "zone_cycles_per_s": {k: ncell / (v * 1e-3) for k, v in ms.items()},
                  "speedup_mixed_vs_fp64": ms["fp64"] / ms["mixed"]}
    return {"grid": list(shape), "steps": list(steps), "rel_l1_vs_fp64": rep, "timing": timing,
            "reconstruction_flux_emf": "fp32", "update_con2prim_metric_cooling": "fp64"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n1", type=int, default=192)
    ap.add_argument("--n2", type=int, default=192)
    ap.add_argument("--n3", type=int, default=128)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="precision of the headline run (default: the plain FP64 step)")
    ap.add_argument("--mixed", choices=["auto", "off"], default="auto",
                    help="also time the mixed-precision step on the same grid")
    ap.add_argument("--precision-report", choices=["auto", "off"], default="auto",
                    help="per-primitive L1 of the mixed step vs FP64 on the 256x256x192 torus after 1, 10, 100 steps")
    ap.add_argument("--report-grid", default="256x256x192")
    ap.add_argument("--cpu-baseline", choices=["auto", "off"], default="auto")
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--_cpu-worker", action="store_true", dest="cpu_worker")
    ap.add_argument("--worker", type=int, default=0)
    ap.add_argument("--slab", type=int, default=4)
    ap.add_argument("--slab-steps", type=int, default=4)
    ap.add_argument("--cpu-kind", choices=["reference", "port"], default="reference")
    ap.add_argument("--graph", choices=["on", "off"], default="on", help="replay timed steps from a HIP graph")
    ap.add_argument("--tune", choices=["on", "off"], default="on", help="autotune launch shapes before timing")
    ap.add_argument("--prof-steps", type=int, default=3, help="steps of the per-kernel profiled pass")
    ap.add_argument("--bbh", choices=["auto", "off"], default="auto",
                    help="circumbinary-disk report on the binary metric")
    ap.add_argument("--bbh-grid", default="300x160x400")
    ap.add_argument("--bbh-steps", type=int, default=100)
    ap.add_argument("--c2p", choices=["auto", "off"], default="auto",
                    help="batched con-to-prim report (2^c2p-log2 zones)")
    ap.add_argument("--c2p-log2", type=int, default=24)
    ap.add_argument("--c2p-reps", type=int, default=3)
    ap.add_argument("--c2p-train-steps", type=int, default=1500)
    ap.add_argument("--c2p-cpu-zones", type=int, default=1 << 19)
    ap.add_argument("--_c2p-cpu-worker", action="store_true", dest="c2p_cpu_worker")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="launch one plain device copy of this many MiB before the run "
                         "(calibrates the HBM byte counters of a profiler run)")
    args = ap.parse_args()

    if args.cpu_worker:
        cpu_worker(args)
        return
    if args.c2p_cpu_worker:
        c2p_cpu_worker(args)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU halo exchange is a design note, not part of this build)", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    from coolharm3d_amd import torus_grid, torus

    n1, n2, n3 = args.n1, args.n2, args.n3
    if args.copy_probe_mib > 0:
        import numpy as np
        from coolharm3d_amd.c2p import Con2PrimBatch
        B = Con2PrimBatch()
        nb = args.copy_probe_mib * (1 << 20)
        src, dst = B.empty((nb,), np.uint8), B.empty((nb,), np.uint8)
        B._chk(B.lib.hgpu_dev_copy_kernel(dst.ptr, src.ptr, nb, None), "hgpu_dev_copy_kernel")
        B.sync()
        del src, dst
    g = torus_grid(n1, n2, n3)
    p0 = torus(g)
    el, kern, finite, shapes = timed_run(args, g, p0, args.precision)
    assert finite, "non-finite state after the timed steps"
    ncell = n1 * n2 * n3
    value = ncell * args.steps / el

    tag = "fp64" if args.precision == "fp64" else "mixed"
    units = units_per_launch(n1, n2, n3)
    model = ALGO_BYTES if tag == "fp64" else ALGO_BYTES_MIXED
    per_launch_s = {k: kern[k][0] / max(kern[k][1], 1) / 1e3 for k in ("flux", "emf", "update")}
    workload = f"torus{n1}x{n2}x{n3}_{tag}"
    traffic_tab = measured_traffic(workload)
    # the dominant kernel: the conserved update + inversion (k_update), the longest
    # single kernel of the step; its algorithmic bytes are the state's 160 B per zone
    avg_s = per_launch_s["update"]
    algo = UPDATE_STATE_BYTES * units["update"]
    achieved = algo / avg_s / 1e9
    traffic = traffic_tab.get("update") if traffic_tab else None
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "k_update",
                "algo_bytes_per_unit": UPDATE_STATE_BYTES, "unit_of_work": "interior zone",
                "units_per_launch": units["update"],
                "traffic_source": f"profiles/hbm_traffic_{workload}.md" if traffic else None,
                "avg_launch_ms": avg_s * 1e3, "algo_bytes_per_launch": algo,
                "kernel_ms_per_step": {k: kern[k][0] / args.prof_steps
                                       for k in ("flux", "emf", "update", "recover", "bounds", "fixup")},
                "state_roofline_frac": value * STATE_BYTES_PER_ZONE_CYCLE / (HBM_PEAK_GBS * 1e9),
                "kernel_models": {
                    k: {"bytes_per_unit": model[k], "units_per_launch": units[k],
                        "avg_launch_ms": per_launch_s[k] * 1e3,
                        "achieved_GBs": model[k] * units[k] / per_launch_s[k] / 1e9,
                        "frac": model[k] * units[k] / per_launch_s[k] / 1e9 / HBM_PEAK_GBS,
                        "traffic": traffic_tab.get(k) if traffic_tab else None}
                    for k in ("flux", "emf", "update")},
                "launch_shapes": shapes, "timed_steps_graph": args.graph == "on"}

    mixed = None
    if args.mixed == "auto" and args.precision == "fp64":
        el_m, kern_m, fin_m, shapes_m = timed_run(args, g, p0, "mixed")
        v_m = ncell * args.steps / el_m
        mixed = {"value": v_m, "unit": "zone-cycles/s", "ms_per_step": el_m / args.steps * 1e3,
                 "speedup_vs_fp64": v_m / value, "finite": fin_m, "launch_shapes": shapes_m,
                 "kernel_ms_per_step": {k: kern_m[k][0] / args.prof_steps
                                        for k in ("flux", "emf", "update", "recover", "bounds", "fixup")}}

    prec_rep = None
    if args.precision_report == "auto":
        prec_rep = precision_report(tuple(int(x) for x in args.report_grid.split("x")))

    cpu = None
    if rank == 0 and args.cpu_baseline == "auto":
        cpu = cpu_baseline(args)

    c2p = con2prim_report(args) if args.c2p == "auto" else None
    bbh = bbh_report(args) if args.bbh == "auto" else None

    line = {
        "metric": METRIC, "value": value, "unit": "zone-cycles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if tag == "fp64" else "f64+f32",
        "data": "synthetic (Fishbone-Moncrief torus, seeded +-2% u perturbation)",
        "config": {"workload": f"torus_{n1}x{n2}x{n3}_{tag}", "grid": [n1, n2, n3], "spin": 0.9,
                   "recon": "PPM+MC", "flux": "HLL", "ct": "PPM EMFs", "cooling": "h/r=0.1",
                   "precision": tag},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mixed_precision": mixed,
        "precision_variant": prec_rep,
        "con2prim": c2p,
        "bbh": bbh,
    }
    if rank == 0:
        print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
