"""The drop-in: the reference's own driver, initial data and diagnostics with
step_ch() served by coolharm3d_amd/dropin/step_ch_hgpu.c (oracle/Makefile,
`make dropin`).  The reference globals are read back through the same
harness as the pure reference build, so these tests check the boundary a
maintainer would use, not the Python wrapper.

The *_host builds link the host build of the kernel bodies (CPU); the *_gpu
builds link libhgpu.so and run on the GPU.  The shim uploads the reference's
own geometry and connection, so the torus agrees with the reference to the
device/host libm (pow) only.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import golden, interior, assert_rel_l1, nfail_equal_modulo_cooling_init, ROOT

SCRIPT = r"""
import sys, numpy as np
sys.path[:0] = [{root!r}, {oracle!r}]
from refwrap import RefRun, torus_params
cfg = {cfg!r}
out = {{}}
if "c0" in cfg:
    R = RefRun(cfg, 1024, 1, 1)
    R.set_time(0.0, 1.0e-5, 0.4)
else:
    from coolharm3d_amd.problems import torus, torus_grid
    g = torus_grid(24, 24, 16, coords="diag3" if "c1d3" in cfg else "mks")
    R = RefRun(cfg, 24, 24, 16, params=torus_params(g))
    R.set_prims(torus(g))
    R.fixup_bounds()
    R.set_time(0.0, 1.0e-5, g.cour)
R.reset_nfail()
out["p0"] = R.prims()
done = 0
for s in (1, 10, 100):
    R.step(s - done)
    done = s
    sc = R.scalars()
    out[f"s{{s}}_prims"] = R.prims()
    out[f"s{{s}}_pflag"] = R.pflag()
    out[f"s{{s}}_nfail"] = R.nfail()
    out[f"s{{s}}_t"] = sc["t"]
    out[f"s{{s}}_dt"] = sc["dt"]
    out[f"s{{s}}_mtot"] = R.diag()["M_tot"]
np.savez({path!r}, **out)
"""


def _built(cfg):
    return os.path.exists(os.path.join(ROOT, "oracle", "_ref", cfg, "libref.so"))


def _run(cfg, tmp_path, stages=False):
    """The reference driver with the drop-in; stages: step_ch() through the
    per-stage hooks a decomposed run uses (HGPU_STAGES=1)."""
    path = str(tmp_path / f"{cfg}.npz")
    code = SCRIPT.format(root=ROOT, oracle=os.path.join(ROOT, "oracle"), cfg=cfg, path=path)
    env = dict(os.environ, HGPU_STAGES="1" if stages else "0")
    print(f"[drop-in] running the reference driver with {cfg} (stages {stages})", flush=True)
    subprocess.run([sys.executable, "-c", code], check=True, stdout=subprocess.DEVNULL, timeout=600, env=env)
    return dict(np.load(path))


def _check_m_tot(R, cfg_ref, tmp_path, rel):
    """M_tot, which the driver prints every step (src/main.c:71), against the
    reference build's own after the same steps: rel = 0 where the sum runs in
    the reference's loop order (host build), else the order's round-off."""
    ref = _run(cfg_ref, tmp_path)
    for s in (1, 10, 100):
        a, b = float(R[f"s{s}_mtot"]), float(ref[f"s{s}_mtot"])
        assert b != 0.
        assert abs(a - b) <= rel * abs(b), (s, a, b)


def _check_c0(R):
    G = golden("c0_balsara1024.npz")
    assert np.array_equal(interior(R["p0"]), interior(G["p0"]))
    for s in (1, 10, 100):
        assert np.array_equal(interior(R[f"s{s}_prims"]), G[f"s{s}_prims"]), f"state after {s} steps"
        assert np.array_equal(interior(R[f"s{s}_pflag"]), G[f"s{s}_pflag"].astype(np.int32))
        assert np.array_equal(R[f"s{s}_nfail"], G[f"s{s}_nfail"])
        assert float(R[f"s{s}_t"]) == float(G[f"s{s}_t"]) and float(R[f"s{s}_dt"]) == float(G[f"s{s}_dt"])


def _check_c1(R, tol100, fixture="c1_torus24x24x16.npz"):
    G = golden(fixture)
    assert np.array_equal(R["p0"], G["p0"])
    for s, tol in ((1, 1e-12), (10, 1e-12), (100, tol100)):
        assert_rel_l1(G[f"s{s}_prims"], interior(R[f"s{s}_prims"]), tol, f"drop-in torus after {s} steps")
        assert np.array_equal(interior(R[f"s{s}_pflag"]), G[f"s{s}_pflag"].astype(np.int32))
        ok, info = nfail_equal_modulo_cooling_init(G[f"s{s}_nfail"], R[f"s{s}_nfail"], 30)
        assert ok, info
        assert float(R[f"s{s}_dt"]) == pytest.approx(float(G[f"s{s}_dt"]), rel=1e-10)


@pytest.mark.skipif(not _built("dropin_c0_host"), reason="drop-in build not present (make -C oracle dropin)")
def test_dropin_host_balsara_bit_identical(tmp_path):
    R = _run("dropin_c0_host", tmp_path)
    _check_c0(R)
    _check_m_tot(R, "c0", tmp_path, 0.)


@pytest.mark.skipif(not _built("dropin_c0_host"), reason="drop-in build not present (make -C oracle dropin)")
def test_dropin_host_balsara_staged_bit_identical(tmp_path):
    """The per-stage path of the shim (hgpu_advance, boundary flags through
    mpi_global_imax2(), hgpu_fixup_stage, hgpu_end_step after
    mpi_global_vmax(), mpi_global_min()) in a single-rank run."""
    _check_c0(_run("dropin_c0_host", tmp_path, stages=True))


@pytest.mark.slow
@pytest.mark.skipif(not _built("dropin_c1_host"), reason="drop-in build not present (make -C oracle dropin)")
def test_dropin_host_torus(tmp_path):
    R = _run("dropin_c1_host", tmp_path)
    _check_c1(R, 1e-12)
    _check_m_tot(R, "c1", tmp_path, 1e-12)


@pytest.mark.gpu
@pytest.mark.skipif(not _built("dropin_c0_gpu"), reason="drop-in build not present (make -C oracle dropin)")
def test_dropin_gpu_balsara_bit_identical(tmp_path):
    R = _run("dropin_c0_gpu", tmp_path)
    _check_c0(R)
    _check_m_tot(R, "c0", tmp_path, 1e-14)


@pytest.mark.gpu
@pytest.mark.skipif(not _built("dropin_c1_gpu"), reason="drop-in build not present (make -C oracle dropin)")
def test_dropin_gpu_torus(tmp_path):
    R = _run("dropin_c1_gpu", tmp_path)
    _check_c1(R, 1e-9)
    _check_m_tot(R, "c1", tmp_path, 1e-12)


@pytest.mark.slow
@pytest.mark.skipif(not _built("dropin_c1d3_host"), reason="drop-in build not present (make -C oracle dropin)")
def test_dropin_host_torus_diag3(tmp_path):
    """The reference driver built with COORD_DIAGONAL3 (the binary-disk set-up's
    theta map) and the drop-in: the shim maps the coordinates and their
    diag3_exponent/diag3_factor globals."""
    _check_c1(_run("dropin_c1d3_host", tmp_path), 1e-12, "c1d3_torus24x24x16.npz")


@pytest.mark.gpu
@pytest.mark.skipif(not _built("dropin_c1d3_gpu"), reason="drop-in build not present (make -C oracle dropin)")
def test_dropin_gpu_torus_diag3(tmp_path):
    _check_c1(_run("dropin_c1d3_gpu", tmp_path), 1e-9, "c1d3_torus24x24x16.npz")


@pytest.mark.gpu
@pytest.mark.skipif(not _built("dropin_c1_gpu"), reason="drop-in build not present (make -C oracle dropin)")
def test_dropin_gpu_torus_staged(tmp_path):
    _check_c1(_run("dropin_c1_gpu", tmp_path, stages=True), 1e-9)


DYN_OBJ = os.path.join(ROOT, "oracle", "_ref", "dropin_dyn", "step_ch_hgpu.o")


@pytest.mark.skipif(not os.path.exists(DYN_OBJ), reason="dynamic-metric shim not compiled (make -C oracle dropin)")
def test_dropin_maps_the_binary_metric_and_diag3():
    """The shim compiled against the reference's headers with
    METRIC_GENERAL_DYNAMIC + METRIC_DYNAMIC_FULLPN_NZ and COORD_DIAGONAL3 (the
    binary-disk configuration, whose metric bodies the reference does not
    ship, so it cannot link): it reads the binary and coordinate globals the
    library's binary metric needs and uploads no host geometry."""
    out = subprocess.run(["nm", DYN_OBJ], capture_output=True, text=True, check=True).stdout
    undefined = {l.split()[-1] for l in out.splitlines() if l.split()[:1] == ["U"]}
    for g in ("m_bh1", "m_bh2", "initial_bbh_separation", "phi_0_bbh", "diag3_exponent", "diag3_factor"):
        assert g in undefined, g
    assert "hgpu_upload_geometry" not in undefined and "geom_arr" not in undefined


SYNC_SCRIPT = r"""
import sys, numpy as np
sys.path[:0] = [{root!r}, {oracle!r}]
from refwrap import RefRun
R = RefRun({cfg!r}, 1024, 1, 1)
R.set_time(0.0, 1.0e-5, 0.4)
R.set_n_restart(7)
R.reset_nfail()
out = {{"p0": R.prims()}}
for s in range(1, 11):
    R.step(1)
    out[f"s{{s}}"] = R.prims()
np.savez({path!r}, **out)
"""


@pytest.mark.skipif(not _built("dropin_c0_host"), reason="drop-in build not present (make -C oracle dropin)")
def test_dropin_host_sync_when_due(tmp_path):
    """HGPU_HOST_SYNC=2: the host arrays are refreshed exactly after the steps
    whose state the driver reads next (restart_write() every n_restart = 7
    steps, diag(OUT_EVERY)'s energy dump every 10th step), and are stale
    otherwise."""
    path = str(tmp_path / "sync.npz")
    code = SYNC_SCRIPT.format(root=ROOT, oracle=os.path.join(ROOT, "oracle"), cfg="dropin_c0_host", path=path)
    subprocess.run([sys.executable, "-c", code], check=True, stdout=subprocess.DEVNULL, timeout=600,
                   env=dict(os.environ, HGPU_HOST_SYNC="2"), cwd=str(tmp_path))
    R = dict(np.load(path))
    G = golden("c0_balsara1024.npz")
    assert np.array_equal(R["s6"], R["p0"])          # steps 1-6: nothing due, host untouched
    assert not np.array_equal(R["s7"], R["p0"])      # step 7: restart due
    assert np.array_equal(R["s9"], R["s7"])          # steps 8-9: stale again
    assert np.array_equal(interior(R["s10"]), G["s10_prims"])   # step 10: energy dump due


RESTART_SCRIPT = r"""
import sys, numpy as np
sys.path[:0] = [{root!r}, {oracle!r}]
from refwrap import RefRun
phase = {phase!r}
if phase == "write":
    R = RefRun({cfg!r}, 1024, 1, 1)
    R.set_time(0.0, 1.0e-5, 0.4)
    R.step(5)
    R.restart_write()
    R.step(5)
else:
    R = RefRun({cfg!r}, 1024, 1, 1, restart=True)
    R.step(5)
np.savez({path!r}, p=R.prims(), t=R.scalars()["t"], dt=R.scalars()["dt"], nstep=R.nstep())
"""


def _restart_pair(cfg, tmp_path):
    d = tmp_path / cfg
    d.mkdir()
    res = {}
    for phase in ("write", "read"):
        path = str(d / f"{phase}.npz")
        code = RESTART_SCRIPT.format(root=ROOT, oracle=os.path.join(ROOT, "oracle"), cfg=cfg, path=path, phase=phase)
        subprocess.run([sys.executable, "-c", code], check=True, stdout=subprocess.DEVNULL, timeout=600,
                       cwd=str(d), env=dict(os.environ, HGPU_HOST_SYNC="1"))
        res[phase] = dict(np.load(path))
    assert (d / "rdump0.0000").exists()
    return res


@pytest.mark.skipif(not (_built("dropin_c0_host") and _built("c0")), reason="reference/drop-in builds not present")
def test_dropin_host_ascii_restart_round_trip(tmp_path):
    """The reference's ASCII checkpoint (restart_write() / restart_init(),
    src/restart.c) through the drop-in: 5 steps, checkpoint, 5 more steps; a
    fresh process reads the checkpoint (restart_init() instead of init(), the
    shim uploads what it read on its first step) and runs 5 steps.  The
    restarted run equals the pure reference build doing the same round trip,
    bit for bit, and continues where the uninterrupted run continues (to the
    precision the checkpoint format keeps)."""
    ref = _restart_pair("c0", tmp_path)
    drop = _restart_pair("dropin_c0_host", tmp_path)
    assert np.array_equal(drop["write"]["p"], ref["write"]["p"])
    assert np.array_equal(drop["read"]["p"], ref["read"]["p"])
    assert float(drop["read"]["t"]) == float(ref["read"]["t"]) and int(drop["read"]["nstep"]) == 10
    assert float(drop["read"]["dt"]) == float(ref["read"]["dt"])
    e = np.abs(drop["read"]["p"] - drop["write"]["p"]).max()
    assert e <= 1e-12 * np.abs(drop["write"]["p"]).max()
