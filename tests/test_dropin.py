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
    g = torus_grid(24, 24, 16)
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


def _check_c0(R):
    G = golden("c0_balsara1024.npz")
    assert np.array_equal(interior(R["p0"]), interior(G["p0"]))
    for s in (1, 10, 100):
        assert np.array_equal(interior(R[f"s{s}_prims"]), G[f"s{s}_prims"]), f"state after {s} steps"
        assert np.array_equal(interior(R[f"s{s}_pflag"]), G[f"s{s}_pflag"].astype(np.int32))
        assert np.array_equal(R[f"s{s}_nfail"], G[f"s{s}_nfail"])
        assert float(R[f"s{s}_t"]) == float(G[f"s{s}_t"]) and float(R[f"s{s}_dt"]) == float(G[f"s{s}_dt"])


def _check_c1(R, tol100):
    G = golden("c1_torus24x24x16.npz")
    assert np.array_equal(R["p0"], G["p0"])
    for s, tol in ((1, 1e-12), (10, 1e-12), (100, tol100)):
        assert_rel_l1(G[f"s{s}_prims"], interior(R[f"s{s}_prims"]), tol, f"drop-in torus after {s} steps")
        assert np.array_equal(interior(R[f"s{s}_pflag"]), G[f"s{s}_pflag"].astype(np.int32))
        ok, info = nfail_equal_modulo_cooling_init(G[f"s{s}_nfail"], R[f"s{s}_nfail"], 30)
        assert ok, info
        assert float(R[f"s{s}_dt"]) == pytest.approx(float(G[f"s{s}_dt"]), rel=1e-10)


@pytest.mark.skipif(not _built("dropin_c0_host"), reason="drop-in build not present (make -C oracle dropin)")
def test_dropin_host_balsara_bit_identical(tmp_path):
    _check_c0(_run("dropin_c0_host", tmp_path))


@pytest.mark.skipif(not _built("dropin_c0_host"), reason="drop-in build not present (make -C oracle dropin)")
def test_dropin_host_balsara_staged_bit_identical(tmp_path):
    """The per-stage path of the shim (hgpu_advance, boundary flags through
    mpi_global_imax2(), hgpu_fixup_stage, hgpu_end_step after
    mpi_global_vmax(), mpi_global_min()) in a single-rank run."""
    _check_c0(_run("dropin_c0_host", tmp_path, stages=True))


@pytest.mark.slow
@pytest.mark.skipif(not _built("dropin_c1_host"), reason="drop-in build not present (make -C oracle dropin)")
def test_dropin_host_torus(tmp_path):
    _check_c1(_run("dropin_c1_host", tmp_path), 1e-12)


@pytest.mark.gpu
@pytest.mark.skipif(not _built("dropin_c0_gpu"), reason="drop-in build not present (make -C oracle dropin)")
def test_dropin_gpu_balsara_bit_identical(tmp_path):
    _check_c0(_run("dropin_c0_gpu", tmp_path))


@pytest.mark.gpu
@pytest.mark.skipif(not _built("dropin_c1_gpu"), reason="drop-in build not present (make -C oracle dropin)")
def test_dropin_gpu_torus(tmp_path):
    _check_c1(_run("dropin_c1_gpu", tmp_path), 1e-9)


@pytest.mark.gpu
@pytest.mark.skipif(not _built("dropin_c1_gpu"), reason="drop-in build not present (make -C oracle dropin)")
def test_dropin_gpu_torus_staged(tmp_path):
    _check_c1(_run("dropin_c1_gpu", tmp_path, stages=True), 1e-9)
