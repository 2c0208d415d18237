"""check_for_crash() on the device (hgpu_get_crash) against the reference's
own crash test: the shock tube with rho < 0 in the first 600, 520 or 510
zones, stepped with the driver's check_for_crash() after every step
(tests/golden/c0c_balsara1024_crash.npz, written by the reference build).
The reference ends the run at step 1 for 600 and 520 zones (n_bad 1199 and
1039 Utoprim_1d_ee failures against half of 2 x 1024) and survives 5 steps
with 510; the library latches the same step and n_bad, on the host build of
the kernels and on the GPU."""
import numpy as np
import pytest

from helpers import golden


def _run(hostcheck, lo, hi, steps):
    from coolharm3d_amd import Harm3D, balsara1_grid
    G0 = golden("c0_balsara1024.npz")
    H = Harm3D(balsara1_grid(1024, cour=0.4), hostcheck=hostcheck)
    p = G0["p0"].copy()
    p[3 + lo:3 + hi, :, :, 0] = -1.0e-3
    H.set_prims(p)
    H.set_time(0.0, 1.0e-5)
    H.step_ch(steps)
    return H.crash()


def _check(hostcheck):
    G = golden("c0c_balsara1024_crash.npz")
    for (lo, hi), step, n_bad in zip(G["cases"], G["crash_step"], G["n_bad"]):
        print(f"[crash check] rho < 0 in zones [{lo}, {hi}): reference step {step} n_bad {n_bad}", flush=True)
        assert _run(hostcheck, int(lo), int(hi), int(G["steps"])) == (int(step), int(n_bad))


def test_crash_check_host_build_matches_reference():
    _check(True)


@pytest.mark.gpu
def test_gpu_crash_check_matches_reference():
    _check(False)
