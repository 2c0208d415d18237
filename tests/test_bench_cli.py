"""bench.py's command-line contract and host-side helpers (no GPU needed)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    """One GPU only: --gpus above 1 exits with status 2 and says so."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_units_per_launch():
    """Faces per flux launch (mean of the three directions, transverse ghosts
    included), edges per EMF launch, interior zones per update launch."""
    u = bench.units_per_launch(4, 3, 2)
    n1t, n2t, n3t = 10, 9, 8
    assert u["flux"] == (5 * n2t * n3t + n1t * 4 * n3t + n1t * n2t * 3) / 3.0
    assert u["emf"] == 5 * 4 * 3
    assert u["update"] == 4 * 3 * 2


def test_committed_traffic_profiles_hold_the_algorithmic_bytes():
    """The committed counter summaries parse, and each timed scope's HBM bytes
    per launch are at least its algorithmic bytes (re-reads only add)."""
    units = bench.units_per_launch(192, 192, 128)
    for workload, algo in (("torus192x192x128_fp64", bench.ALGO_BYTES),
                           ("torus192x192x128_mixed", bench.ALGO_BYTES_MIXED)):
        t = bench.measured_traffic(workload)
        assert t is not None and set(t) == {"flux", "emf", "update"}
        for scope in t:
            assert t[scope] >= 0.95 * algo[scope] * units[scope], (workload, scope)
