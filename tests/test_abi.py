"""The C ABI library loads and exports every entry point include/hgpu.h declares.

No compute calls are made on the product library here (CPU-only machines have
no HIP device); its only call is hgpu_create() on a CPU machine, which must
fail with an error message rather than fall back to anything.  The argument
and configuration errors of the boundary are exercised on the host build of
the same C ABI.
"""
import ctypes
import os
import re
import subprocess

import pytest

from helpers import ROOT

HDR = os.path.join(ROOT, "include", "hgpu.h")
LIB = os.path.join(ROOT, "coolharm3d_amd", "libhgpu.so")


def declared_functions():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hgpu_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("hgpu_create", "hgpu_step", "hgpu_substep", "hgpu_upload_prims", "hgpu_download_prims",
                 "hgpu_download_pflag", "hgpu_download_nfail", "hgpu_con2prim_batch", "hgpu_fixup_bounds"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libhgpu.so not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, f"not exported: {missing}"
    lib = ctypes.CDLL(LIB)
    for n in declared_functions():
        assert hasattr(lib, n)


def test_no_cpu_fallback_without_gpu():
    """On a machine without a GPU the product library refuses to create a
    context (the Python layer raises), instead of computing on the CPU."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except Exception:
        pass
    if not os.path.exists(LIB):
        pytest.skip("libhgpu.so not built")
    from coolharm3d_amd import Harm3D, balsara1_grid
    from coolharm3d_amd._lib import HgpuError
    with pytest.raises(HgpuError, match="no HIP device"):
        Harm3D(balsara1_grid(16))


# The create-time checks and argument errors of the boundary, on the host build
# of the same C ABI (the GPU build runs the same checks before any device call).
_BAD_CONFIGS = [
    ("n1", 0, "grid sizes must be >= 1"),
    ("metric", 1, "metric must be 0"),
    ("coords", 2, "coords must be 0"),
    ("recon", 3, "recon must be 2"),
    ("flux", 2, "flux must be 0"),
    ("bc", 5, "bc must be 0 or 10"),
    ("precision", 2, "precision must be 0"),
    ("bc_nbr", 64, "bc_nbr holds 6 face bits"),
]


@pytest.mark.parametrize("field,value,msg", _BAD_CONFIGS)
def test_create_rejects_invalid_configurations(field, value, msg):
    from coolharm3d_amd import _lib, torus_grid
    from coolharm3d_amd.harm import config_for
    lib = _lib.load(hostcheck=True)
    cfg = config_for(torus_grid(8, 8, 8))
    setattr(cfg, field, value)
    h = ctypes.c_void_p()
    rc = lib.hgpu_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc < 0 and h.value is None
    assert msg in lib.hgpu_last_error().decode()


def test_create_rejects_inconsistent_options():
    from coolharm3d_amd import _lib, torus_grid, balsara1_grid
    from coolharm3d_amd.harm import config_for
    lib = _lib.load(hostcheck=True)
    h = ctypes.c_void_p()
    cases = [(torus_grid(8, 8, 8), dict(recon=2, ct_para=1), "recon must be 5"),
             (balsara1_grid(16), dict(cooling=1), "needs a black-hole metric"),
             (torus_grid(8, 2, 8), {}, "at least NG interior zones")]
    for grid, kw, msg in cases:
        cfg = config_for(grid, **kw)
        assert lib.hgpu_create(ctypes.byref(cfg), ctypes.byref(h)) < 0
        assert msg in lib.hgpu_last_error().decode()
    cfg = config_for(torus_grid(8, 8, 8, coords="diag3"))
    cfg.diag3_exponent = 4
    assert lib.hgpu_create(ctypes.byref(cfg), ctypes.byref(h)) < 0
    assert "odd diag3_exponent" in lib.hgpu_last_error().decode()
    assert lib.hgpu_create(None, ctypes.byref(h)) == -1


def test_call_argument_errors():
    """Errors come back as negative status with a message; the context stays
    usable after them."""
    import numpy as np
    from coolharm3d_amd import _lib, torus_grid
    from coolharm3d_amd.harm import config_for
    lib = _lib.load(hostcheck=True)
    cfg = config_for(torus_grid(8, 8, 8))
    h = ctypes.c_void_p()
    assert lib.hgpu_create(ctypes.byref(cfg), ctypes.byref(h)) == 0
    try:
        assert lib.hgpu_step(h, -1) < 0 and b"nsteps" in lib.hgpu_last_error()
        assert lib.hgpu_substep(h, 2) < 0 and b"stage must be 0 or 1" in lib.hgpu_last_error()
        assert lib.hgpu_step(None, 1) < 0
        lo = np.array([0, 0, 0], np.int32)
        hi = np.array([100, 3, 3], np.int32)
        buf = np.zeros(8 * 101 * 4 * 4)
        rc = lib.hgpu_download_box(h, 0, lo.ctypes.data_as(_lib.c_int_p), hi.ctypes.data_as(_lib.c_int_p),
                                   buf.ctypes.data_as(_lib.c_double_p), None)
        assert rc < 0 and b"outside the grid" in lib.hgpu_last_error()
        assert lib.hgpu_step(h, 0) == 0
        nb = ctypes.c_longlong()
        assert lib.hgpu_device_bytes(h, ctypes.byref(nb)) == 0 and nb.value > 0
        assert lib.hgpu_device_bytes(h, None) < 0
    finally:
        assert lib.hgpu_destroy(h) == 0
