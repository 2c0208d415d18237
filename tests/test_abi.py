"""The C ABI library loads and exports every entry point include/hgpu.h declares.

No compute calls are made here (CPU-only machines have no HIP device); the
only call is hgpu_create() on a CPU machine, which must fail with an error
message rather than fall back to anything.
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
