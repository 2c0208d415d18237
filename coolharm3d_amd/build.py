"""Builds the native libraries in-tree.

* ``coolharm3d_amd/libhgpu.so`` -- the product: HIP kernels for gfx950 behind the
  C ABI of ``include/hgpu.h`` (hipcc): the GRMHD advance (hgpu_kernels.hip) and
  the batched con-to-prim with its learned guess (hgpu_c2p.hip), and the
  binary-black-hole metric (hgpu_pn.hip).
* ``coolharm3d_amd/libhgpu_hostcheck.so`` -- the same kernel bodies compiled for
  the host (g++), used only by the CPU tests to compare the kernels' arithmetic
  with the reference where no GPU exists.

FP64 parity with the reference (gcc x86-64, no FMA contraction) needs
``-ffp-contract=off`` on both; packed FP32 instructions are disabled on the
device (``-packed-fp32-ops``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_FLAGS = ["-O3", "--offload-arch=gfx950", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-O2", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
              "-I" + os.path.join(ROCM, "include")]

GPU_LIB = os.path.join(PKG, "libhgpu.so")
HOSTCHECK_LIB = os.path.join(PKG, "libhgpu_hostcheck.so")


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_gpu():
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    tmp = GPU_LIB + ".tmp"
    _run([hipcc] + HIP_FLAGS + ["-o", tmp, os.path.join(CSRC, "hgpu_kernels.hip"),
                                os.path.join(CSRC, "hgpu_c2p.hip"), os.path.join(CSRC, "hgpu_pn.hip")])
    os.replace(tmp, GPU_LIB)


def build_hostcheck():
    cxx = shutil.which("g++") or "g++"
    tmp = HOSTCHECK_LIB + ".tmp"
    _run([cxx] + HOST_FLAGS + ["-o", tmp, os.path.join(CSRC, "hgpu_hostcheck.cpp"),
                               os.path.join(CSRC, "hgpu_c2p_host.cpp"), os.path.join(CSRC, "hgpu_pn_host.cpp")])
    os.replace(tmp, HOSTCHECK_LIB)


def build_all():
    build_gpu()
    build_hostcheck()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"all": build_all, "gpu": build_gpu, "hostcheck": build_hostcheck}[what]()
