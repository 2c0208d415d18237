"""The stage split and halo hooks of the C ABI (hgpu_advance,
hgpu_fixup_stage, hgpu_end_step, hgpu_download_box/hgpu_upload_box): a grid
cut into blocks, stepped as a domain-decomposed step_ch() steps
(coolharm3d_amd.blocks.BlockGrid: the exchange order of bounds_mpi(),
src/harm_mpi.c:1631-1744; the flag, speed and dt reductions of
src/bounds.c:2164-2166 and src/step_ch.c:561, :618), gives the single-block
result bit for bit: state, pflag, nfail, t and dt.  The failure fixtures put
flagged zones on block faces, so the pflag exchange and the second exchange
after the fixups are exercised.  hgpu_substep twice + hgpu_end_step equals
hgpu_step.

The host build of the kernel bodies runs these on the CPU; the gpu-marked
copies run the GPU library.
"""
import os

import numpy as np
import pytest

from helpers import golden, interior, failure_grid, ROOT

HOSTLIB = os.path.join(ROOT, "coolharm3d_amd", "libhgpu_hostcheck.so")
needs_host = pytest.mark.skipif(not os.path.exists(HOSTLIB), reason="host check build missing")


def _whole(which, hostcheck, steps):
    from coolharm3d_amd import Harm3D
    grid, kw, G = failure_grid(which)
    H = Harm3D(grid, hostcheck=hostcheck, **kw)
    H.set_prims(G["p0"])
    H.set_time(0.0, float(G["dt0"]))
    out = []
    done = 0
    for s in steps:
        H.step_ch(s - done)
        done = s
        out.append((interior(H.prims()), interior(H.pflag()), interior(H.nfail().transpose(1, 2, 3, 0)),
                    H.nfailtot(), H.time()))
    return out


def _blocks(which, ncpux, hostcheck, steps):
    from coolharm3d_amd.blocks import BlockGrid
    grid, kw, G = failure_grid(which)
    B = BlockGrid(grid, ncpux, hostcheck=hostcheck, **kw)
    B.set_prims(G["p0"])
    B.set_time(0.0, float(G["dt0"]))
    out = []
    done = 0
    for s in steps:
        B.step_ch(s - done)
        done = s
        out.append((interior(B.prims()), interior(B.pflag()), interior(B.nfail().transpose(1, 2, 3, 0)),
                    B.nfailtot(), B.time()))
    return out


def _compare(whole, split, steps):
    for s, (w, b) in zip(steps, zip(whole, split)):
        assert np.array_equal(w[0], b[0]), f"state after {s} steps: max |diff| {np.abs(w[0] - b[0]).max():.3e}"
        assert np.array_equal(w[1], b[1]), f"pflag after {s} steps"
        assert np.array_equal(w[2], b[2]), f"nfail after {s} steps"
        assert w[4] == b[4], f"clock after {s} steps"
    # the reference's x1 inflow check counts GAMMA_CALC in physical ghost zones:
    # the totals still agree
    for s, (w, b) in zip(steps, zip(whole, split)):
        assert np.array_equal(w[3], b[3]), f"nfailtot after {s} steps"


def _check_split_touches_faces(which, ncpux):
    """The fixture's flagged zones of step one lie within NG zones of an
    internal face of the decomposition (so boundary_mpi_pflag is raised)."""
    grid, _, G = failure_grid(which)
    nb = (grid.n1 // ncpux[0], grid.n2 // ncpux[1], grid.n3 // ncpux[2])
    idx = np.argwhere(G["s1_pflag"] != 0)
    near = False
    for d in range(3):
        if ncpux[d] > 1:
            r = idx[:, d] % nb[d]
            near |= bool(((r < 3) | (r >= nb[d] - 3)).any())
    assert near


@needs_host
@pytest.mark.parametrize("which,ncpux", [("c1f", (2, 2, 2)), ("c1f", (1, 3, 1)), ("c0f", (8, 1, 1))])
def test_blocks_equal_single_block_host_build(which, ncpux):
    _check_split_touches_faces(which, ncpux)
    steps = (1, 2, 5)
    _compare(_whole(which, True, steps), _blocks(which, ncpux, True, steps), steps)


@needs_host
@pytest.mark.parametrize("which", ["c0f", "c1f"])
def test_substeps_equal_step_host_build(which):
    _substeps_equal_step(which, True)


def _substeps_equal_step(which, hostcheck):
    from coolharm3d_amd import Harm3D
    grid, kw, G = failure_grid(which)
    A = Harm3D(grid, hostcheck=hostcheck, **kw)
    S = Harm3D(grid, hostcheck=hostcheck, **kw)
    for H in (A, S):
        H.set_prims(G["p0"])
        H.set_time(0.0, float(G["dt0"]))
    for n in range(3):
        A.step_ch(1)
        S.advance(0)
        S.advance(1)
        S.end_step()
        assert np.array_equal(A.prims(), S.prims()), f"state, step {n + 1}"
        assert np.array_equal(A.pflag(), S.pflag()), f"pflag, step {n + 1}"
        assert np.array_equal(A.nfail(), S.nfail()), f"nfail, step {n + 1}"
        assert np.array_equal(A.gamma(), S.gamma()), f"gamma, step {n + 1}"
        assert A.time() == S.time()
    # the substep path reproduces the reference's flag map of step one
    S2 = Harm3D(grid, hostcheck=hostcheck, **kw)
    S2.set_prims(G["p0"])
    S2.set_time(0.0, float(G["dt0"]))
    S2.advance(0)
    S2.advance(1)
    assert np.array_equal(interior(S2.pflag()), G["s1_pflag"].astype(np.int32))


@needs_host
def test_box_transfer_round_trip_and_bounds_checks():
    from coolharm3d_amd import Harm3D
    from coolharm3d_amd._lib import HgpuError
    grid, kw, G = failure_grid("c1f")
    H = Harm3D(grid, hostcheck=True, **kw)
    H.set_prims(G["p0"])
    lo, hi = (2, 5, 0), (7, 9, 21)
    pr, pf = H.download_box(0, lo, hi)
    assert np.array_equal(pr, G["p0"][2:8, 5:10, 0:22])
    pr2 = pr * 2.0
    pf2 = np.arange(pf.size, dtype=np.int32).reshape(pf.shape)
    H.upload_box(0, lo, hi, pr2, pf2)
    full = H.prims()
    assert np.array_equal(full[2:8, 5:10, 0:22], pr2)
    assert np.array_equal(H.pflag()[2:8, 5:10, 0:22], pf2)
    assert np.array_equal(full[8:], G["p0"][8:])
    for bad in (((0, 0, 0), (30, 3, 3)), ((-1, 0, 0), (3, 3, 3)), ((4, 0, 0), (3, 3, 3))):
        with pytest.raises(HgpuError):
            H.download_box(0, *bad)


@pytest.mark.gpu
@pytest.mark.parametrize("which,ncpux", [("c1f", (2, 2, 2)), ("c0f", (8, 1, 1))])
def test_gpu_blocks_equal_single_block(which, ncpux):
    print(f"[gpu decomposition] {which} {ncpux}", flush=True)
    steps = (1, 2, 5)
    _compare(_whole(which, False, steps), _blocks(which, ncpux, False, steps), steps)


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["c0f", "c1f"])
def test_gpu_substeps_equal_step(which):
    print(f"[gpu substeps] {which}", flush=True)
    _substeps_equal_step(which, False)
