"""Batched con-to-prim of independent zones with a learned initial guess.

Python side of the ``hgpu_c2p_*`` / ``hgpu_mlp_*`` entry points of
``include/hgpu.h``.  The per-zone algorithm is the reference's
``Utoprim_2d_fast()`` (src/utoprim_2d_fast.c:135-408, called from
recover_primitives(), src/step_ch.c:832); arrays are structure-of-arrays
(``U[8][n]``, ``rth[2][n]``, ``prim[8][n]``).

On the GPU library every array is a ``torch`` CUDA tensor (torch only provides
device memory and the stream: calls are enqueued on torch's current stream, so
``torch.cuda.Event`` timing brackets them).  With ``hostcheck=True`` the host
build of the same bodies runs on numpy arrays; only tests use that.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

RETURN_CODES = {
    0: "success", -100: "U[RHO] <= 0", 2: "initial guess has u~^2 < 0 or too large",
    1: "Newton ended above the tolerance", 101: "Newton did not converge",
    201: "Newton produced a non-finite residual", 3: "W <= 0 or W too big", 4: "v^2 >= 1",
}


class Con2PrimBatch:
    def __init__(self, hostcheck: bool = False):
        self.host = hostcheck
        self.lib = _lib.load(hostcheck=hostcheck)
        sizes = (ctypes.c_int * 6)()
        self._chk(self.lib.hgpu_mlp_sizes(sizes), "hgpu_mlp_sizes")
        self.n_param, self.n_train, self.train_off, self.n_in, self.n_hidden, self.n_out = list(sizes)
        if not hostcheck:
            import torch
            if not torch.cuda.is_available():
                raise _lib.HgpuError("no HIP device available (the hgpu library runs only on a GPU)")
            self.torch = torch
            self.dev = torch.device("cuda:0")

    # -- memory plumbing --------------------------------------------------------
    def _chk(self, rc, what):
        _lib.check(self.lib, rc, what)

    def empty(self, shape, dtype):
        if self.host:
            return np.empty(shape, dtype=dtype)
        tdt = {np.float64: self.torch.float64, np.float32: self.torch.float32, np.int32: self.torch.int32,
               np.uint8: self.torch.uint8}[np.dtype(dtype).type]
        return self.torch.empty(shape, dtype=tdt, device=self.dev)

    def zeros(self, shape, dtype):
        a = self.empty(shape, dtype)
        a[...] = 0
        return a

    def asarray(self, a, dtype):
        """Device (or host-build) copy of an array-like, contiguous."""
        if self.host:
            return np.ascontiguousarray(np.asarray(a, dtype=dtype))
        if isinstance(a, self.torch.Tensor):
            return a.to(device=self.dev, dtype=self.empty((1,), dtype).dtype).contiguous()
        return self.torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=dtype))).to(self.dev)

    @staticmethod
    def ptr(a):
        if a is None:
            return None
        if isinstance(a, np.ndarray):
            if not a.flags["C_CONTIGUOUS"]:
                raise ValueError("arrays must be contiguous")
            return a.ctypes.data
        if not a.is_contiguous():
            raise ValueError("arrays must be contiguous")
        return a.data_ptr()

    def stream(self):
        return None if self.host else self.torch.cuda.current_stream().cuda_stream

    def to_numpy(self, a):
        return a if self.host else a.cpu().numpy()

    def sync(self):
        if not self.host:
            self.torch.cuda.synchronize()

    # -- samples and solve --------------------------------------------------------
    def samples(self, n, seed, a=0.9, gam=5.0 / 3.0, guess=False):
        """U[8][n], rth[2][n], true prim[8][n] (and the perturbed-truth guess)."""
        U = self.empty((8, n), np.float64)
        rth = self.empty((2, n), np.float64)
        prim = self.empty((8, n), np.float64)
        g = self.empty((8, n), np.float64) if guess else None
        self._chk(self.lib.hgpu_c2p_samples(self.stream(), n, seed, a, gam, self.ptr(U), self.ptr(rth),
                                            self.ptr(prim), self.ptr(g)), "hgpu_c2p_samples")
        out = {"U": U, "rth": rth, "prim": prim}
        if guess:
            out["guess"] = g
        return out

    def workspace(self, n, compact=True):
        nbytes = int(self.lib.hgpu_c2p_workspace_bytes(n, 1 if compact else 0))
        return self.empty((max(nbytes, 8),), np.uint8), nbytes

    def solve(self, U, rth, a=0.9, gam=5.0 / 3.0, mlp=None, plain_guess="cold", prim_in=None, compact=True,
              budget_first=0, budget_next=0, work=None, out=None):
        """Utoprim_2d_fast() on every zone.

        plain_guess: "cold" (rho = D, u = 1e-3 D, u~ = 0) or "prim" (prim_in, the
        reference's own guess).  mlp: parameter blob of a trained network, or
        None for the plain guess only.  Returns (prim, code, iters)."""
        n = U.shape[1]
        mode = {"prim": 0, "cold": 1}[plain_guess]
        if mode == 0 and prim_in is None:
            raise ValueError("plain_guess='prim' needs prim_in")
        if out is None:
            prim = self.empty((8, n), np.float64)
            code = self.empty((n,), np.int32)
            iters = self.empty((n,), np.int32)
        else:
            prim, code, iters = out
        if mode == 0:
            prim[...] = prim_in
        if compact and work is None:
            work = self.workspace(n, True)
        wptr, wbytes = (self.ptr(work[0]), work[1]) if compact else (None, 0)
        self._chk(self.lib.hgpu_c2p_solve(self.stream(), n, a, gam, self.ptr(U), self.ptr(rth), self.ptr(mlp), mode,
                                          1 if compact else 0, budget_first, budget_next, wptr, wbytes,
                                          self.ptr(prim), self.ptr(code), self.ptr(iters)), "hgpu_c2p_solve")
        return prim, code, iters

    # -- network ------------------------------------------------------------------
    def features(self, U, rth, a=0.9, gam=5.0 / 3.0, prim=None):
        n = U.shape[1]
        feat = self.empty((self.n_in, n), np.float32)
        target = self.empty((self.n_out, n), np.float32) if prim is not None else None
        self._chk(self.lib.hgpu_mlp_features(self.stream(), n, a, gam, self.ptr(U), self.ptr(rth), self.ptr(prim),
                                             self.ptr(feat), self.ptr(target)), "hgpu_mlp_features")
        return feat, target

    def forward(self, w, feat):
        n = feat.shape[1]
        y = self.empty((self.n_out, n), np.float32)
        self._chk(self.lib.hgpu_mlp_forward(self.stream(), n, self.ptr(w), self.ptr(feat), self.ptr(y)),
                  "hgpu_mlp_forward")
        return y

    def grad(self, w, feat, target):
        n = feat.shape[1]
        nbytes = int(self.lib.hgpu_mlp_workspace_bytes(n))
        work = self.empty((nbytes,), np.uint8)
        g = self.empty((self.n_train,), np.float32)
        loss = self.empty((1,), np.float32)
        self._chk(self.lib.hgpu_mlp_grad(self.stream(), n, self.ptr(w), self.ptr(feat), self.ptr(target),
                                         self.ptr(work), nbytes, self.ptr(g), self.ptr(loss)), "hgpu_mlp_grad")
        return g, loss

    def init(self, seed, feat, target):
        w = self.empty((self.n_param,), np.float32)
        self._chk(self.lib.hgpu_mlp_init(self.stream(), seed, feat.shape[1], self.ptr(feat), self.ptr(target),
                                         self.ptr(w)), "hgpu_mlp_init")
        return w

    def train(self, w, feat, target, seed, steps, batch, lr):
        n = feat.shape[1]
        nbytes = int(self.lib.hgpu_mlp_workspace_bytes(batch))
        work = self.empty((nbytes,), np.uint8)
        hist = self.empty((max(steps, 1),), np.float32)
        self._chk(self.lib.hgpu_mlp_train(self.stream(), n, self.ptr(feat), self.ptr(target), seed, steps, batch, lr,
                                          self.ptr(work), nbytes, self.ptr(w), self.ptr(hist)), "hgpu_mlp_train")
        return hist[:steps]

    def train_guess(self, n_train, seed, a=0.9, gam=5.0 / 3.0, steps=2000, batch=4096, lr=3e-3):
        """Trains the guess network on zones the library samples and solves itself.

        Zones are drawn from the workload's distribution (seed), solved from the
        cold guess; converged solutions (code 0) are the training set.  Returns
        (w, loss history)."""
        d = self.samples(n_train, seed, a, gam)
        prim, code, _ = self.solve(d["U"], d["rth"], a, gam, plain_guess="cold", compact=False)
        ok = (code == 0)
        U = d["U"][:, ok].contiguous() if not self.host else np.ascontiguousarray(d["U"][:, ok])
        rth = d["rth"][:, ok].contiguous() if not self.host else np.ascontiguousarray(d["rth"][:, ok])
        P = prim[:, ok].contiguous() if not self.host else np.ascontiguousarray(prim[:, ok])
        feat, target = self.features(U, rth, a, gam, prim=P)
        w = self.init(seed ^ 0x5EED, feat, target)
        hist = self.train(w, feat, target, seed ^ 0xBA7C4, steps, batch, lr)
        return w, hist


def stats(code, iters):
    """Mean Newton iterations and failure rate (code != 0) of a solve."""
    code = np.asarray(code)
    iters = np.asarray(iters)
    return {"mean_iters": float(iters.mean()), "max_iters": int(iters.max()),
            "failure_rate": float((code != 0).mean()), "n": int(code.size)}
