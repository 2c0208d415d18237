"""Batched con-to-prim of independent zones with a learned initial guess.

Python side of the ``hgpu_c2p_*`` / ``hgpu_mlp_*`` entry points of
``include/hgpu.h``.  The per-zone algorithm is the reference's
``Utoprim_2d_fast()`` (src/utoprim_2d_fast.c:135-408, called from
recover_primitives(), src/step_ch.c:832); arrays are structure-of-arrays
(``U[8][n]``, ``rth[2][n]``, ``prim[8][n]``).

Device memory, copies and timing go through the library's own plumbing
(``hgpu_dev_*``, ``hgpu_event_*``), so no other GPU runtime is loaded beside
it.  All work is enqueued on the null stream.  With ``hostcheck=True`` the host
build of the same bodies runs, its "device" buffers being host memory; only
tests use that.
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


class DevArray:
    """A buffer of the library's device (host memory in the host build)."""

    def __init__(self, lib, shape, dtype):
        self.lib = lib
        self.shape = tuple(int(x) for x in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _lib.check(lib, lib.hgpu_dev_alloc(self.nbytes, ctypes.byref(p)), "hgpu_dev_alloc")
        self.ptr = p.value

    def __del__(self):
        if getattr(self, "ptr", None):
            self.lib.hgpu_dev_free(self.ptr)
            self.ptr = None

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _lib.check(self.lib, self.lib.hgpu_dev_copy(out.ctypes.data, self.ptr, self.nbytes, 1, None), "hgpu_dev_copy")
        return out

    def upload(self, a):
        a = np.ascontiguousarray(np.asarray(a, self.dtype))
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _lib.check(self.lib, self.lib.hgpu_dev_copy(self.ptr, a.ctypes.data, self.nbytes, 0, None), "hgpu_dev_copy")
        _lib.check(self.lib, self.lib.hgpu_stream_sync(None), "hgpu_stream_sync")
        return self

    def copy_from(self, other):
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _lib.check(self.lib, self.lib.hgpu_dev_copy(self.ptr, other.ptr, self.nbytes, 2, None), "hgpu_dev_copy")
        return self


class Timer:
    """Elapsed time of work enqueued between start() and stop(), from events."""

    def __init__(self, lib):
        self.lib = lib
        self.ev = [ctypes.c_void_p(), ctypes.c_void_p()]
        for e in self.ev:
            _lib.check(lib, lib.hgpu_event_create(ctypes.byref(e)), "hgpu_event_create")

    def start(self):
        _lib.check(self.lib, self.lib.hgpu_event_record(self.ev[0], None), "hgpu_event_record")

    def stop(self):
        """Records the end event and returns the milliseconds in between."""
        _lib.check(self.lib, self.lib.hgpu_event_record(self.ev[1], None), "hgpu_event_record")
        ms = ctypes.c_float()
        _lib.check(self.lib, self.lib.hgpu_event_elapsed(self.ev[0], self.ev[1], ctypes.byref(ms)),
                   "hgpu_event_elapsed")
        return float(ms.value)

    def __del__(self):
        for e in getattr(self, "ev", []):
            self.lib.hgpu_event_destroy(e)


class Con2PrimBatch:
    def __init__(self, hostcheck: bool = False):
        self.host = hostcheck
        self.lib = _lib.load(hostcheck=hostcheck)
        if not hostcheck:
            n = ctypes.c_int()
            self._chk(self.lib.hgpu_dev_count(ctypes.byref(n)), "hgpu_dev_count")
            if n.value < 1:
                raise _lib.HgpuError("no HIP device available (the hgpu library runs only on a GPU)")
        sizes = (ctypes.c_int * 6)()
        self._chk(self.lib.hgpu_mlp_sizes(sizes), "hgpu_mlp_sizes")
        self.n_param, self.n_train, self.train_off, self.n_in, self.n_hidden, self.n_out = list(sizes)

    # -- memory plumbing --------------------------------------------------------
    def _chk(self, rc, what):
        _lib.check(self.lib, rc, what)

    def empty(self, shape, dtype):
        return DevArray(self.lib, shape, dtype)

    def upload(self, a, dtype=None):
        a = np.asarray(a, dtype) if dtype is not None else np.asarray(a)
        return DevArray(self.lib, a.shape, a.dtype).upload(a)

    @staticmethod
    def ptr(a):
        return None if a is None else a.ptr

    def sync(self):
        self._chk(self.lib.hgpu_stream_sync(None), "hgpu_stream_sync")

    def timer(self):
        return Timer(self.lib)

    # -- samples and solve --------------------------------------------------------
    def samples(self, n, seed, a=0.9, gam=5.0 / 3.0, guess=False):
        """U[8][n], rth[2][n], true prim[8][n] (and the perturbed-truth guess)."""
        U = self.empty((8, n), np.float64)
        rth = self.empty((2, n), np.float64)
        prim = self.empty((8, n), np.float64)
        g = self.empty((8, n), np.float64) if guess else None
        self._chk(self.lib.hgpu_c2p_samples(None, n, seed, a, gam, U.ptr, rth.ptr, prim.ptr, self.ptr(g)),
                  "hgpu_c2p_samples")
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
            out = (self.empty((8, n), np.float64), self.empty((n,), np.int32), self.empty((n,), np.int32))
        prim, code, iters = out
        if mode == 0:
            prim.copy_from(prim_in)
        if compact and work is None:
            work = self.workspace(n, True)
        wptr, wbytes = (work[0].ptr, work[1]) if compact else (None, 0)
        self._chk(self.lib.hgpu_c2p_solve(None, n, a, gam, U.ptr, rth.ptr, self.ptr(mlp), mode,
                                          1 if compact else 0, budget_first, budget_next, wptr, wbytes,
                                          prim.ptr, code.ptr, iters.ptr), "hgpu_c2p_solve")
        return prim, code, iters

    # -- network ------------------------------------------------------------------
    def features(self, U, rth, a=0.9, gam=5.0 / 3.0, prim=None):
        n = U.shape[1]
        feat = self.empty((self.n_in, n), np.float32)
        target = self.empty((self.n_out, n), np.float32) if prim is not None else None
        self._chk(self.lib.hgpu_mlp_features(None, n, a, gam, U.ptr, rth.ptr, self.ptr(prim), feat.ptr,
                                             self.ptr(target)), "hgpu_mlp_features")
        return feat, target

    def forward(self, w, feat):
        n = feat.shape[1]
        y = self.empty((self.n_out, n), np.float32)
        self._chk(self.lib.hgpu_mlp_forward(None, n, w.ptr, feat.ptr, y.ptr), "hgpu_mlp_forward")
        return y

    def grad(self, w, feat, target):
        n = feat.shape[1]
        nbytes = int(self.lib.hgpu_mlp_workspace_bytes(n))
        work = self.empty((nbytes,), np.uint8)
        g = self.empty((self.n_train,), np.float32)
        loss = self.empty((1,), np.float32)
        self._chk(self.lib.hgpu_mlp_grad(None, n, w.ptr, feat.ptr, target.ptr, work.ptr, nbytes, g.ptr, loss.ptr),
                  "hgpu_mlp_grad")
        return g, loss

    def init(self, seed, feat, target):
        w = self.empty((self.n_param,), np.float32)
        self._chk(self.lib.hgpu_mlp_init(None, seed, feat.shape[1], feat.ptr, target.ptr, w.ptr), "hgpu_mlp_init")
        return w

    def train(self, w, feat, target, seed, steps, batch, lr):
        n = feat.shape[1]
        nbytes = int(self.lib.hgpu_mlp_workspace_bytes(batch))
        work = self.empty((nbytes,), np.uint8)
        hist = self.empty((max(steps, 1),), np.float32)
        self._chk(self.lib.hgpu_mlp_train(None, n, feat.ptr, target.ptr, seed, steps, batch, lr, work.ptr, nbytes,
                                          w.ptr, hist.ptr), "hgpu_mlp_train")
        return hist.numpy()[:steps]

    def train_guess(self, n_train, seed, a=0.9, gam=5.0 / 3.0, steps=2000, batch=4096, lr=3e-3):
        """Trains the guess network on zones the library samples and solves itself.

        Zones are drawn from the workload's distribution (seed) and solved from
        the cold guess; the converged ones (code 0) are the training set.
        Returns (w, loss history)."""
        d = self.samples(n_train, seed, a, gam)
        prim, code, _ = self.solve(d["U"], d["rth"], a, gam, plain_guess="cold", compact=False)
        ok = code.numpy() == 0
        U = self.upload(d["U"].numpy()[:, ok])
        rth = self.upload(d["rth"].numpy()[:, ok])
        P = self.upload(prim.numpy()[:, ok])
        feat, target = self.features(U, rth, a, gam, prim=P)
        w = self.init(seed ^ 0x5EED, feat, target)
        hist = self.train(w, feat, target, seed ^ 0xBA7C4, steps, batch, lr)
        return w, hist


def stats(code, iters):
    """Mean Newton iterations and failure rate (code != 0) of a solve."""
    code = code.numpy() if isinstance(code, DevArray) else np.asarray(code)
    iters = iters.numpy() if isinstance(iters, DevArray) else np.asarray(iters)
    return {"mean_iters": float(iters.mean()), "max_iters": int(iters.max()),
            "failure_rate": float((code != 0).mean()), "n": int(code.size)}
