"""Oracle of the batched con-to-prim workload and of the guess network.

TEST INFRASTRUCTURE: imported only by tests/, the smoke check and the bench's
CPU baseline, as the checker; the product never calls it.

* ``philox4x32_10`` / ``uniforms``: the counter-based generator (Salmon et al.,
  SC11) in numpy, pinned by the published known-answer vectors.
* ``samples``: the zone sampler (r, theta, Lorentz factor, magnetisation, u/rho,
  isotropic directions) restated with the oracle's Kerr-Schild record
  (``orc_ks_geom_rth``, src/metric.c:1168-1252) and ``primtoflux``
  (src/phys.c:28-63).  The sampler has no counterpart in the reference: its
  parity is with its own definition, pinned by invariants (b^2/rho, Lorentz
  factor, unit directions) and by inverting the conserved variables back with
  the oracle's Utoprim_2d_fast() (src/utoprim_2d_fast.c:135-408).
* ``mlp_forward`` / ``mlp_loss_grad``: the 8-64-64-2 GELU network and its
  backward pass in float64 numpy, pinned by central finite differences.  The
  network has no counterpart in the reference.
"""
from __future__ import annotations

import numpy as np

import pyoracle

MASK32 = np.uint64(0xFFFFFFFF)
PHILOX_M = (np.uint64(0xD2511F53), np.uint64(0xCD9E8D57))
PHILOX_W = (np.uint64(0x9E3779B9), np.uint64(0xBB67AE85))


def philox4x32_10(ctr, key):
    """ctr: (N, 4) uint32-valued, key: (2,) or (N, 2).  Returns (N, 4) uint64."""
    c = [np.asarray(ctr, np.uint64)[:, i].copy() for i in range(4)]
    key = np.asarray(key, np.uint64)
    k0 = np.broadcast_to(key[..., 0], c[0].shape).copy()
    k1 = np.broadcast_to(key[..., 1], c[0].shape).copy()
    for _ in range(10):
        p0 = PHILOX_M[0] * c[0]
        p1 = PHILOX_M[1] * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK32
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
        k0 = (k0 + PHILOX_W[0]) & MASK32
        k1 = (k1 + PHILOX_W[1]) & MASK32
    return np.stack(c, axis=1)


def u01(hi, lo):
    m = ((np.asarray(hi, np.uint64) << np.uint64(21)) ^ (np.asarray(lo, np.uint64) >> np.uint64(11)))
    m &= np.uint64((1 << 53) - 1)
    return (m.astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def uniforms(seed, stream, idx, count):
    """(len(idx), count) uniforms: block b of draw n has counter (n_lo, n_hi, stream, b)."""
    idx = np.asarray(idx, np.uint64)
    nblk = (count + 1) // 2
    out = np.empty((idx.size, 2 * nblk))
    key = np.array([seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF], np.uint64)
    for b in range(nblk):
        ctr = np.stack([idx & MASK32, idx >> np.uint64(32), np.full(idx.shape, stream, np.uint64),
                        np.full(idx.shape, b, np.uint64)], axis=1)
        r = philox4x32_10(ctr, key)
        out[:, 2 * b] = u01(r[:, 0], r[:, 1])
        out[:, 2 * b + 1] = u01(r[:, 2], r[:, 3])
    return out[:, :count]


# sampling ranges of the batched inversion workload
R_MIN, R_MAX, TH_PAD = 1.3, 100.0, 0.05
GAMMA_MAX = 50.0
SIG_MIN, SIG_MAX = 1e-4, 1e2
UOR_MIN, UOR_MAX = 1e-4, 1e1
JITTER = 0.1


def ks_geom(a, r, th):
    return pyoracle.ks_geom_rth(a, r, th)


def gcov_of(g42):
    return np.asarray(g42[16:32]).reshape(4, 4)


def gcon_of(g42):
    return np.asarray(g42[0:16]).reshape(4, 4)


def bsq_of(g42, prim, gam):
    """b^2 of primitives prim in the record g42 (get_state(), src/phys.c)."""
    gcov = gcov_of(g42)
    alpha = g42[34]
    ut = np.array([0.0, prim[2], prim[3], prim[4]])
    utsq = ut[1:] @ gcov[1:, 1:] @ ut[1:]
    gamma = np.sqrt(1.0 + utsq)
    ncon = np.asarray(g42[38:42])
    ucon = gamma * ncon + ut
    ucov = gcov @ ucon
    B = np.array([0.0, prim[5], prim[6], prim[7]])
    bt = B @ ucov
    bcon = np.empty(4)
    bcon[0] = bt
    bcon[1:] = (B[1:] + bt * ucon[1:]) / ucon[0]
    return float(bcon @ gcov @ bcon), float(gamma), alpha


def samples(seed, idx, a, gam):
    """The sampler restated: (U[8][m], rth[2][m], prim[8][m], guess[8][m]) of draws idx."""
    idx = np.asarray(idx)
    u = uniforms(seed, 0, idx, 14)
    m = idx.size
    U = np.empty((8, m))
    rth = np.empty((2, m))
    P = np.empty((8, m))
    G = np.empty((8, m))
    for z in range(m):
        uz = u[z]
        r = np.exp(np.log(R_MIN) + uz[0] * (np.log(R_MAX) - np.log(R_MIN)))
        th = TH_PAD + uz[1] * (np.pi - 2 * TH_PAD)
        gamma = np.exp(uz[2] * np.log(GAMMA_MAX))
        sig = np.exp(np.log(SIG_MIN) + uz[3] * (np.log(SIG_MAX) - np.log(SIG_MIN)))
        uor = np.exp(np.log(UOR_MIN) + uz[4] * (np.log(UOR_MAX) - np.log(UOR_MIN)))
        cv, pv = 2 * uz[5] - 1, 2 * np.pi * uz[6]
        cb, pb = 2 * uz[7] - 1, 2 * np.pi * uz[8]
        g42 = ks_geom(a, r, th)
        L = np.linalg.cholesky(gcov_of(g42)[1:, 1:])
        ev = np.array([np.sqrt(max(0.0, 1 - cv * cv)) * np.cos(pv), np.sqrt(max(0.0, 1 - cv * cv)) * np.sin(pv), cv])
        eb = np.array([np.sqrt(max(0.0, 1 - cb * cb)) * np.cos(pb), np.sqrt(max(0.0, 1 - cb * cb)) * np.sin(pb), cb])
        xv = np.linalg.solve(L.T, ev)
        xb = np.linalg.solve(L.T, eb)
        prim = np.zeros(8)
        prim[0], prim[1] = 1.0, uor
        prim[2:5] = np.sqrt((gamma - 1) * (gamma + 1)) * xv
        prim[5:8] = xb
        bsq, _, _ = bsq_of(g42, prim, gam)
        prim[5:8] *= np.sqrt(sig * prim[0] / bsq)
        U[:, z] = pyoracle.prim_to_flux(g42, gam, prim, 0)
        rth[:, z] = (r, th)
        P[:, z] = prim
        G[:5, z] = prim[:5] * (1 + JITTER * (2 * uz[9:14] - 1))
        G[5:, z] = prim[5:]
    return U, rth, P, G


def utoprim_zones(a, gam, U, rth, guess):
    """Oracle Utoprim_2d_fast() per zone: (prim[8][m], code[m], iters[m])."""
    m = U.shape[1]
    P = np.empty((8, m))
    code = np.empty(m, np.int32)
    iters = np.empty(m, np.int32)
    for z in range(m):
        g42 = ks_geom(a, rth[0, z], rth[1, z])
        ret, prim, _, _, it = pyoracle.utoprim_2d(g42, gam, U[:, z], guess[:, z])
        P[:, z] = prim
        code[z] = ret
        iters[z] = it
    return P, code, iters


# ---------------------------------------------------------------------------
# the guess network, float64
N_IN, N_H, N_OUT = 8, 64, 2
OFF_MU, OFF_ISD, OFF_W1 = 0, 8, 16
OFF_B1 = OFF_W1 + N_IN * N_H
OFF_W2 = OFF_B1 + N_H
OFF_B2 = OFF_W2 + N_H * N_H
OFF_W3 = OFF_B2 + N_H
OFF_B3 = OFF_W3 + N_H * N_OUT
OFF_YMU = OFF_B3 + N_OUT
OFF_YSD = OFF_YMU + N_OUT
N_PARAM = OFF_YSD + N_OUT
TRAIN_OFF, N_TRAIN = OFF_W1, OFF_YMU - OFF_W1
GELU_K, GELU_C = 1.5957691216057308, 0.044715


def unpack(w):
    w = np.asarray(w, np.float64)
    return dict(mu=w[OFF_MU:OFF_MU + N_IN], isd=w[OFF_ISD:OFF_ISD + N_IN],
                W1=w[OFF_W1:OFF_B1].reshape(N_IN, N_H), b1=w[OFF_B1:OFF_W2],
                W2=w[OFF_W2:OFF_B2].reshape(N_H, N_H), b2=w[OFF_B2:OFF_W3],
                W3=w[OFF_W3:OFF_B3].reshape(N_H, N_OUT), b3=w[OFF_B3:OFF_YMU],
                ymu=w[OFF_YMU:OFF_YSD], ysd=w[OFF_YSD:OFF_YSD + N_OUT])


def gelu(z):
    return z / (1.0 + np.exp(-GELU_K * (z + GELU_C * z ** 3)))


def gelu_grad(z):
    s = 1.0 / (1.0 + np.exp(-GELU_K * (z + GELU_C * z ** 3)))
    return s + z * s * (1 - s) * GELU_K * (1 + 3 * GELU_C * z * z)


def mlp_forward(w, feat):
    """feat[8][n] -> y[2][n] (de-normalised outputs)."""
    p = unpack(w)
    x = (np.asarray(feat, np.float64).T - p["mu"]) * p["isd"]
    h1 = gelu(x @ p["W1"] + p["b1"])
    h2 = gelu(h1 @ p["W2"] + p["b2"])
    o = h2 @ p["W3"] + p["b3"]
    return (o * p["ysd"] + p["ymu"]).T


def mlp_loss(w, feat, target):
    p = unpack(w)
    x = (np.asarray(feat, np.float64).T - p["mu"]) * p["isd"]
    o = gelu(gelu(x @ p["W1"] + p["b1"]) @ p["W2"] + p["b2"]) @ p["W3"] + p["b3"]
    tn = (np.asarray(target, np.float64).T - p["ymu"]) / p["ysd"]
    return 0.5 * np.sum((o - tn) ** 2) / x.shape[0]


def mlp_loss_grad(w, feat, target):
    """Mean loss 0.5|o - t_n|^2 and its gradient over the trainable slots."""
    p = unpack(w)
    x = (np.asarray(feat, np.float64).T - p["mu"]) * p["isd"]
    n = x.shape[0]
    z1 = x @ p["W1"] + p["b1"]
    h1 = gelu(z1)
    z2 = h1 @ p["W2"] + p["b2"]
    h2 = gelu(z2)
    o = h2 @ p["W3"] + p["b3"]
    tn = (np.asarray(target, np.float64).T - p["ymu"]) / p["ysd"]
    dy = (o - tn) / n
    loss = 0.5 * np.sum((o - tn) ** 2) / n
    gW3 = h2.T @ dy
    gb3 = dy.sum(0)
    dz2 = (dy @ p["W3"].T) * gelu_grad(z2)
    gW2 = h1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ p["W2"].T) * gelu_grad(z1)
    gW1 = x.T @ dz1
    gb1 = dz1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
