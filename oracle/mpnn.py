"""CPU oracle of the learned-start network (test infrastructure only: tests/
import it as the checker of the device kernels; nothing in the product calls it).

The network has no counterpart in the reference -- its plan names a learned
starting block as future work (docs/REQUIREMENTS.md) -- so this module is the
definition the device code (sparse-eigs-mex_amd/csrc/mpnn.hip) is checked
against, and its own gradient is pinned by finite differences
(tests/test_mpnn_oracle.py).  Float64 throughout; the weights enter the
forward pass rounded to BF16 as on the device.

  A_hat = D^-1 |A_off|                      (graph operator, row i: |a_ij| / a_ii)
  H0    = features (n x 32)                 (log a_ii - mean, log s_i - mean,
                                             (a_ii - s_i) / a_ii, 8 N(0,1) noise, 4 noise
                                             channels diffused by A_hat^8 and by A_hat^64,
                                             each scaled to unit RMS, zeros)
  M_l   = A_hat H_l,  Z_l = [H_l | M_l] W_l^T + b_l,  H_{l+1} = relu(Z_l) (last: Z_l)
  L     = 1 - tr(S^-1 X^T U U^T X) / k,  S = X^T X + ridge I,  X = H_L
"""
from __future__ import annotations

import numpy as np

WIDTH = 32
NOISE = 16
KIN = 2 * WIDTH
LAYER_P = WIDTH * KIN + WIDTH


def bf16_round(p: np.ndarray) -> np.ndarray:
    """Round float32 values to BF16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(p, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray, layers: int):
    out = []
    for l in range(layers):
        blk = params[l * LAYER_P:(l + 1) * LAYER_P]
        out.append((blk[:WIDTH * KIN].reshape(WIDTH, KIN), blk[WIDTH * KIN:]))
    return out


def graph(Ap, Ai, Ax):
    """Full symmetric |A_off| as a dense-free edge list (rows, cols, |a|) and d, s."""
    n = len(Ap) - 1
    col = np.repeat(np.arange(n), np.diff(Ap))
    up = Ai <= col
    r, c, a = Ai[up], col[up], Ax[up]
    diag = r == c
    d = np.zeros(n)
    d[c[diag]] = a[diag]
    off = ~diag
    r, c, w = r[off], c[off], np.abs(a[off])
    rows = np.concatenate([r, c])
    cols = np.concatenate([c, r])
    vals = np.concatenate([w, w])
    s = np.zeros(n)
    np.add.at(s, rows, vals)
    return rows, cols, vals, d, s


def _spmv(rows, cols, vals, H, n):
    out = np.zeros((n, H.shape[1]))
    np.add.at(out, rows, vals[:, None] * H[cols])
    return out


def features(Ap, Ai, Ax, seed: int, gen):
    rows, cols, vals, d, s = graph(Ap, Ai, Ax)
    n = len(Ap) - 1
    H = np.zeros((n, WIDTH))
    ld = np.where(d > 0, np.log(np.where(d > 0, d, 1.0)), 0.0)
    ls = np.where(s > 0, np.log(np.where(s > 0, s, 1.0)), 0.0)
    md = ld[d > 0].mean() if np.any(d > 0) else 0.0
    ms = ls[s > 0].mean() if np.any(s > 0) else 0.0
    H[:, 0] = np.where(d > 0, ld - md, 0.0)
    H[:, 1] = np.where(s > 0, ls - ms, 0.0)
    H[:, 2] = np.where(d > 0, (d - s) / np.where(d > 0, d, 1.0), 0.0)
    ids = (np.arange(n, dtype=np.int64)[:, None] * NOISE + np.arange(NOISE)[None, :]).ravel()
    z = gen.normal01(seed, ids).reshape(n, NOISE)
    H[:, 3:11] = z[:, 0:8]
    dinv = np.where(d > 0, 1.0 / np.where(d > 0, d, 1.0), 0.0)
    Q = z[:, 8:12].copy()
    done = 0
    for target, c0 in ((8, 11), (64, 15)):
        while done < target:
            Q = dinv[:, None] * _spmv(rows, cols, vals, Q, n)
            done += 1
        rms = np.sqrt(np.mean(Q * Q, axis=0))
        H[:, c0:c0 + 4] = Q / np.where(rms > 0, rms, 1.0)
    return H.astype(np.float32).astype(np.float64)


def forward(Wb, A, H0):
    """Wb: list of (W, b) float64 (already BF16-rounded W); returns Hs, Ms."""
    Ap, Ai, Ax = A
    rows, cols, vals, d, s = graph(Ap, Ai, Ax)
    n = len(Ap) - 1
    dinv = np.where(d > 0, 1.0 / np.where(d > 0, d, 1.0), 0.0)
    Hs, Ms = [H0], []
    for l, (W, b) in enumerate(Wb):
        M = dinv[:, None] * _spmv(rows, cols, vals, Hs[-1], n)
        Ms.append(M)
        Z = np.concatenate([Hs[-1], M], axis=1) @ W.T + b
        Hs.append(np.maximum(Z, 0.0) if l + 1 < len(Wb) else Z)
    return Hs, Ms


def ridge_of(X):
    """1e-6 times the mean diagonal of X^T X: a constant of the loss (its
    dependence on X is not differentiated, here or on the device)."""
    t = float(np.sum(X * X))
    return 1e-6 * (t / WIDTH if t > 0 else 1.0)


def loss_and_dX(X, U, ridge=None):
    k = U.shape[1]
    S = X.T @ X
    if ridge is None:
        ridge = ridge_of(X)
    S = S + ridge * np.eye(WIDTH)
    Si = np.linalg.inv(S)
    B = X.T @ U
    C = B @ B.T
    f = np.trace(Si @ C)
    dX = -(2.0 / k) * (U @ B.T @ Si - X @ Si @ C @ Si)
    return 1.0 - f / k, dX


def loss_grad(Wb, A, H0, U, ridge=None):
    """loss and [(dW, db)] per layer for weights Wb (float64)."""
    Ap, Ai, Ax = A
    rows, cols, vals, d, s = graph(Ap, Ai, Ax)
    n = len(Ap) - 1
    dinv = np.where(d > 0, 1.0 / np.where(d > 0, d, 1.0), 0.0)
    Hs, Ms = forward(Wb, A, H0)
    L, dZ = loss_and_dX(Hs[-1], U, ridge)
    grads = [None] * len(Wb)
    for l in range(len(Wb) - 1, -1, -1):
        W, _ = Wb[l]
        In = np.concatenate([Hs[l], Ms[l]], axis=1)
        grads[l] = (dZ.T @ In, dZ.sum(axis=0))
        if l == 0:
            break
        dIn = dZ @ W
        dH = dIn[:, :WIDTH] + _spmv(rows, cols, vals, dinv[:, None] * dIn[:, WIDTH:], n)
        dZ = dH * (Hs[l] > 0)
    return L, grads


def flat(grads):
    return np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])
