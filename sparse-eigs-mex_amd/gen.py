"""Synthetic matrices for the speigs workloads (int64 CSC, sorted rows, FP64).

Node order is x-fastest, c = x + N*(y + N*z), and a full-storage column lists
its entries as (z-1, y-1, x-1, diagonal, x+1, y+1, z+1) -- the order of the
reference test builder lap3d (test/test_speigs.c:31-47), so column-major CSC
rows come out sorted.

Random edge weights come from a counter-based hash (splitmix64 of seed and
edge id), so any subset of a matrix can be regenerated in any order, by numpy
here or by C, and agree bit for bit.
"""
from __future__ import annotations

import math

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser (Steele, Lea and Flood 2014) on a uint64 array."""
    with np.errstate(over="ignore"):
        z = (x + GOLDEN).astype(np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, ids: np.ndarray) -> np.ndarray:
    """U[0,1) doubles: top 53 bits of splitmix64(seed * 2^32 + id)."""
    with np.errstate(over="ignore"):
        key = np.uint64(seed & 0xFFFFFFFF) * np.uint64(1 << 32) + ids.astype(np.uint64)
    return (splitmix64(key) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal01(seed: int, ids: np.ndarray) -> np.ndarray:
    """N(0,1) by Box-Muller over two counter streams."""
    u1 = uniform01(seed, 2 * ids)
    u2 = uniform01(seed, 2 * ids + 1)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


# --------------------------------------------------------------------- stencils
def _grid3(N: int):
    n = N * N * N
    c = np.arange(n, dtype=np.int64)
    x = c % N
    y = (c // N) % N
    z = c // (N * N)
    return n, c, x, y, z


def _assemble(n: int, rows: np.ndarray, vals: np.ndarray, valid: np.ndarray):
    """rows/vals/valid are (n, s) slot arrays already in sorted column order."""
    cnt = valid.sum(axis=1)
    Ap = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(cnt, out=Ap[1:])
    return Ap, rows[valid].astype(np.int64), vals[valid].astype(np.float64)


def lap3d(N: int, dirichlet: bool = True):
    """3-D 7-point Laplacian, full storage: Dirichlet diag 6, Neumann diag = degree."""
    n, c, x, y, z = _grid3(N)
    nb = np.stack([c - N * N, c - N, c - 1, c, c + 1, c + N, c + N * N], axis=1)
    valid = np.stack([z > 0, y > 0, x > 0, np.ones(n, bool), x < N - 1, y < N - 1, z < N - 1], axis=1)
    vals = np.full((n, 7), -1.0)
    deg = valid.sum(axis=1) - 1
    vals[:, 3] = 6.0 if dirichlet else deg.astype(np.float64)
    return _assemble(n, nb, vals, valid)


def lap2d(M: int, dirichlet: bool = True):
    """2-D 5-point Laplacian, full storage, x-fastest."""
    n = M * M
    c = np.arange(n, dtype=np.int64)
    x = c % M
    y = c // M
    nb = np.stack([c - M, c - 1, c, c + 1, c + M], axis=1)
    valid = np.stack([y > 0, x > 0, np.ones(n, bool), x < M - 1, y < M - 1], axis=1)
    vals = np.full((n, 5), -1.0)
    vals[:, 2] = 4.0 if dirichlet else (valid.sum(axis=1) - 1).astype(np.float64)
    return _assemble(n, nb, vals, valid)


def upper_only(Ap, Ai, Ax):
    """Keep entries with row <= column (the reference test_upper_only strip)."""
    n = len(Ap) - 1
    col = np.repeat(np.arange(n, dtype=np.int64), np.diff(Ap))
    keep = Ai <= col
    cnt = np.bincount(col[keep], minlength=n)
    Bp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(cnt, out=Bp[1:])
    return Bp, Ai[keep].copy(), Ax[keep].copy()


def exact_spectrum3(N: int, dirichlet: bool = True) -> np.ndarray:
    """Sorted closed-form spectrum: sums over axes of 2-2cos(p pi/(N+1)) (Dirichlet)
    or 2-2cos(p pi/N), p from 0 (Neumann)."""
    p = np.arange(N, dtype=np.float64)
    e1 = 2.0 - 2.0 * np.cos((p + 1) * np.pi / (N + 1)) if dirichlet else 2.0 - 2.0 * np.cos(p * np.pi / N)
    return np.sort((e1[:, None, None] + e1[None, :, None] + e1[None, None, :]).ravel())


def dirichlet_mode3(N: int, p: int, q: int, r: int) -> np.ndarray:
    """Separable eigenvector sin(p)(x) sin(q)(y) sin(r)(z), 1-based wave numbers,
    normalised; eigenvalue sum of 2-2cos(k pi/(N+1))."""
    t = np.arange(1, N + 1, dtype=np.float64)
    sx = np.sin(p * np.pi * t / (N + 1))
    sy = np.sin(q * np.pi * t / (N + 1))
    sz = np.sin(r * np.pi * t / (N + 1))
    v = (sz[:, None, None] * sy[None, :, None] * sx[None, None, :]).ravel()
    return v / np.linalg.norm(v)


# ------------------------------------------------------- weighted graph Laplacian
def weighted_lap3d(N: int, seed: int, wlo: float = 0.1, whi: float = 10.0, shift: float = 1e-3,
                   upper: bool = True, weights: np.ndarray | None = None):
    """Weighted 3-D graph Laplacian plus shift*I: A = L_w + shift*I.

    Edge e between neighbours gets weight exp(U[ln wlo, ln whi]) from the counter
    stream (seed, edge id); edge ids are x-edges, then y-edges, then z-edges, each
    in node order of their lower endpoint.  Diagonal = sum of incident weights +
    shift, so the constant vector is an eigenvector with eigenvalue `shift`.
    `weights`, if given, is a (3, n) array of per-node forward-edge weights that
    replaces the random ones (used by the Gaussian-random-field family).
    """
    n, c, x, y, z = _grid3(N)
    if weights is None:
        ids = np.arange(3 * n, dtype=np.int64).reshape(3, n)
        u = uniform01(seed, ids.ravel()).reshape(3, n)
        w = np.exp(math.log(wlo) + u * (math.log(whi) - math.log(wlo)))
    else:
        w = np.asarray(weights, dtype=np.float64)
    has = np.stack([x < N - 1, y < N - 1, z < N - 1])           # forward edge exists
    wf = np.where(has, w, 0.0)                                   # forward weights per node
    # backward weights: the edge from the lower neighbour
    wb = np.zeros_like(wf)
    wb[0, 1:] = wf[0, :-1]
    wb[1, N:] = wf[1, :-N]
    wb[2, N * N:] = wf[2, :-N * N]
    diag = wf.sum(axis=0) + wb.sum(axis=0) + shift
    if upper:
        nb = np.stack([c - N * N, c - N, c - 1, c], axis=1)
        valid = np.stack([z > 0, y > 0, x > 0, np.ones(n, bool)], axis=1)
        vals = np.stack([-wb[2], -wb[1], -wb[0], diag], axis=1)
    else:
        nb = np.stack([c - N * N, c - N, c - 1, c, c + 1, c + N, c + N * N], axis=1)
        valid = np.stack([z > 0, y > 0, x > 0, np.ones(n, bool), x < N - 1, y < N - 1, z < N - 1], axis=1)
        vals = np.stack([-wb[2], -wb[1], -wb[0], diag, -wf[0], -wf[1], -wf[2]], axis=1)
    return _assemble(n, nb, vals, valid)


def grf_weights(N: int, seed: int, corr: float = 8.0, logstd: float = 1.0) -> np.ndarray:
    """exp of a Gaussian random field (Gaussian covariance, correlation length
    `corr` cells, log-std `logstd`) sampled per node; a forward edge takes the
    geometric mean of its endpoints' coefficients.  Returns (3, n)."""
    n = N * N * N
    g = normal01(seed, np.arange(n, dtype=np.int64)).reshape(N, N, N)
    k = np.fft.fftfreq(N)
    kk = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    filt = np.exp(-0.5 * (2 * np.pi) ** 2 * kk * corr * corr / 2.0)
    f = np.real(np.fft.ifftn(np.fft.fftn(g) * filt))
    f = (f - f.mean()) / (f.std() + 1e-300) * logstd
    a = np.exp(f)                                   # [z, y, x]
    w = np.zeros((3, N, N, N))
    w[0, :, :, :-1] = np.sqrt(a[:, :, :-1] * a[:, :, 1:])
    w[1, :, :-1, :] = np.sqrt(a[:, :-1, :] * a[:, 1:, :])
    w[2, :-1, :, :] = np.sqrt(a[:-1, :, :] * a[1:, :, :])
    return w.reshape(3, n)


def grf_lap3d(N: int, seed: int, corr: float = 8.0, logstd: float = 1.0, shift: float = 1e-3, upper: bool = True):
    """The learned-start family: 3-D graph Laplacian on an N^3 grid whose edge
    coefficients are exp of a seeded Gaussian random field (correlation length
    `corr` cells, log-std `logstd`), plus shift*I (constant eigenvector, eigenvalue
    `shift`, as in the weighted family)."""
    return weighted_lap3d(N, seed, shift=shift, upper=upper, weights=grf_weights(N, seed, corr, logstd))


# ------------------------------------------------------------- batched blocks
def block_diag(blocks):
    """Assemble CSC blocks (Ap, Ai, Ax) block-diagonally."""
    n = sum(len(b[0]) - 1 for b in blocks)
    nnz = sum(int(b[0][-1]) for b in blocks)
    Ap = np.zeros(n + 1, dtype=np.int64)
    Ai = np.empty(nnz, dtype=np.int64)
    Ax = np.empty(nnz, dtype=np.float64)
    r0 = 0
    z0 = 0
    for bp, bi, bx in blocks:
        nb = len(bp) - 1
        nz = int(bp[-1])
        Ap[r0 + 1:r0 + nb + 1] = bp[1:] + z0
        Ai[z0:z0 + nz] = bi + r0
        Ax[z0:z0 + nz] = bx
        r0 += nb
        z0 += nz
    return Ap, Ai, Ax


def weighted_grid(dims, seed: int, wlo: float, whi: float, ghost: bool = True):
    """Weighted 2-D or 3-D grid Laplacian, full storage, with Dirichlet-type ghost
    weights on the boundary so the block is SPD (diag = sum of all incident weights,
    boundary ghosts included)."""
    dims = tuple(int(d) for d in dims)
    n = int(np.prod(dims))
    c = np.arange(n, dtype=np.int64)
    coords = []
    s = 1
    strides = []
    for d in dims:
        coords.append((c // s) % d)
        strides.append(s)
        s *= d
    D = len(dims)
    ids = np.arange(2 * D * n, dtype=np.int64).reshape(2 * D, n)
    u = uniform01(seed, ids.ravel()).reshape(2 * D, n)
    w = np.exp(math.log(wlo) + u * (math.log(whi) - math.log(wlo)))
    wf = np.zeros((D, n))
    wb = np.zeros((D, n))
    diag = np.zeros(n)
    for a in range(D):
        fwd = coords[a] < dims[a] - 1
        wf[a] = np.where(fwd, w[a], 0.0)
        st = strides[a]
        wb[a, st:] = wf[a, :-st]
        if ghost:
            lo = coords[a] == 0
            hi = coords[a] == dims[a] - 1
            diag += np.where(lo, w[D + a], 0.0) + np.where(hi, w[D + a], 0.0)
    diag += wf.sum(axis=0) + wb.sum(axis=0)
    nbs = []
    valids = []
    vals = []
    for a in reversed(range(D)):
        nbs.append(c - strides[a]); valids.append(coords[a] > 0); vals.append(-wb[a])
    nbs.append(c); valids.append(np.ones(n, bool)); vals.append(diag)
    for a in range(D):
        nbs.append(c + strides[a]); valids.append(coords[a] < dims[a] - 1); vals.append(-wf[a])
    return _assemble(n, np.stack(nbs, 1), np.stack(vals, 1), np.stack(valids, 1))


def batched_components(ncomp: int = 512, seed: int = 4, nmin: int = 500, nmax: int = 50_000):
    """Block-diagonal matrix of `ncomp` grid Laplacians with log-uniform sizes;
    even index -> 2-D m x m, odd -> 3-D N^3.  Returns (Ap, Ai, Ax, sizes)."""
    u = uniform01(seed, np.arange(ncomp, dtype=np.int64))
    sizes = np.exp(math.log(nmin) + u * (math.log(nmax) - math.log(nmin)))
    blocks = []
    dims_all = []
    for i, s in enumerate(sizes):
        if i % 2 == 0:
            m = max(2, int(round(math.sqrt(s))))
            dims = (m, m)
        else:
            N = max(2, int(round(s ** (1.0 / 3.0))))
            dims = (N, N, N)
        dims_all.append(dims)
        blocks.append(weighted_grid(dims, seed * 100003 + i + 1, 0.5, 2.0))
    Ap, Ai, Ax = block_diag(blocks)
    return Ap, Ai, Ax, [int(np.prod(d)) for d in dims_all]


def csc_to_dense(Ap, Ai, Ax, n=None, symmetrize_upper: bool = False) -> np.ndarray:
    n = len(Ap) - 1 if n is None else n
    D = np.zeros((n, n))
    col = np.repeat(np.arange(n), np.diff(Ap))
    if symmetrize_upper:
        keep = Ai <= col
        D[Ai[keep], col[keep]] = Ax[keep]
        D[col[keep], Ai[keep]] = Ax[keep]
    else:
        D[Ai, col] = Ax
    return D


def spmv_upper(Ap, Ai, Ax, x):
    """y = A x reading only the upper triangle (numpy, for checks)."""
    n = len(Ap) - 1
    col = np.repeat(np.arange(n, dtype=np.int64), np.diff(Ap))
    keep = Ai <= col
    r, c, v = Ai[keep], col[keep], Ax[keep]
    y = np.zeros((n,) + x.shape[1:])
    np.add.at(y, r, v[:, None] * x[c] if x.ndim > 1 else v * x[c])
    off = r != c
    np.add.at(y, c[off], v[off][:, None] * x[r[off]] if x.ndim > 1 else v[off] * x[r[off]])
    return y
