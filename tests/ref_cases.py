"""The reference validation suite (test/test_speigs.c, 13 tests / 73 checks),
restated as functions over a `solve` callable so the same cases run against
the CPU oracle (tests/test_oracle.py) and the MI355X library
(tests/test_gpu_sm.py, tests/test_gpu_lm.py, tests/test_gpu_chosen.py).  Tolerances are the
reference's own.

solve(A, k, which, opts=None, vectors=True) -> speigs_amd.Result
solve_z(Ap, Ai, Axr, Axi, k, which) -> speigs_amd.Result
"""
import math

import numpy as np

from conftest import load_pkg

gen = load_pkg().gen


def check_dirichlet(solve, N=16, k=6):
    """test_dirichlet (test/test_speigs.c:65-88)."""
    A = gen.lap3d(N, True)
    ex = gen.exact_spectrum3(N, True)
    r = solve(A, k, "sm")
    assert r.rc == 0
    for i in range(k):
        assert abs(r.lam[i] - ex[i]) <= 1e-9 * max(1.0, abs(ex[i])), (i, r.lam[i], ex[i])
    assert r.resid.max() <= 1e-8
    assert r.info["flag"] == 0
    return r


def check_neumann_singular(solve, N=16, k=5, shift=True):
    """test_neumann_singular (test/test_speigs.c:90-113).  shift=False drops the
    test's check that the factor needed a shift (info.shift > 0): an FP32
    factor can pass its pivot test at delta = 0 with a rounding-level last
    pivot, and is then only the preconditioner of the block eigensolver."""
    A = gen.lap3d(N, False)
    ex = gen.exact_spectrum3(N, False)
    r = solve(A, k, "sm")
    assert r.rc == 0
    if shift:
        assert r.info["shift"] > 0.0
    assert r.info["flag"] == 0
    assert r.info["nops"] < 400
    assert abs(r.lam[0]) <= 1e-9
    for i in range(1, k):
        assert abs(r.lam[i] - ex[i]) <= 1e-7 * max(1.0, abs(ex[i]))
    v0 = r.V[:, 0]
    assert abs(v0.max() - v0.min()) <= 1e-7
    return r


def check_multiplicity(solve, N=16):
    """test_multiplicity (test/test_speigs.c:115-141)."""
    A = gen.lap3d(N, True)
    ex = gen.exact_spectrum3(N, True)
    r = solve(A, 6, "sm")
    assert abs(ex[1] - ex[2]) < 1e-12 and abs(ex[2] - ex[3]) < 1e-12
    for i in (1, 2, 3):
        assert abs(r.lam[i] - ex[i]) <= 1e-9 * max(1.0, abs(ex[i]))
    worst = 0.0
    for a in (1, 2, 3):
        for b in range(a + 1, 4):
            worst = max(worst, abs(float(r.V[:, a] @ r.V[:, b])))
    assert worst <= 1e-8
    return r


def check_lm(solve, N=16, k=4):
    """test_lm (test/test_speigs.c:143-158)."""
    A = gen.lap3d(N, True)
    ex = gen.exact_spectrum3(N, True)
    r = solve(A, k, "lm", vectors=False)
    assert r.info["path"] == 3
    for i in range(k):
        want = ex[-1 - i]
        assert abs(r.lam[i] - want) <= 1e-9 * abs(want), (i, r.lam[i], want)
    return r


def check_diagonal(solve):
    """test_diagonal (test/test_speigs.c:160-175)."""
    n = 5000
    i = np.arange(n)
    Ap = np.arange(n + 1, dtype=np.int64)
    Ai = i.astype(np.int64)
    Ax = (i + 1.0) * np.where(i % 2 == 1, 1.0, -1.0)
    r = solve((Ap, Ai, Ax), 4, "sm", vectors=False)
    assert r.info["path"] == 0
    assert abs(abs(r.lam[0]) - 1.0) < 1e-12
    assert abs(abs(r.lam[3]) - 4.0) < 1e-12
    return r


def _shift_diag(A, c):
    Ap, Ai, Ax = (np.array(a) for a in A)
    col = np.repeat(np.arange(len(Ap) - 1), np.diff(Ap))
    Ax = Ax.copy()
    Ax[Ai == col] -= c
    return Ap, Ai, Ax


def check_indefinite(solve, N=12, k=4):
    """test_indefinite (test/test_speigs.c:177-205)."""
    ex = gen.exact_spectrum3(N, True)
    cnt = len(ex)
    mid = cnt // 2
    while mid + 1 < cnt and ex[mid + 1] - ex[mid] < 1e-6:
        mid += 1
    c = 0.5 * (ex[mid] + ex[mid + 1])
    A = _shift_diag(gen.lap3d(N, True), c)
    byabs = np.sort(np.abs(ex - c))
    r = solve(A, k, "sm", vectors=False)
    assert r.info["path"] == 5
    assert r.info["flag"] == 0
    for i in range(k):
        assert abs(abs(r.lam[i]) - byabs[i]) <= 1e-7 * max(1.0, byabs[i])
    return r


def check_pathological(solve, N=12, k=3):
    """test_pathological (test/test_speigs.c:211-227)."""
    ex = gen.exact_spectrum3(N, True)
    c = ex[len(ex) // 2]
    A = _shift_diag(gen.lap3d(N, True), c)
    r = solve(A, k, "sm", opts={"maxit": 40}, vectors=False)
    assert r.rc == 0
    assert abs(r.lam[0]) < 1e-6
    assert r.info["flag"] == 1
    return r


def check_blocks(solve, N1=12, N2=10, k=6):
    """test_blocks (test/test_speigs.c:231-272)."""
    A = gen.block_diag([gen.lap3d(N1, True), gen.lap3d(N2, True)])
    n1 = N1 ** 3
    allx = np.sort(np.concatenate([gen.exact_spectrum3(N1), gen.exact_spectrum3(N2)]))
    r = solve(A, k, "sm")
    assert r.info["path"] == 4
    assert r.info["ncomp"] == 2
    for i in range(k):
        assert abs(r.lam[i] - allx[i]) <= 1e-9 * max(1.0, abs(allx[i]))
    for t in range(k):
        a = float(np.sum(r.V[:n1, t] ** 2))
        b = float(np.sum(r.V[n1:, t] ** 2))
        assert min(a, b) <= 1e-16
    return r


def tridiag_1d(n, diag=2.0, off=-1.0):
    Ap = [0]
    Ai = []
    Ax = []
    for c in range(n):
        if c > 0:
            Ai.append(c - 1); Ax.append(off)
        Ai.append(c); Ax.append(diag)
        if c < n - 1:
            Ai.append(c + 1); Ax.append(off)
        Ap.append(len(Ai))
    return np.array(Ap, np.int64), np.array(Ai, np.int64), np.array(Ax, float)


def check_banded(solve, n=20000, k=4):
    """test_banded (test/test_speigs.c:275-296)."""
    A = tridiag_1d(n)
    r = solve(A, k, "sm", vectors=False)
    assert r.info["path"] == 6
    for i in range(k):
        want = 2.0 - 2.0 * math.cos((i + 1) * math.pi / (n + 1))
        assert abs(r.lam[i] - want) <= 1e-9 * max(1e-6, abs(want))
    return r


def check_isolated(solve):
    """test_isolated (test/test_speigs.c:299-325)."""
    m = 2000
    Ap, Ai, Ax = (list(a) for a in tridiag_1d(m))
    for v in (1e-7, -5e-8, 3e-7):
        Ai.append(len(Ap) - 1); Ax.append(v); Ap.append(len(Ai))
    Ai.append(len(Ap) - 1); Ax.append(0.0); Ap.append(len(Ai))
    A = (np.array(Ap, np.int64), np.array(Ai, np.int64), np.array(Ax, float))
    r = solve(A, 4, "sm", vectors=False, resid=False)
    assert r.info["ncomp"] == 5
    assert abs(r.lam[0]) < 1e-14
    assert abs(r.lam[1] + 5e-8) < 1e-18
    assert abs(r.lam[2] - 1e-7) < 1e-18
    assert abs(r.lam[3] - 3e-7) < 1e-18
    return r


def check_hermitian(solve_z, N=12, k=5):
    """test_hermitian (test/test_speigs.c:331-389)."""
    Ap, Ai, Ax = gen.lap3d(N, True)
    n = N ** 3
    ex = gen.exact_spectrum3(N, True)
    i = np.arange(n, dtype=np.int64)
    th = 2.399963 * ((i * 2654435761) % 1000).astype(float) / 1000.0   # int64 product, as in C
    col = np.repeat(np.arange(n), np.diff(Ap))
    d = th[Ai] - th[col]
    Ar = Ax * np.cos(d)
    Aim = Ax * np.sin(d)
    r = solve_z(Ap, Ai, Ar, Aim, k, "sm")
    assert r.rc == 0
    for t in range(k):
        assert abs(r.lam[t] - ex[t]) <= 1e-8 * max(1.0, abs(ex[t]))
    Ac = np.zeros((n, n), complex)
    Ac[Ai, col] = Ar + 1j * Aim
    worst = max(np.linalg.norm(Ac @ r.V[:, t] - r.lam[t] * r.V[:, t]) for t in range(k))
    assert worst <= 1e-8
    ov = max(abs(np.vdot(r.V[:, a], r.V[:, b])) for a in range(k) for b in range(a + 1, k))
    assert ov <= 1e-7
    return r


def check_nonsymmetric(solve):
    """test_nonsymmetric (test/test_speigs.c:393-416)."""
    n = 5000
    Ap, Ai, Ax = tridiag_1d(n, 5.0, -1.0)
    col = np.repeat(np.arange(n), np.diff(Ap))
    Ax = np.where(Ai > col, -4.0, Ax)         # sub != super
    r = solve((Ap, Ai, Ax), 3, "sm", vectors=False)
    assert r.rc == -6
    r2 = solve((Ap, Ai, Ax), 3, "sm", opts={"check_sym": 0}, vectors=False)
    assert abs(r2.lam[0] - 3.0) < 1e-3
    return r2


def check_upper_only(solve, N=12, k=4):
    """test_upper_only (test/test_speigs.c:419-439)."""
    A = gen.upper_only(*gen.lap3d(N, True))
    ex = gen.exact_spectrum3(N, True)
    r = solve(A, k, "sm", vectors=False)
    assert r.rc == 0
    for i in range(k):
        assert abs(r.lam[i] - ex[i]) <= 1e-9 * max(1.0, abs(ex[i]))
    return r


def subspace_angle_max(V, U):
    """Largest principal angle (radians) between span(V) and span(U)."""
    qv, _ = np.linalg.qr(V)
    qu, _ = np.linalg.qr(U)
    # sine form: arccos of the cosines loses half the digits near zero angle
    resid = qv - qu @ (qu.T @ qv)
    return float(np.arcsin(min(1.0, np.linalg.norm(resid, 2))))


def dirichlet_eigenspace(N, lam, tol=1e-9):
    """Orthonormal basis of the closed-form eigenspace of the N^3 Dirichlet
    Laplacian for eigenvalue lam (all separable sine modes with that value)."""
    e1 = 2.0 - 2.0 * np.cos(np.arange(1, N + 1) * np.pi / (N + 1))
    cols = []
    for p in range(N):
        for q in range(N):
            for s in range(N):
                if abs(e1[p] + e1[q] + e1[s] - lam) < tol:
                    cols.append(gen.dirichlet_mode3(N, p + 1, q + 1, s + 1))
    return np.stack(cols, axis=1)
