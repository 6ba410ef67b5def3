"""CPU oracle pinned against the reference's own tests (closed-form spectra),
numpy, and mathematical invariants.  Runs without a GPU."""
import ctypes

import numpy as np
import pytest

import ref_cases as rc
from conftest import ORACLE_SO, load_pkg

pkg = load_pkg()
gen = pkg.gen


@pytest.fixture(scope="module")
def solve(oracle):
    def _solve(A, k, which, opts=None, vectors=True, resid=True):
        return pkg.speigs(A, k, which, opts=opts, vectors=vectors, resid=resid, lib=oracle, prefix="oracle_")
    return _solve


@pytest.fixture(scope="module")
def solve_z(oracle):
    def _solve(Ap, Ai, Axr, Axi, k, which):
        return pkg.speigs_z(Ap, Ai, Axr, Axi, k, which, lib=oracle, prefix="oracle_")
    return _solve


# ---- the 13 reference tests (test/test_speigs.c:441-464), at the reference's sizes
def test_ref1_dirichlet(solve):
    r = rc.check_dirichlet(solve)
    assert r.info["path"] == 2


def test_ref2_neumann_singular(solve):
    rc.check_neumann_singular(solve)


def test_ref3_multiplicity(solve):
    r = rc.check_multiplicity(solve)
    # the triple's computed span equals the closed-form eigenspace
    ex = gen.exact_spectrum3(16, True)
    U = rc.dirichlet_eigenspace(16, ex[1])
    assert rc.subspace_angle_max(r.V[:, 1:4], U) <= 1e-8


def test_ref4_lm(solve):
    rc.check_lm(solve)


def test_ref5_diagonal(solve):
    rc.check_diagonal(solve)


def test_ref6_indefinite(solve):
    rc.check_indefinite(solve)


def test_ref7_pathological(solve):
    rc.check_pathological(solve)


def test_ref8_blocks(solve):
    rc.check_blocks(solve)


def test_ref9_banded(solve):
    rc.check_banded(solve)


def test_ref10_isolated(solve):
    rc.check_isolated(solve)


def test_ref11_hermitian(solve_z):
    rc.check_hermitian(solve_z)


def test_ref12_nonsymmetric(solve):
    rc.check_nonsymmetric(solve)


def test_ref13_upper_only(solve):
    rc.check_upper_only(solve)


# ---- configuration 1: 20^3 Dirichlet, upper-only CSC, k=6 'sm'
def test_config1_closed_form(solve):
    A = gen.upper_only(*gen.lap3d(20, True))
    r = solve(A, 6, "sm")
    want = np.array([0.067015042649, 0.133531083527, 0.133531083527, 0.133531083527,
                     0.200047124405, 0.200047124405])
    assert r.rc == 0 and r.info["flag"] == 0
    assert np.abs(r.lam - want).max() <= 1e-11
    ex = gen.exact_spectrum3(20, True)
    assert np.abs(r.lam - ex[:6]).max() <= 1e-12
    # residual against A computed independently here
    Vs = r.V
    R = gen.spmv_upper(*A, Vs) - Vs * r.lam
    assert np.linalg.norm(R, axis=0).max() <= 1e-10 * 12
    # subspace angles: the 0.1335 triple exactly, the cut 0.2000 pair inside its triple
    assert rc.subspace_angle_max(Vs[:, 1:4], rc.dirichlet_eigenspace(20, ex[1])) <= 1e-8
    U3 = rc.dirichlet_eigenspace(20, ex[4])
    proj = U3 @ (U3.T @ Vs[:, 4:6])
    assert np.linalg.norm(Vs[:, 4:6] - proj) <= 1e-8


# ---- pieces
def _xorshift_np(seed, count):
    s = seed
    out = []
    for _ in range(count):
        x = s
        x ^= x >> 12
        x ^= (x << 25) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 27
        s = x
        out.append((x * 0x2545F4914F6CDD1D) & 0xFFFFFFFFFFFFFFFF)
    return out


def test_rng_known_answers(oracle):
    oracle.oracle_rng_next.restype = ctypes.c_uint64
    oracle.oracle_rng_next.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
    s = ctypes.c_uint64(88172645463325252)
    got = [oracle.oracle_rng_next(ctypes.byref(s)) for _ in range(5)]
    assert got == _xorshift_np(88172645463325252, 5)
    # the default-seed uniform start vector (first entries) in [-0.5, 0.5)
    oracle.oracle_rand_vec.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_int64,
                                       ctypes.POINTER(ctypes.c_double)]
    v = np.zeros(1000)
    s = ctypes.c_uint64(88172645463325252)
    oracle.oracle_rand_vec(ctypes.byref(s), 1000, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    ref = np.array([(x >> 11) / 9007199254740992.0 - 0.5 for x in _xorshift_np(88172645463325252, 1000)])
    assert np.array_equal(v, ref)
    assert v.min() >= -0.5 and v.max() < 0.5


def _gf2_mul(A, B):
    """product of 64 x 64 bit matrices over GF(2), each a list of 64 column bitmasks"""
    out = []
    for col in B:
        acc, i = 0, 0
        while col:
            if col & 1:
                acc ^= A[i]
            col >>= 1
            i += 1
        out.append(acc)
    return out


def _gf2_pow(A, e):
    R = [1 << b for b in range(64)]
    while e:
        if e & 1:
            R = _gf2_mul(R, A)
        A = _gf2_mul(A, A)
        e >>= 1
    return R


def test_rng_state_map_has_full_period(oracle):
    """The oracle's generator state update (reference src/speigs_core.c:49-54,
    xorshift64* after Vigna 2016) is linear over GF(2); what fixes it
    mathematically is that its 64 x 64 bit matrix T has order exactly 2^64 - 1,
    i.e. the state runs through every nonzero value (T^(2^64-1) = I and
    T^((2^64-1)/p) != I for each prime p of 2^64 - 1).  A wrong shift, a
    shift in the wrong direction or a dropped step breaks that.  The
    multiplier of the output step is the published decimal constant."""
    oracle.oracle_rng_next.restype = ctypes.c_uint64
    oracle.oracle_rng_next.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
    T = []
    for b in range(64):
        s = ctypes.c_uint64(1 << b)
        oracle.oracle_rng_next(ctypes.byref(s))
        T.append(s.value)
    primes = [3, 5, 17, 257, 641, 65537, 6700417]
    order = (1 << 64) - 1
    assert np.prod(np.array(primes, dtype=object)) == order
    eye = [1 << b for b in range(64)]
    assert _gf2_pow(T, order) == eye
    for p in primes:
        assert _gf2_pow(T, order // p) != eye, p
    # linearity: the map of a sum of states is the sum of the maps
    s = ctypes.c_uint64(0x9E3779B97F4A7C15)
    oracle.oracle_rng_next(ctypes.byref(s))
    acc = 0
    for b in range(64):
        if (0x9E3779B97F4A7C15 >> b) & 1:
            acc ^= T[b]
    assert s.value == acc
    # output = state * M_32 mod 2^64, M_32 = 2685821657736338717 (Vigna 2016)
    s = ctypes.c_uint64(1)
    out = oracle.oracle_rng_next(ctypes.byref(s))
    assert out == (s.value * 2685821657736338717) % (1 << 64)


@pytest.mark.parametrize("fn", ["oracle_syev", "oracle_syev_jacobi"])
def test_dense_eig_vs_numpy(oracle, fn):
    rng = np.random.default_rng(3)
    f = getattr(oracle, fn)
    f.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    for m in (1, 2, 5, 41, 120):
        B = rng.standard_normal((m, m))
        S = B + B.T
        A = np.asfortranarray(np.triu(S) + np.tril(rng.standard_normal((m, m)), -1))  # lower is junk
        w = np.zeros(m)
        assert f(m, A.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), w.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
        wref = np.linalg.eigvalsh(S)
        assert np.allclose(w, wref, atol=1e-12 * max(1, np.abs(wref).max()))
        assert np.allclose(A.T @ A, np.eye(m), atol=1e-12)
        assert np.allclose(S @ A, A * w, atol=1e-11 * max(1, np.abs(wref).max()))


def test_jacobi_graded_relative_accuracy(oracle):
    """A positive definite matrix with eigenvalues 1e8 .. 1: Jacobi keeps the
    small eigenvalues to relative accuracy (what the Lanczos projected problem
    with theta_0 = 1/delta needs)."""
    f = oracle.oracle_syev_jacobi
    f.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    rng = np.random.default_rng(1)
    m = 20
    Q, _ = np.linalg.qr(rng.standard_normal((m, m)))
    d = np.logspace(0, 8, m)
    D = np.diag(np.sqrt(d))
    # scaled-diagonally-dominant: D (I + small) D
    E = rng.standard_normal((m, m)) * 1e-2
    S = D @ (np.eye(m) + (E + E.T) / 2) @ D
    A = np.asfortranarray(S.copy())
    w = np.zeros(m)
    f(m, A.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    # reference for the small eigenvalues: the largest eigenvalues of S^-1 =
    # D^-1 M^-1 D^-1 (M well conditioned), which eigvalsh gets to relative accuracy
    M = np.eye(m) + (E + E.T) / 2
    Dinv = np.diag(1.0 / np.sqrt(d))
    mu = np.linalg.eigvalsh(Dinv @ np.linalg.inv(M) @ Dinv)
    wref = np.sort(1.0 / mu)
    assert np.all(np.abs(w[:3] - wref[:3]) / wref[:3] < 1e-12)
    # and a normwise solver cannot promise that: its bound is u*||S|| ~ 1e-8


def _chol(o):
    """The oracle's sparse Cholesky entry points with their argument types (each
    test binds them itself: the tests may run in any order or process)."""
    o.oracle_chol_factor.restype = ctypes.c_void_p
    o.oracle_chol_factor.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_double, ctypes.POINTER(ctypes.c_int)]
    o.oracle_chol_solve.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    o.oracle_chol_free.argtypes = [ctypes.c_void_p]
    return o


def test_sparse_cholesky_vs_dense(oracle):
    """L L^T = P(A+dI)P^T and the solve, against numpy on a small weighted grid."""
    o = _chol(oracle)
    for upper in (True, False):
        Ap, Ai, Ax = gen.weighted_lap3d(7, seed=5, upper=upper)
        n = len(Ap) - 1
        st = ctypes.c_int(0)
        F = o.oracle_chol_factor(n, Ap.ctypes.data, Ai.ctypes.data, Ax.ctypes.data, 0.25, ctypes.byref(st))
        assert st.value == 0
        D = gen.csc_to_dense(Ap, Ai, Ax, symmetrize_upper=True) + 0.25 * np.eye(n)
        rng = np.random.default_rng(0)
        b = rng.standard_normal(n)
        x = np.zeros(n)
        o.oracle_chol_solve(F, b.ctypes.data, x.ctypes.data)
        assert np.linalg.norm(D @ x - b) <= 1e-13 * np.linalg.norm(D, 1) * np.linalg.norm(x)
        assert np.allclose(x, np.linalg.solve(D, b), rtol=1e-11, atol=1e-12)
        o.oracle_chol_free(F)


def test_sparse_cholesky_known_answer(oracle):
    """(A + d I)^-1 v = v / (lambda + d) for a separable Dirichlet eigenvector."""
    o = _chol(oracle)
    N = 14
    Ap, Ai, Ax = gen.lap3d(N, True)
    n = N ** 3
    st = ctypes.c_int(0)
    F = o.oracle_chol_factor(n, Ap.ctypes.data, Ai.ctypes.data, Ax.ctypes.data, 0.0, ctypes.byref(st))
    assert st.value == 0
    v = gen.dirichlet_mode3(N, 2, 3, 1)
    lam = sum(2 - 2 * np.cos(p * np.pi / (N + 1)) for p in (2, 3, 1))
    x = np.zeros(n)
    o.oracle_chol_solve(F, v.ctypes.data, x.ctypes.data)
    assert np.abs(x - v / lam).max() <= 1e-12 / lam
    o.oracle_chol_free(F)


def test_neumann_not_pd_detected(oracle):
    o = _chol(oracle)
    Ap, Ai, Ax = gen.lap3d(8, False)
    st = ctypes.c_int(0)
    F = o.oracle_chol_factor(512, Ap.ctypes.data, Ai.ctypes.data, Ax.ctypes.data, 0.0, ctypes.byref(st))
    assert st.value == 1
    o.oracle_chol_free(F)


def test_weighted_laplacian_constant_mode(solve):
    """A = L_w + 1e-3 I: lambda_0 = 1e-3 exactly with the constant eigenvector."""
    A = gen.weighted_lap3d(12, seed=11)
    r = solve(A, 4, "sm")
    assert r.rc == 0 and r.info["flag"] == 0
    assert abs(r.lam[0] - 1e-3) <= 1e-12
    v = r.V[:, 0] * np.sign(r.V[0, 0])
    assert np.abs(v - 1 / np.sqrt(len(v))).max() <= 1e-9


MF_SO = ORACLE_SO.parent / "liboracle_mf.so"


@pytest.fixture(scope="module")
def oracle_mf(oracle):
    if not MF_SO.exists():
        pytest.skip("the multifrontal CPU baseline needs MKL's BLAS (oracle/Makefile)")
    lib = ctypes.CDLL(str(MF_SO), mode=ctypes.RTLD_GLOBAL)
    pkg.bind_api(lib, "oracle_")
    lib.oracle_mf_factor.restype = ctypes.c_void_p
    lib.oracle_mf_factor.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_double, ctypes.POINTER(ctypes.c_int)]
    lib.oracle_mf_solve.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.oracle_mf_free.argtypes = [ctypes.c_void_p]
    lib.oracle_mf_set_threads.argtypes = [ctypes.c_int]
    return lib


@pytest.mark.parametrize("case", ["weighted12", "disconnected", "neumann"])
def test_multifrontal_baseline_solve_vs_numpy(oracle_mf, case):
    """The CPU baseline's supernodal factor (oracle/o_mf.c): solves to
    rounding against numpy's dense solve, on a weighted 3-D Laplacian, on two
    disconnected blocks (several tree roots) and with the singular Neumann
    matrix reported as not positive definite (status 1)."""
    A = {"weighted12": pkg.gen.weighted_lap3d(12, seed=3, upper=True),
         "disconnected": pkg.gen.block_diag([pkg.gen.lap3d(7, True), pkg.gen.lap3d(5, True)]),
         "neumann": pkg.gen.lap3d(10, False)}[case]
    Ap, Ai, Ax = (np.ascontiguousarray(a) for a in A)
    n = len(Ap) - 1
    st = ctypes.c_int(-9)
    F = oracle_mf.oracle_mf_factor(n, Ap.ctypes.data, Ai.ctypes.data, Ax.ctypes.data, 0.0, ctypes.byref(st))
    try:
        if case == "neumann":
            assert st.value == 1
            return
        assert st.value == 0
        D = pkg.gen.csc_to_dense(Ap, Ai, Ax, symmetrize_upper=True)
        b = np.random.default_rng(3).standard_normal(n)
        x = np.zeros(n)
        oracle_mf.oracle_mf_solve(F, b.ctypes.data, x.ctypes.data)
        assert np.linalg.norm(D @ x - b) <= 1e-13 * np.linalg.norm(b) * np.linalg.cond(D)
        assert np.abs(x - np.linalg.solve(D, b)).max() <= 1e-10 * np.abs(x).max()
    finally:
        oracle_mf.oracle_mf_free(F)


@pytest.mark.parametrize("threads", [1, 4])
def test_multifrontal_baseline_speigs_equals_oracle(oracle, oracle_mf, threads):
    """The driver over the multifrontal factor takes the simplicial oracle's
    path exactly: same operation counts and shift, eigenvalues to 1e-12."""
    oracle_mf.oracle_mf_set_threads(threads)
    for A, k in ((pkg.gen.weighted_lap3d(16, seed=3), 12), (pkg.gen.lap3d(14, False), 5)):
        r = pkg.speigs(A, k, "sm", opts={"tol": 1e-10}, lib=oracle_mf, prefix="oracle_")
        o = pkg.speigs(A, k, "sm", opts={"tol": 1e-10}, lib=oracle, prefix="oracle_")
        assert r.rc == 0 and o.rc == 0
        assert r.info["nops"] == o.info["nops"] and r.info["shift"] == o.info["shift"]
        assert np.abs(r.lam - o.lam).max() <= 1e-12 * np.abs(o.lam).max()
