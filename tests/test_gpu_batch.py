"""Batched components on the GPU (the blocks path, reference
src/speigs_core.c:532-632): all components that do not take the host
short-circuits share one block-diagonal device operator and run their
thick-restart Lanczos recurrences in lockstep.  Checked against the
one-component-at-a-time device path, against the CPU oracle, against closed
forms, and for the fallback when a component is not positive definite."""
import numpy as np
import pytest

from conftest import load_pkg
import ref_cases as rc

pkg = load_pkg()
gen = pkg.gen
pytestmark = pytest.mark.gpu


def _solve(A, k, which, opts=None, env=None, monkeypatch=None, **kw):
    if env and monkeypatch:
        for key, val in env.items():
            monkeypatch.setenv(key, val)
    print(f"[case] n={len(A[0]) - 1} k={k} which={which} opts={opts} env={env} {kw}", flush=True)
    r = pkg.speigs(A, k, which, opts=opts, **kw)
    if r.rc < 0:
        print(f"[error] rc={r.rc} at {pkg.last_error()}", flush=True)
    return r


def _norm1(A):
    """||A||_1 of the symmetric matrix whose upper triangle A stores (the
    blocks path leaves info.anorm at 0, as the reference's does)"""
    Ap, Ai, Ax = A
    col = np.repeat(np.arange(len(Ap) - 1), np.diff(Ap))
    s = np.zeros(len(Ap) - 1)
    up = Ai <= col
    np.add.at(s, col[up], np.abs(Ax[up]))
    off = up & (Ai < col)
    np.add.at(s, Ai[off], np.abs(Ax[off]))
    return s.max()


def _comps(sizes, seed):
    blocks = []
    for i, s in enumerate(sizes):
        if i % 2 == 0:
            m = int(round(s ** 0.5))
            blocks.append(gen.weighted_grid((m, m), seed * 1000 + i, 0.5, 2.0))
        else:
            N = int(round(s ** (1 / 3)))
            blocks.append(gen.weighted_grid((N, N, N), seed * 1000 + i, 0.5, 2.0))
    return gen.block_diag(blocks)


@pytest.mark.parametrize("which", ["sm", "lm"])
def test_batched_matches_serial(hiplib, monkeypatch, which):
    """mixed sizes (host dense components and device components): the batched
    device solve and the one-call-per-component device solve give the same
    eigenvalues (1e-10 relative) and residuals below tol * ||A_c||."""
    A = _comps([900, 1331, 2500, 3375, 4900, 700, 1600, 1000], seed=5)
    k = 8
    rb = _solve(A, k, which, opts={"tol": 1e-11}, env={"SPEIGS_HIP_BATCH": "1"}, monkeypatch=monkeypatch)
    rs = _solve(A, k, which, opts={"tol": 1e-11}, env={"SPEIGS_HIP_BATCH": "0"}, monkeypatch=monkeypatch)
    print("batched", rb.info, "serial", rs.info)
    assert rb.rc == 0 and rs.rc == 0
    assert rb.info["path"] == 4 and rb.info["ncomp"] == 8 and rb.info["flag"] == 0
    assert np.abs(rb.lam - rs.lam).max() <= 1e-10 * np.abs(rs.lam).max()
    R = gen.spmv_upper(*A, rb.V) - rb.V * rb.lam
    res = np.linalg.norm(R, axis=0)
    print("max residual", res.max())
    assert res.max() <= 1e-9 * max(1.0, np.abs(rb.lam).max())


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_batched_vs_oracle(hiplib, oracle, precision):
    """12 components, k = 6 'sm': eigenvalues equal the CPU oracle's to 1e-10
    relative, residuals within tol * ||A||, with the shared factor in FP64
    (precision 0) and in FP32 with every apply refined to FP64 (precision 1;
    precision 2, the fast path, runs the batch on the FP64 factor: the block
    eigensolver that its FP32 factor serves is not batched)."""
    A = _comps([1000, 1200, 1500, 2000, 1728, 2197, 900, 1100, 1331, 1600, 2744, 1024], seed=7)
    r = _solve(A, 6, "sm", opts={"tol": 1e-11}, hip_opts={"precision": precision})
    ro = pkg.speigs(A, 6, "sm", opts={"tol": 1e-11}, lib=oracle, prefix="oracle_")
    print("gpu", r.lam, r.info, r.hip_info, "oracle", ro.lam, ro.info)
    assert r.rc == 0 and ro.rc == 0 and r.info["path"] == 4 and r.info["flag"] == 0
    assert r.hip_info["precision_used"] == (1 if precision == 1 else 0)
    if precision == 1:
        assert r.hip_info["ir_steps"] > 0
    assert np.abs(r.lam - ro.lam).max() <= 1e-10 * np.abs(ro.lam).max()
    R = gen.spmv_upper(*A, r.V) - r.V * r.lam
    assert np.linalg.norm(R, axis=0).max() <= 1e-11 * _norm1(A)


def test_config4_full_size(hiplib, oracle):
    """The 512-component workload at its full size (n = 5.08e6, sizes
    log-uniform in [500, 50,000], half 2-D, half 3-D), k = 6 'sm', with the
    FP32 factor refined to FP64 (precision 1): every returned vector lies in
    one component, every residual is within tol * ||A||_1, the values equal the
    FP64 batched run's to 1e-10 relative, and each value is one of the oracle's
    k smallest of the component that holds its vector.  Precision 2 (the fast
    path) runs the batch in FP64, the faster operator for lockstep Lanczos."""
    Ap, Ai, Ax, sizes = gen.batched_components(512, seed=4)
    A = (Ap, Ai, Ax)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    r32 = _solve(A, 6, "sm", opts={"tol": 1e-10}, hip_opts={"precision": 1})
    r64 = _solve(A, 6, "sm", opts={"tol": 1e-10}, hip_opts={"precision": 0})
    rf = _solve(A, 6, "sm", opts={"tol": 1e-10}, hip_opts={"precision": 2})
    print("fp32", r32.info, r32.hip_info, "fp64", r64.info)
    assert r32.hip_info["precision_used"] == 1 and r32.hip_info["ir_steps"] > 0
    assert rf.rc == 0 and rf.hip_info["precision_used"] == 0 and np.array_equal(rf.lam, r64.lam)
    for r in (r32, r64):
        assert r.rc == 0 and r.info["flag"] == 0 and r.info["ncomp"] == 512 and r.info["path"] == 4
    assert np.abs(r32.lam - r64.lam).max() <= 1e-10 * np.abs(r64.lam).max()
    R = gen.spmv_upper(*A, r32.V) - r32.V * r32.lam
    assert np.linalg.norm(R, axis=0).max() <= 1e-10 * _norm1(A)
    for t in range(6):
        mass = np.array([np.sum(r32.V[offs[c]:offs[c + 1], t] ** 2) for c in range(512)])
        c = int(np.argmax(mass))
        assert np.sort(mass)[-2] <= 1e-20
        lo, hi = offs[c], offs[c + 1]
        sub = (Ap[lo:hi + 1] - Ap[lo], Ai[Ap[lo]:Ap[hi]] - lo, Ax[Ap[lo]:Ap[hi]])
        o = pkg.speigs(sub, 6, "sm", opts={"tol": 1e-10}, lib=oracle, prefix="oracle_")
        assert o.rc == 0
        assert np.min(np.abs(o.lam - r32.lam[t])) <= 1e-10 * abs(r32.lam[t]), (t, c, r32.lam[t], o.lam)


def test_batched_closed_form(hiplib):
    """three Dirichlet Laplacians (13^3, 11^3, 12^3) on the device batch: the 8
    smallest eigenvalues are the closed-form union."""
    A = gen.block_diag([gen.lap3d(13, True), gen.lap3d(11, True), gen.lap3d(12, True)])
    allx = np.sort(np.concatenate([gen.exact_spectrum3(13), gen.exact_spectrum3(11), gen.exact_spectrum3(12)]))
    r = pkg.speigs(A, 8, "sm", opts={"tol": 1e-12})
    assert r.rc == 0 and r.info["path"] == 4 and r.info["ncomp"] == 3
    assert np.abs(r.lam - allx[:8]).max() <= 1e-10 * allx[7]


def test_batched_singular_component_falls_back(hiplib, oracle):
    """the reference's Neumann test matrix (16^3, singular; test/test_speigs.c:90-113)
    next to two Dirichlet blocks: the shared factorization is not positive
    definite, so every device component is solved by its own call (shift
    escalation per component).  Operator counts within one restart cycle of the
    oracle's (the count of the singular block depends on the rounding of its
    factor, i.e. on the fill-reducing order), the reference test's tolerances
    against the closed form."""
    A = gen.block_diag([gen.lap3d(16, False), gen.lap3d(12, True), gen.lap3d(10, True)])
    ex = gen.exact_spectrum3(16, False)
    r = pkg.speigs(A, 5, "sm")
    ro = pkg.speigs(A, 5, "sm", lib=oracle, prefix="oracle_")
    print("gpu", r.lam, r.info, "oracle", ro.lam, ro.info)
    assert r.rc == 0 and ro.rc == 0 and r.info["flag"] == 0
    assert abs(r.info["nops"] - ro.info["nops"]) <= 20     # ncv = 20 Lanczos vectors per cycle
    assert abs(r.lam[0]) <= 1e-9
    assert np.abs(r.lam[1:] - ex[1:5]).max() <= 1e-7


def test_ref8_blocks_batched(hiplib):
    """the reference's test_blocks on components large enough for the device batch."""
    rc.check_blocks(lambda A, k, which: pkg.speigs(A, k, which), N1=12, N2=10, k=6)


def test_many_components_batched_small(hiplib, monkeypatch):
    """the many-component generator at 64 components (sizes 500..5,000): batched and serial
    device solves agree; every returned vector lives in one component."""
    Ap, Ai, Ax, sizes = gen.batched_components(64, seed=4, nmin=500, nmax=5000)
    A = (Ap, Ai, Ax)
    rb = _solve(A, 6, "sm", opts={"tol": 1e-10}, env={"SPEIGS_HIP_BATCH": "1"}, monkeypatch=monkeypatch)
    rs = _solve(A, 6, "sm", opts={"tol": 1e-10}, env={"SPEIGS_HIP_BATCH": "0"}, monkeypatch=monkeypatch)
    assert rb.rc == 0 and rs.rc == 0 and rb.info["ncomp"] == 64
    assert np.abs(rb.lam - rs.lam).max() <= 1e-10 * np.abs(rs.lam).max()
    offs = np.concatenate([[0], np.cumsum(sizes)])
    for t in range(6):
        mass = np.array([np.sum(rb.V[offs[c]:offs[c + 1], t] ** 2) for c in range(64)])
        assert np.sort(mass)[-2] <= 1e-20


@pytest.mark.parametrize("which", ["sm", "lm"])
def test_dense_components_on_device(hiplib, oracle, monkeypatch, which):
    """the components the recursion sends to the dense short-circuit
    (n <= 800, src/speigs_core.c:318-353) are tridiagonalised and
    back-transformed on the device in one batch: with a 9-row (host), five
    81-729-row (device dense batch) and a 1,728-row (device sparse batch)
    component, the k = 10 values equal the oracle's to 1e-12 relative and the
    host dense path's (SPEIGS_HIP_DENSE_DEVICE=0) to 1e-13, and every returned
    vector's residual against A is within 1e-11 ||A||_1."""
    sizes = [9, 81, 125, 400, 343, 729, 1728]
    A = _comps(sizes, 11)
    k = 10
    r = _solve(A, k, which)
    o = pkg.speigs(A, k, which, lib=oracle, prefix="oracle_")
    h = _solve(A, k, which, env={"SPEIGS_HIP_DENSE_DEVICE": "0"}, monkeypatch=monkeypatch)
    assert r.rc == 0 and o.rc == 0 and h.rc == 0 and r.info["path"] == 4 and r.info["ncomp"] == len(sizes)
    scale = np.abs(o.lam).max()
    assert np.abs(r.lam - o.lam).max() <= 1e-12 * scale
    assert np.abs(r.lam - h.lam).max() <= 1e-13 * scale
    R = gen.spmv_upper(*A, r.V) - r.V * r.lam
    assert np.linalg.norm(R, axis=0).max() <= 1e-11 * _norm1(A)
    assert np.abs(np.linalg.norm(r.V, axis=0) - 1.0).max() <= 1e-12
