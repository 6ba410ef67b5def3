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
    print(f"[case] n={len(A[0]) - 1} k={k} which={which} opts={opts} env={env}", flush=True)
    return pkg.speigs(A, k, which, opts=opts, **kw)


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


def test_batched_vs_oracle(hiplib, oracle):
    """12 components, k = 6 'sm': eigenvalues equal the CPU oracle's to 1e-10 relative."""
    A = _comps([1000, 1200, 1500, 2000, 1728, 2197, 900, 1100, 1331, 1600, 2744, 1024], seed=7)
    r = pkg.speigs(A, 6, "sm", opts={"tol": 1e-11})
    ro = pkg.speigs(A, 6, "sm", opts={"tol": 1e-11}, lib=oracle, prefix="oracle_")
    print("gpu", r.lam, r.info, "oracle", ro.lam, ro.info)
    assert r.rc == 0 and ro.rc == 0
    assert np.abs(r.lam - ro.lam).max() <= 1e-10 * np.abs(ro.lam).max()


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
