"""'sm' SHIFTINV path on the GPU: device multifrontal factor + triangular solves
inside the thick-restart Lanczos, against closed forms, numpy and the oracle."""
import ctypes

import numpy as np
import pytest

import ref_cases as rc
from conftest import load_pkg

pkg = load_pkg()
gen = pkg.gen
pytestmark = pytest.mark.gpu


def _solve(A, k, which, opts=None, vectors=True, resid=True, hip_opts=None):
    print(f"[case] n={len(A[0]) - 1} k={k} which={which} opts={opts} hip={hip_opts}", flush=True)
    r = pkg.speigs(A, k, which, opts=opts, vectors=vectors, resid=resid, hip_opts=hip_opts)
    if r.rc < 0:
        print(f"[error] rc={r.rc} at {pkg.last_error()}", flush=True)
    return r


def _op(hiplib, A, mode, precision=0, delta=0.0):
    hiplib.speigs_hip_op_create.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                            ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(pkg.HipInfo)]
    hiplib.speigs_hip_op_apply.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    hiplib.speigs_hip_op_destroy.argtypes = [ctypes.c_void_p]
    Ap, Ai, Ax = (np.ascontiguousarray(a) for a in A)
    op = ctypes.c_void_p()
    hi = pkg.HipInfo()
    rc_ = hiplib.speigs_hip_op_create(len(Ap) - 1, Ap.ctypes.data, Ai.ctypes.data, Ax.ctypes.data, mode, precision,
                                      delta, ctypes.byref(op), ctypes.byref(hi))
    if rc_ < 0:
        print(f"[error] op_create rc={rc_} at {pkg.last_error()}", flush=True)
    return rc_, op, hi


def _apply(hiplib, op, B, X):
    rc_ = hiplib.speigs_hip_op_apply(op, B.ctypes.data, X.ctypes.data, 1 if B.ndim == 1 else B.shape[1])
    if rc_ < 0:
        print(f"[error] op_apply rc={rc_} at {pkg.last_error()}", flush=True)
    return rc_


def test_solve_wait_failure_is_reported(hiplib, monkeypatch):
    """Host-forced failure of the blocked FP64 solve's bounded waits: with
    SPEIGS_HIP_WAIT_TEST=1 every wait looks for a flag value no work item
    publishes and gives up after 64 polls, so every kernel still exits.  The
    failed solve must be reported by the apply that ran it (not only by a later
    one) and by speigs(), with the site in speigs_hip_last_error(); a factor
    made without the knob solves normally afterwards."""
    A = gen.weighted_lap3d(24, seed=9, upper=True)
    n = len(A[0]) - 1
    B = np.asfortranarray(np.random.default_rng(5).standard_normal((n, 1)))
    X = np.zeros((n, 1), order="F")
    monkeypatch.setenv("SPEIGS_HIP_SOLVE_MODE", "1")      # the blocked solve has the waits
    monkeypatch.setenv("SPEIGS_HIP_WAIT_TEST", "1")
    rc_, op, _ = _op(hiplib, A, 1, 0, 0.0)
    assert rc_ == 0
    assert _apply(hiplib, op, B, X) == -5
    assert "bounded wait" in pkg.last_error()
    hiplib.speigs_hip_op_destroy(op)
    r = _solve(A, 4, "sm")
    assert r.rc == -5 and "bounded wait" in pkg.last_error()
    monkeypatch.delenv("SPEIGS_HIP_WAIT_TEST")
    rc_, op, _ = _op(hiplib, A, 1, 0, 0.0)
    assert rc_ == 0 and _apply(hiplib, op, B, X) == 0 and pkg.last_error() == ""
    hiplib.speigs_hip_op_destroy(op)
    assert np.linalg.norm(gen.spmv_upper(*A, X) - B) <= 1e-11 * np.linalg.norm(B)


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("precision", [0, 1])
def test_solve_vs_numpy(hiplib, upper, precision):
    A = gen.weighted_lap3d(11, seed=5, upper=upper)
    n = len(A[0]) - 1
    delta = 0.05
    rc_, op, hi = _op(hiplib, A, 1, precision, delta)
    assert rc_ == 0, rc_
    D = gen.csc_to_dense(*A, symmetrize_upper=True) + delta * np.eye(n)
    B = np.asfortranarray(np.random.default_rng(2).standard_normal((n, 3)))
    X = np.zeros((n, 3), order="F")
    assert _apply(hiplib, op, B, X) == 0
    Xref = np.linalg.solve(D, B)
    err = np.abs(X - Xref).max() / np.abs(Xref).max()
    print("solve rel err", err, "nnzL", hi.nnz_L, "nsuper", hi.nsuper)
    assert err <= 1e-11
    hiplib.speigs_hip_op_destroy(op)


def test_solve_known_answer(hiplib):
    """(A)^-1 v = v / lambda for a separable Dirichlet eigenvector (20^3)."""
    N = 20
    rc_, op, hi = _op(hiplib, gen.lap3d(N, True), 1)
    assert rc_ == 0
    v = np.ascontiguousarray(gen.dirichlet_mode3(N, 3, 1, 2))
    lam = sum(2 - 2 * np.cos(p * np.pi / (N + 1)) for p in (3, 1, 2))
    x = np.zeros_like(v)
    assert _apply(hiplib, op, v, x) == 0
    assert np.abs(x - v / lam).max() <= 1e-12 / lam
    hiplib.speigs_hip_op_destroy(op)


def test_not_pd_detected(hiplib):
    rc_, op, hi = _op(hiplib, gen.lap3d(10, False), 1)   # singular Neumann Laplacian
    assert rc_ == -3


def test_ref1_dirichlet(hiplib):
    r = rc.check_dirichlet(_solve)
    assert r.info["path"] == 2


def test_ref2_neumann_singular(hiplib):
    rc.check_neumann_singular(_solve)


def test_ref6_indefinite(hiplib):
    """reference test 6 on the product library: A - cI with c between two
    eigenvalues is indefinite, no shift makes it positive definite, and the call
    falls back to an LU at sigma = 0 (path 5) and returns the k eigenvalues
    nearest zero (test/test_speigs.c:177-205)."""
    r = rc.check_indefinite(_solve)
    assert r.rc == 0


def test_ref7_pathological(hiplib):
    """reference test 7: sigma exactly on an eigenvalue (A - cI singular): the
    call returns cleanly, finds the zero eigenvalue and reports partial
    convergence (test/test_speigs.c:211-227)."""
    rc.check_pathological(_solve)


def _between(ex, mid):
    while mid + 1 < len(ex) and ex[mid + 1] - ex[mid] < 1e-6:
        mid += 1
    return 0.5 * (ex[mid] + ex[mid + 1])


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_indefinite_beyond_host_lu(hiplib, precision):
    """n = 13,824, past the 6,000 of the dense host LU (which returned
    SPEIGS_ERR_FACTOR here before): A - cI with c between two eigenvalues in
    the middle of the spectrum is factored on the GPU as A = L S L^T and every
    apply is refined against A; path 5 (the reference's LU path), converged at
    the reference's default tol (1e-14 ||A||), the 6 eigenvalues nearest zero
    equal to the closed form, residuals checked here against A.  The FP32
    precisions (1, and 2, whose block eigensolver needs a definite operator)
    run the same fallback over an FP32 signed factor."""
    N, k = 24, 6
    ex = gen.exact_spectrum3(N, True)
    c = _between(ex, len(ex) // 2)
    A = rc._shift_diag(gen.lap3d(N, True), c)
    want = np.sort(np.abs(ex - c))[:k]
    r = _solve(A, k, "sm", hip_opts={"precision": precision})
    print(r.info, r.hip_info)
    assert r.rc == 0 and r.info["path"] == 5 and r.info["flag"] == 0
    assert np.abs(np.abs(r.lam) - want).max() <= 1e-10 * max(1.0, want.max())
    res = np.linalg.norm(gen.spmv_upper(*A, r.V) - r.V * r.lam, axis=0)
    assert res.max() <= 1e-12 * r.info["anorm"]
    assert r.hip_info["ir_steps"] > 0


def test_indefinite_weighted_vs_dense(hiplib):
    """the variable-coefficient family shifted into its spectrum (n = 1,728):
    the 8 eigenvalues nearest zero equal numpy's dense ones to 1e-10, and
    their eigenvectors span the same spaces."""
    A0 = gen.weighted_lap3d(12, seed=5)
    D = gen.csc_to_dense(*A0, symmetrize_upper=True)
    ev, U = np.linalg.eigh(D)
    c = _between(ev, len(ev) // 3)
    A = rc._shift_diag(A0, c)
    order = np.argsort(np.abs(ev - c), kind="stable")[:8]
    want = ev[order] - c
    r = _solve(A, 8, "sm")
    assert r.rc == 0 and r.info["path"] == 5 and r.info["flag"] == 0
    got = r.lam[np.argsort(np.abs(r.lam), kind="stable")]
    assert np.abs(np.sort(got) - np.sort(want)).max() <= 1e-10 * np.abs(ev).max()
    P = U[:, order]
    assert np.linalg.norm(r.V - P @ (P.T @ r.V)) <= 1e-8


def test_ref3_multiplicity(hiplib):
    r = rc.check_multiplicity(_solve)
    ex = gen.exact_spectrum3(16, True)
    assert rc.subspace_angle_max(r.V[:, 1:4], rc.dirichlet_eigenspace(16, ex[1])) <= 1e-8


def test_ref8_blocks(hiplib):
    rc.check_blocks(_solve)


def test_ref13_upper_only(hiplib):
    rc.check_upper_only(_solve)


def test_config1_closed_form(hiplib, oracle):
    A = gen.upper_only(*gen.lap3d(20, True))
    want = np.array([0.067015042649, 0.133531083527, 0.133531083527, 0.133531083527,
                     0.200047124405, 0.200047124405])
    r = _solve(A, 6, "sm")
    assert r.rc == 0 and r.info["flag"] == 0 and r.info["path"] == 2
    assert np.abs(r.lam - want).max() <= 1e-11
    o = pkg.speigs(A, 6, "sm", lib=oracle, prefix="oracle_")
    assert np.abs(r.lam - o.lam).max() <= 1e-10 * np.abs(o.lam).max()
    ex = gen.exact_spectrum3(20, True)
    assert rc.subspace_angle_max(r.V[:, 1:4], rc.dirichlet_eigenspace(20, ex[1])) <= 1e-8
    U3 = rc.dirichlet_eigenspace(20, ex[4])
    assert np.linalg.norm(r.V[:, 4:6] - U3 @ (U3.T @ r.V[:, 4:6])) <= 1e-8


def test_config2_sm_closed_form(hiplib):
    A = gen.lap3d(46, True)
    want = np.array([0.013398728261, 0.026777509198, 0.026777509198, 0.026777509198,
                     0.040156290135, 0.040156290135])
    r = _solve(A, 6, "sm", hip_opts={})
    assert r.rc == 0 and r.info["flag"] == 0
    assert np.abs(r.lam - want).max() <= 1e-10 * want.max() + 1e-12
    ex = gen.exact_spectrum3(46, True)
    assert np.abs(r.lam - ex[:6]).max() <= 1e-10 * ex[5]
    R = gen.spmv_upper(*A, r.V) - r.V * r.lam
    assert np.linalg.norm(R, axis=0).max() <= 1e-10 * 12
    print("info", r.info, "hip", r.hip_info)


@pytest.mark.parametrize("precision", [0, 1])
def test_weighted_vs_oracle(hiplib, oracle, precision):
    """A = L_w + 1e-3 I at 16^3: lambda_0 = 1e-3 exactly, the rest equal the oracle."""
    A = gen.weighted_lap3d(16, seed=21)
    r = _solve(A, 10, "sm", opts={"tol": 1e-12}, hip_opts={"precision": precision})
    o = pkg.speigs(A, 10, "sm", opts={"tol": 1e-12}, lib=oracle, prefix="oracle_")
    assert r.rc == 0 and r.info["flag"] == 0
    assert abs(r.lam[0] - 1e-3) <= 1e-13
    assert np.abs(r.lam - o.lam).max() <= 1e-10 * np.abs(o.lam).max()
    assert r.resid.max() <= 1e-12 * r.info["anorm"]
    print("precision", precision, "nops", r.info["nops"], o.info["nops"], "hip", r.hip_info)


@pytest.mark.parametrize("prec", [0, 2])
def test_potrf_waves_bitwise(hiplib, monkeypatch, prec):
    """The panel factorization on one wave and with its columns split over four
    waves applies the same operations to every element in the same order: the
    factors, and so the solves, agree to the bit (FP64 factor, and FP32 with
    block solves).  The blocked panel kernel (16-column blocks, products on the
    matrix cores) sums in another order: its solve agrees to rounding and
    meets the same residual bound."""
    A = gen.weighted_lap3d(28, seed=12, upper=True)
    n = len(A[0]) - 1
    B = np.asfortranarray(np.random.default_rng(8).standard_normal((n, 3)))
    out = []
    for waves in ("1", "4", "16"):
        monkeypatch.setenv("SPEIGS_HIP_POTRF_WAVES", waves)
        rc_, op, hi = _op(hiplib, A, 1, prec, 0.0)
        assert rc_ == 0
        X = np.zeros((n, 3), order="F")
        assert _apply(hiplib, op, B, X) == 0
        hiplib.speigs_hip_op_destroy(op)
        out.append(X)
    for X in out[1:]:
        R = gen.spmv_upper(*A, X) - B
        assert np.linalg.norm(R) / np.linalg.norm(B) <= (1e-12 if prec == 0 else 1e-3)
    assert np.array_equal(out[0], out[1])
    diff = np.linalg.norm(out[2] - out[1]) / np.linalg.norm(out[1])
    print("blocked vs unblocked panel: relative solution difference", diff)
    assert diff <= (1e-11 if prec == 0 else 1e-3)


@pytest.mark.parametrize("prec,solve_mode", [(2, None), (0, "3")])
def test_solve_halves_two_streams_bitwise(hiplib, monkeypatch, prec, solve_mode):
    """Below the root the elimination tree splits into two halves that share no
    front, and the solves run the halves' level chains on two streams (joined
    before the root's forward step and after the backward sweep).  Each front
    does the same work either way: the FP32 multi-right-hand-side solve and the
    FP64 single-vector solve (inverted diagonal blocks) give the same bits as
    with the split planned away (SPEIGS_HIP_RS_SPLIT=0) and as with both halves
    on one stream (=1), over two applies of one operator."""
    A = gen.weighted_lap3d(28, seed=12, upper=True)
    n = len(A[0]) - 1
    if solve_mode:
        monkeypatch.setenv("SPEIGS_HIP_SOLVE_MODE", solve_mode)
    B = np.asfortranarray(np.random.default_rng(8).standard_normal((n, 3 if prec == 2 else 1)))
    Bv = B if prec == 2 else np.ascontiguousarray(B[:, 0])
    out = {}
    for split in ("2", "1", "0"):
        monkeypatch.setenv("SPEIGS_HIP_RS_SPLIT", split)
        rc_, op, hi = _op(hiplib, A, 1, prec, 0.0)
        assert rc_ == 0
        xs = []
        for _ in range(2):
            X = np.zeros_like(Bv, order="F")
            assert _apply(hiplib, op, Bv, X) == 0
            xs.append(X)
        hiplib.speigs_hip_op_destroy(op)
        assert np.array_equal(xs[0], xs[1])
        out[split] = xs[0]
    R = gen.spmv_upper(*A, out["2"]) - Bv
    assert np.linalg.norm(R) / np.linalg.norm(Bv) <= (1e-12 if prec == 0 else 1e-3)
    assert np.array_equal(out["2"], out["0"]) and np.array_equal(out["1"], out["0"])


@pytest.mark.parametrize("mode,mfma,prec", [("0", "0", 0), ("1", "0", 0), ("1", "1", 0), ("1", "0", 1), ("1", "1", 1),
                                            ("2", "1", 0), ("2", "1", 1), ("3", "1", 0), ("3", "1", 1)])
def test_kernel_variants_agree(hiplib, monkeypatch, mode, mfma, prec):
    """one-workgroup-per-front vs blocked multi-workgroup solves, FMA vs MFMA
    Schur updates: the same solution to rounding (46^3 has fronts of ~2000)."""
    monkeypatch.setenv("SPEIGS_HIP_SOLVE_MODE", mode)
    monkeypatch.setenv("SPEIGS_HIP_MFMA", mfma)
    A = gen.weighted_lap3d(24, seed=9, upper=True)
    n = len(A[0]) - 1
    rc_, op, hi = _op(hiplib, A, 1, prec, 0.0)
    assert rc_ == 0
    B = np.asfortranarray(np.random.default_rng(4).standard_normal((n, 2)))
    X = np.zeros((n, 2), order="F")
    assert _apply(hiplib, op, B, X) == 0
    R = gen.spmv_upper(*A, X) - B
    rel = np.linalg.norm(R) / np.linalg.norm(B)
    print("mode", mode, "mfma", mfma, "residual", rel)
    assert rel <= 1e-12
    hiplib.speigs_hip_op_destroy(op)


def test_vector_solve_lanczos_matches_chained(hiplib, monkeypatch):
    """the single-vector solve by the inverted diagonal blocks over the block
    solve's plan (mode 3: chosen by default at tolerances >= 1e-12, where one
    such solve per apply suffices) and the blocked chained solve (mode 1, the
    default at tighter tolerances) give the same eigenpairs: at the reference's
    default tolerance (1e-14 ||A||, mode 3 forced, its applies refined against
    A) the values agree to 1e-12 relative, at 1e-10 to 1e-9."""
    A = gen.weighted_lap3d(30, seed=12, upper=True)
    h = {"precision": 0}
    r3l = _solve(A, 8, "sm", opts={"tol": 1e-10}, hip_opts=h)
    monkeypatch.setenv("SPEIGS_HIP_SOLVE_MODE", "3")
    r3 = _solve(A, 8, "sm", hip_opts=h)
    monkeypatch.setenv("SPEIGS_HIP_SOLVE_MODE", "1")
    r1 = _solve(A, 8, "sm", hip_opts=h)
    r1l = _solve(A, 8, "sm", opts={"tol": 1e-10}, hip_opts=h)
    for r in (r3, r3l, r1, r1l):
        assert r.rc == 0 and r.info["flag"] == 0
    assert np.abs(r3.lam - r1.lam).max() <= 1e-12 * np.abs(r1.lam).max()
    assert np.abs(r3l.lam - r1l.lam).max() <= 1e-9 * np.abs(r1.lam).max()
    print("ir steps", r3.hip_info["ir_steps"], r3l.hip_info["ir_steps"], r1.hip_info["ir_steps"])
    assert r3.hip_info["ir_steps"] > 0 and r1.hip_info["ir_steps"] == 0


def test_inverse_solve_weighted_40(hiplib, monkeypatch):
    """inverse-based (chain-free) FP64 solve on a 40^3 weighted 3-D Laplacian: residual at rounding level."""
    monkeypatch.setenv("SPEIGS_HIP_SOLVE_MODE", "2")
    A = gen.weighted_lap3d(40, seed=3, upper=True)
    n = len(A[0]) - 1
    rc_, op, hi = _op(hiplib, A, 1, 0, 0.0)
    assert rc_ == 0
    B = np.asfortranarray(np.random.default_rng(8).standard_normal((n, 1)))
    X = np.zeros((n, 1), order="F")
    assert _apply(hiplib, op, B, X) == 0
    rel = np.linalg.norm(gen.spmv_upper(*A, X) - B) / np.linalg.norm(B)
    print("inverse-based FP64 solve residual", rel)
    assert rel <= 1e-11
    hiplib.speigs_hip_op_destroy(op)


def test_block_fp32_path_matches_fp64(hiplib):
    """FP32-factor block eigensolver: same eigenvalues as the FP64 Lanczos path (1e-10 relative)."""
    A = gen.weighted_lap3d(20, seed=13)
    r64 = _solve(A, 12, "sm", opts={"tol": 1e-11})
    rb = _solve(A, 12, "sm", opts={"tol": 1e-11}, hip_opts={"precision": 2})
    print("fp64 nops", r64.info["nops"], "block", rb.info, rb.hip_info)
    assert r64.rc == 0 and rb.rc == 0 and rb.info["flag"] == 0
    assert np.abs(rb.lam - r64.lam).max() <= 1e-10 * np.abs(r64.lam).max()
    assert rb.resid.max() <= 1e-11 * rb.info["anorm"]
    R = gen.spmv_upper(*A, rb.V) - rb.V * rb.lam
    assert np.linalg.norm(R, axis=0).max() <= 1e-11 * rb.info["anorm"] * 1.01


@pytest.mark.parametrize("N,nrhs", [(12, 7), (24, 20), (40, 32), (40, 45)])
def test_block_solve_fp32(hiplib, monkeypatch, N, nrhs):
    """unrefined FP32 multi-right-hand-side solve (precision 2): the MFMA split-K
    kernels with FP32 vectors and the generic FMA tiles with FP64 vectors agree
    to FP32 rounding, and both solve A X = B to ~kappa*u32."""
    A = gen.weighted_lap3d(N, seed=3, upper=True)
    n = len(A[0]) - 1
    B = np.asfortranarray(np.random.default_rng(N).standard_normal((n, nrhs)))
    rc_, op64, _ = _op(hiplib, A, 1, 0, 0.0)
    assert rc_ == 0
    X64 = np.zeros((n, nrhs), order="F")
    assert _apply(hiplib, op64, B, X64) == 0
    hiplib.speigs_hip_op_destroy(op64)
    rc_, op, _ = _op(hiplib, A, 1, 2, 0.0)
    assert rc_ == 0
    Xf = np.zeros((n, nrhs), order="F")
    Xg = np.zeros((n, nrhs), order="F")
    monkeypatch.setenv("SPEIGS_HIP_BLOCK_GENERIC", "0")
    assert _apply(hiplib, op, B, Xf) == 0
    monkeypatch.setenv("SPEIGS_HIP_BLOCK_GENERIC", "1")
    assert _apply(hiplib, op, B, Xg) == 0
    hiplib.speigs_hip_op_destroy(op)
    col = lambda D, R: (np.linalg.norm(D, axis=0) / np.linalg.norm(R, axis=0)).max()
    e_fg, e_g64, e_f64 = col(Xf - Xg, Xg), col(Xg - X64, X64), col(Xf - X64, X64)
    print(f"[block] N={N} nrhs={nrhs} fast-vs-generic {e_fg:.2e} generic-vs-fp64 {e_g64:.2e} "
          f"fast-vs-fp64 {e_f64:.2e}")
    assert e_g64 <= 1e-3
    assert e_f64 <= 1e-3
    assert e_fg <= 1e-4


def test_block_solve_tile_heights_bitwise(hiplib, monkeypatch):
    """the L21 products with 64- and 128-row output tiles (every level fixed by
    SPEIGS_HIP_RS_WIDE = 0 / 1, and the autotuner's per-level choice) give the
    same solution to the bit, on four solves in a row per operator (the
    autotuner times the first two and picks per level from the third on)."""
    A = gen.weighted_lap3d(40, seed=9, upper=True)
    n = len(A[0]) - 1
    B = np.asfortranarray(np.random.default_rng(7).standard_normal((n, 28)))
    out = {}
    for mode in ("0", "1", "auto", None):
        if mode is None:
            monkeypatch.delenv("SPEIGS_HIP_RS_WIDE", raising=False)
        else:
            monkeypatch.setenv("SPEIGS_HIP_RS_WIDE", mode)
        rc_, op, _ = _op(hiplib, A, 1, 2, 0.0)
        assert rc_ == 0
        xs = []
        for rep in range(4):
            X = np.zeros((n, 28), order="F")
            assert _apply(hiplib, op, B, X) == 0
            xs.append(X)
        hiplib.speigs_hip_op_destroy(op)
        for X in xs[1:]:
            assert np.array_equal(X, xs[0])
        out[mode] = xs[0]
    for m in ("1", "auto", None):
        assert np.array_equal(out["0"], out[m])


def test_device_memory_reuse(hiplib, monkeypatch):
    """Device blocks kept between calls (and reused by calls of other sizes)
    leave every result bitwise unchanged: a solve repeated after a smaller one,
    after the kept blocks were released, and with keeping turned off."""
    A = gen.weighted_lap3d(24, seed=5)
    B = gen.weighted_lap3d(20, seed=6)
    ho = {"precision": 2}

    def solve(M):
        r = _solve(M, 10, "sm", opts={"tol": 1e-11}, hip_opts=ho)
        assert r.rc == 0 and r.info["flag"] == 0
        return r

    def same(r, s):
        assert np.array_equal(r.lam, s.lam) and np.array_equal(r.V, s.V) and np.array_equal(r.resid, s.resid)

    ra, rb = solve(A), solve(B)
    same(solve(A), ra)
    same(solve(B), rb)
    pkg.release_cache()
    same(solve(A), ra)
    monkeypatch.setenv("SPEIGS_HIP_CACHE_MB", "1")      # small blocks kept, the large ones returned
    same(solve(A), ra)
    same(solve(B), rb)
    monkeypatch.setenv("SPEIGS_HIP_CACHE_MB", "0")
    same(solve(B), rb)
    same(solve(A), ra)


def _solve_z(Ap, Ai, Axr, Axi, k, which, opts=None):
    print(f"[case] speigs_z n={len(Ap) - 1} k={k} which={which}", flush=True)
    r = pkg.speigs_z(Ap, Ai, Axr, Axi, k, which, opts=opts)
    if r.rc < 0:
        print(f"[error] rc={r.rc} at {pkg.last_error()}", flush=True)
    return r


def test_ref11_hermitian(hiplib, oracle):
    """Reference test 11 (test/test_speigs.c:331-389) on the GPU path: A = U M U*
    with U unitary diagonal has the closed-form Laplacian spectrum; complex
    residuals <= 1e-8 and no two returned vectors within an i-multiple.  The
    values equal the oracle's, the returned vectors are orthonormal in C^n."""
    r = rc.check_hermitian(_solve_z)
    o = rc.check_hermitian(lambda *a: pkg.speigs_z(*a, lib=oracle, prefix="oracle_"))
    assert np.abs(r.lam - o.lam).max() <= 1e-10 * np.abs(o.lam).max()
    G = r.V.conj().T @ r.V
    assert np.abs(G - np.eye(G.shape[0])).max() <= 1e-12
    assert r.info["nconv"] == 5 and r.info["flag"] == 0


@pytest.mark.parametrize("N,k,which", [(10, 8, "sm"), (9, 6, "lm")])
def test_hermitian_multiplets(hiplib, oracle, N, k, which):
    """k reaching into the triple eigenvalues: each complex multiplet appears
    with its full multiplicity (the selection keeps C-linearly independent
    directions only), values equal the closed form and the oracle, the
    vectors of a multiplet span the closed-form eigenspace rotated by U."""
    Ap, Ai, Ax = gen.lap3d(N, True)
    n = N ** 3
    i = np.arange(n, dtype=np.int64)
    th = 0.7 * np.sin(0.37 * i) + 0.001 * i
    col = np.repeat(np.arange(n), np.diff(Ap))
    d = th[Ai] - th[col]
    Ar, Aim = Ax * np.cos(d), Ax * np.sin(d)
    r = _solve_z(Ap, Ai, Ar, Aim, k, which)
    o = pkg.speigs_z(Ap, Ai, Ar, Aim, k, which, lib=oracle, prefix="oracle_")
    assert r.rc == 0 and o.rc == 0 and r.info["flag"] == 0 and r.info["nconv"] == k
    ex = gen.exact_spectrum3(N, True)
    want = ex[:k] if which == "sm" else ex[::-1][:k]
    assert np.abs(r.lam - want).max() <= 1e-9 * np.abs(want).max()
    assert np.abs(r.lam - o.lam).max() <= 1e-10 * np.abs(o.lam).max()
    Ac = np.zeros((n, n), complex)
    Ac[Ai, col] = Ar + 1j * Aim
    res = max(np.linalg.norm(Ac @ r.V[:, t] - r.lam[t] * r.V[:, t]) for t in range(k))
    assert res <= 1e-8
    G = r.V.conj().T @ r.V
    assert np.abs(G - np.eye(k)).max() <= 1e-12
    # the eigenspace of A for value lam is U times M's: compare projectors
    U = np.exp(1j * th)
    for lam in np.unique(np.round(r.lam, 9)):
        sel = np.abs(r.lam - lam) <= 1e-8
        Mb = rc.dirichlet_eigenspace(N, lam)
        if Mb.shape[1] != sel.sum():
            continue                         # a multiplet cut at k
        Eb = U[:, None] * Mb
        Q, _ = np.linalg.qr(Eb)
        P = r.V[:, sel]
        assert np.linalg.norm(P - Q @ (Q.conj().T @ P)) <= 1e-8


def test_hermitian_real_input(hiplib):
    """speigs_z without an imaginary part is speigs."""
    A = gen.weighted_lap3d(12, seed=2, upper=True)
    r = _solve_z(*A, None, 4, "sm")
    s = _solve(A, 4, "sm")
    assert r.rc == 0 and np.array_equal(r.lam, s.lam)
