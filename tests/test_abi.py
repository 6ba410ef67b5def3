"""The C-ABI library loads and exports every symbol include/*.h declares, and
its host-only short-circuits behave like the reference (no GPU needed)."""
import re
from pathlib import Path

import numpy as np
import pytest

import ref_cases as rc
from conftest import ROOT, load_pkg

pkg = load_pkg()


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b([a-z_][a-z0-9_]*)\s*\(", txt):
            nm = m.group(1)
            if nm.startswith("speigs") and not nm.endswith("_t"):
                names.add(nm)
    return names


def test_exports(hiplib):
    declared = _declared()
    assert {"speigs", "speigs_z", "speigs_default_opts", "speigs_errmsg", "speigs_ex",
            "speigs_hip_op_create", "speigs_hip_op_apply"} <= declared
    for nm in sorted(declared):
        assert hasattr(hiplib, nm), nm


def test_errmsg_and_defaults(hiplib):
    assert pkg.errmsg(0) == "ok"
    assert "symmetric" in pkg.errmsg(-6)
    o = pkg.default_opts()
    assert (o.dense_max, o.detect, o.check_sym, o.tol, o.ncv, o.seed) == (-1, 1, 1, 0.0, 0, 0)


def _solve(A, k, which, opts=None, vectors=True, resid=True):
    return pkg.speigs(A, k, which, opts=opts, vectors=vectors, resid=resid)


def test_arg_errors(hiplib):
    A = pkg.gen.lap3d(4, True)
    assert _solve(A, 0, "sm").rc == -1
    assert _solve(A, 65, "sm").rc == -1


def test_host_diag_path(hiplib):
    rc.check_diagonal(_solve)


def test_host_dense_path(hiplib):
    A = pkg.gen.lap3d(8, True)                  # n = 512 <= 800: dense short-circuit
    ex = pkg.gen.exact_spectrum3(8, True)
    r = _solve(A, 6, "sm")
    assert r.info["path"] == 1
    assert np.allclose(r.lam, ex[:6], rtol=1e-12)
    D = pkg.gen.csc_to_dense(*A)
    assert np.linalg.norm(D @ r.V - r.V * r.lam, axis=0).max() < 1e-11


@pytest.mark.parametrize("case", ["random", "multiplet", "lm"])
def test_host_dense_selected_vectors(hiplib, case):
    """the dense short-circuit computes all eigenvalues but only the k returned
    vectors (inverse iteration on the tridiagonal, clusters reorthogonalised):
    eigenvalues equal numpy's to 1e-12 relative, residuals and orthogonality at
    rounding level, also inside the triple eigenvalues of a cube Laplacian."""
    rng = np.random.default_rng(11)
    which = "sm"
    if case == "multiplet":
        A = pkg.gen.lap3d(9, True)              # n = 729: triples and sextuples
        k = 12
    else:
        n = 600
        M = rng.standard_normal((n, n))
        M = (M + M.T) / 2
        M[np.abs(M) < 1.5] = 0.0                 # sparse-ish pattern, stays connected
        M += np.diag(np.arange(n) * 0.01)
        cols = [np.nonzero(M[:, j])[0] for j in range(n)]
        Ap = np.concatenate([[0], np.cumsum([len(c) for c in cols])]).astype(np.int64)
        Ai = np.concatenate(cols).astype(np.int64)
        Ax = np.concatenate([M[c, j] for j, c in enumerate(cols)])
        A = (Ap, Ai, Ax)
        k = 10
        which = "lm" if case == "lm" else "sm"
    D = pkg.gen.csc_to_dense(*A)
    r = _solve(A, k, which)
    assert r.rc == 0 and r.info["path"] == 1
    w = np.linalg.eigvalsh(D)
    ref = w[np.argsort(np.abs(w), kind="stable")]
    ref = ref[:k] if which == "sm" else ref[::-1][:k]
    assert np.abs(np.sort(r.lam) - np.sort(ref)).max() <= 1e-12 * np.abs(w).max()
    nrm = np.abs(w).max()
    assert np.linalg.norm(D @ r.V - r.V * r.lam, axis=0).max() <= 1e-11 * nrm
    assert np.abs(r.V.T @ r.V - np.eye(k)).max() <= 1e-11


def test_host_nonsymmetric_rejected(hiplib):
    n = 5000
    Ap, Ai, Ax = rc.tridiag_1d(n, 5.0, -1.0)
    col = np.repeat(np.arange(n), np.diff(Ap))
    Ax = np.where(Ai > col, -4.0, Ax)
    assert _solve((Ap, Ai, Ax), 3, "sm", vectors=False).rc == -6


def test_release_cache_without_device(hiplib):
    pkg.release_cache()        # nothing kept: no device call, no error


def test_gpu_paths_fail_loudly_without_device(hiplib):
    if hiplib.speigs_hip_device_count() > 0:
        pytest.skip("a HIP device is present")
    A = pkg.gen.lap3d(12, True)
    assert _solve(A, 4, "lm").rc == -5          # SPEIGS_ERR_INTERNAL, no CPU fallback
    assert "no HIP device visible" in pkg.last_error()
    assert _solve(pkg.gen.lap3d(3, True), 2, "sm").rc == 0     # dense short-circuit: host only
    assert pkg.last_error() == ""               # a successful call leaves no record


def test_norm_matches_serial_sum_order(hiplib):
    """info.anorm (max |A| column sum, set before the device is needed) equals
    the column sums accumulated in the order of one serial pass over the
    stored upper triangle, bitwise, with rows coupled across the whole range
    so that the threaded sum must defer entries of earlier column ranges."""
    if hiplib.speigs_hip_device_count() > 0:
        pytest.skip("a HIP device is present: the call would run the eigensolver")
    rng = np.random.default_rng(11)
    n = 300_000
    far = rng.integers(0, np.arange(1, n), size=n - 1)          # a row < j for every column j >= 1
    cols = [np.arange(n), np.arange(1, n), np.arange(1, n)]
    rows = [np.arange(n), np.arange(n - 1), far]
    col = np.concatenate(cols)
    row = np.concatenate(rows)
    o = np.lexsort((row, col))
    col, row = col[o], row[o]
    keep = np.ones(col.size, bool)
    keep[1:] = (col[1:] != col[:-1]) | (row[1:] != row[:-1])
    col, row = col[keep], row[keep]
    val = rng.uniform(-3.0, 3.0, col.size) * 10.0 ** rng.integers(-8, 8, col.size)
    val[row == col] = np.abs(val[row == col]) + 1.0
    Ap = np.zeros(n + 1, np.int64)
    np.add.at(Ap, col + 1, 1)
    Ap = np.cumsum(Ap)
    # serial order: entry q adds |a| to cs[col], then (off the diagonal) to cs[row]
    tgt = np.stack([col, row], 1).ravel()
    add = np.stack([np.abs(val), np.where(row == col, 0.0, np.abs(val))], 1).ravel()
    live = np.ones(tgt.size, bool)
    live[1::2] = row != col
    cs = np.zeros(n)
    np.add.at(cs, tgt[live], add[live])
    r = _solve((Ap, row.astype(np.int64), val), 4, "sm", vectors=False)
    assert r.rc == -5
    assert r.info["anorm"] == cs.max()


@pytest.mark.parametrize("case", ["full", "upper", "weighted_upper", "blocks", "chain"])
def test_fast_rcm_equals_reference_scan(hiplib, oracle, case):
    """The O(n+nnz) RCM of the banded gate reproduces the reference's O(n^2)
    seed-scan order exactly (oracle restatement), on full and upper-only storage."""
    import ctypes
    gen = pkg.gen
    A = {"full": lambda: gen.lap3d(9, True),
         "upper": lambda: gen.upper_only(*gen.lap3d(9, True)),
         "weighted_upper": lambda: gen.weighted_lap3d(8, seed=2, upper=True),
         "blocks": lambda: gen.block_diag([gen.lap3d(5, True), gen.lap2d(7, True), gen.lap3d(4, False)]),
         "chain": lambda: rc.tridiag_1d(300)}[case]()
    Ap, Ai, _ = (np.ascontiguousarray(a) for a in A)
    n = len(Ap) - 1
    p1 = np.zeros(n, np.int64)
    p2 = np.zeros(n, np.int64)
    hiplib.speigs_hip_rcm.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    oracle.oracle_rcm.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    assert hiplib.speigs_hip_rcm(n, Ap.ctypes.data, Ai.ctypes.data, p1.ctypes.data) == 0
    oracle.oracle_rcm(n, Ap.ctypes.data, Ai.ctypes.data, p2.ctypes.data)
    assert np.array_equal(np.sort(p1), np.arange(n))
    assert np.array_equal(p1, p2)


@pytest.mark.parametrize("case", ["grid", "components"])
def test_analyze_deterministic_across_threads(hiplib, monkeypatch, case):
    """host ordering + symbolic analysis (no device): the same permutation and
    supernode partition with 1, 3, 8 and 16 threads (regions are dissected by
    whichever thread takes them), a valid permutation, and no empty supernode
    (every supernode id owns at least one column)."""
    import ctypes
    if case == "grid":
        A = pkg.gen.weighted_lap3d(40, seed=3)
    else:
        Ap, Ai, Ax, _ = pkg.gen.batched_components(48, seed=4, nmin=300, nmax=20000)
        A = (Ap, Ai, Ax)
    Ap, Ai = (np.ascontiguousarray(a) for a in A[:2])
    n = len(Ap) - 1
    hiplib.speigs_hip_analyze.argtypes = [ctypes.c_int64] + [ctypes.c_void_p] * 4 + [ctypes.POINTER(pkg.HipInfo)]
    out = []
    for t in ("1", "3", "8", "16"):
        monkeypatch.setenv("SPEIGS_HIP_THREADS", t)
        perm = np.zeros(n, np.int64)
        snode = np.zeros(n, np.int64)
        hi = pkg.HipInfo()
        assert hiplib.speigs_hip_analyze(n, Ap.ctypes.data, Ai.ctypes.data, perm.ctypes.data, snode.ctypes.data,
                                         ctypes.byref(hi)) == 0
        out.append((perm, snode, hi.nsuper))
    (p1, s1, ns1) = out[0]
    for p8, s8, ns8 in out[1:]:
        assert np.array_equal(p1, p8) and np.array_equal(s1, s8) and ns1 == ns8
    assert np.array_equal(np.sort(p1), np.arange(n))
    assert np.array_equal(np.unique(s1), np.arange(ns1))


@pytest.mark.parametrize("N,seed,perm_sha,snode_sha,nnz_l,nsuper", [
    (30, 3, "680a8cd8cf9ea15d", "e480163c2b78dc3e", 4718718, 1247),
    (24, 7, "a8c2225333b1c43b", "c80161f20d4db7e4", 1975935, 601),
])
def test_analyze_ordering_unchanged(hiplib, N, seed, perm_sha, snode_sha, nnz_l, nsuper):
    """The nested-dissection ordering and supernode partition are integer-only
    and deterministic: pinned here so that a change meant to make the analysis
    faster (prefetching, parallel passes, reused BFS structures) cannot change
    the permutation, the fill or the factor's shape unnoticed."""
    import ctypes
    import hashlib
    A = pkg.gen.weighted_lap3d(N, seed=seed)
    Ap, Ai = (np.ascontiguousarray(a) for a in A[:2])
    n = len(Ap) - 1
    hiplib.speigs_hip_analyze.argtypes = [ctypes.c_int64] + [ctypes.c_void_p] * 4 + [ctypes.POINTER(pkg.HipInfo)]
    perm = np.zeros(n, np.int64)
    snode = np.zeros(n, np.int64)
    hi = pkg.HipInfo()
    assert hiplib.speigs_hip_analyze(n, Ap.ctypes.data, Ai.ctypes.data, perm.ctypes.data, snode.ctypes.data,
                                     ctypes.byref(hi)) == 0
    assert hashlib.sha256(perm.tobytes()).hexdigest()[:16] == perm_sha
    assert hashlib.sha256(snode.tobytes()).hexdigest()[:16] == snode_sha
    assert hi.nnz_L == nnz_l and hi.nsuper == nsuper


def _csc_from_entries(n, entries):
    """CSC arrays with the entries (row, col, value) kept in the given order
    inside each column (so unsorted columns and duplicates can be built)"""
    cols = [[] for _ in range(n)]
    for r, c, v in entries:
        cols[c].append((r, v))
    Ap = np.zeros(n + 1, np.int64)
    Ai, Ax = [], []
    for c in range(n):
        for r, v in cols[c]:
            Ai.append(r)
            Ax.append(v)
        Ap[c + 1] = len(Ai)
    return Ap, np.array(Ai, np.int64), np.array(Ax, float)


def _sym_cases():
    rng = np.random.default_rng(5)
    n = 40
    base = []
    for c in range(n):
        base.append((c, c, 4.0 + c))
        for r in range(c + 1, n):
            if rng.random() < 0.15:
                v = float(rng.standard_normal())
                base.append((r, c, v))
                base.append((c, r, v))
    base.sort(key=lambda e: (e[1], e[0]))
    cases = {"symmetric": list(base)}
    # a duplicated off-diagonal pair (both mirrors twice) and a duplicated diagonal
    r0, c0, v0 = next(e for e in base if e[0] > e[1])
    dup = list(base) + [(r0, c0, 0.5), (c0, r0, 0.5), (7, 7, 1.0)]
    cases["duplicates, mirrored"] = sorted(dup, key=lambda e: (e[1], e[0]))
    dup1 = list(base) + [(r0, c0, 0.5)]
    cases["duplicate without mirror"] = sorted(dup1, key=lambda e: (e[1], e[0]))
    # a symmetric pattern whose column 3 lists its rows out of order
    uns = list(base)
    col3 = [e for e in uns if e[1] == 3]
    rest = [e for e in uns if e[1] != 3]
    cases["unsorted column"] = rest[:0] + sorted(rest + col3[::-1], key=lambda e: e[1])
    # values: within and beyond the 1e-12 relative tolerance
    for name, eps in (("perturbed 1e-13", 1e-13), ("perturbed 1e-11", 1e-11)):
        pert = [(r, c, v * (1 + eps) if (r == r0 and c == c0) else v) for r, c, v in base]
        cases[name] = pert
    cases["missing mirror"] = [e for e in base if not (e[0] == c0 and e[1] == r0)]
    return n, cases


def test_symmetry_check_matches_transpose_compare(hiplib, oracle):
    """The full-storage symmetry test (runs matched by binary search, no
    transpose formed) rejects exactly what the oracle's restatement of the
    reference's transpose comparison (src/speigs_core.c:376-398) rejects:
    duplicates, unsorted columns, value perturbations either side of the
    1e-12 tolerance, a missing mirror entry.  n = 40 takes the host dense
    short-circuit, so accepted inputs are also solved on the host."""
    n, cases = _sym_cases()
    want = {"symmetric": True, "duplicates, mirrored": True, "duplicate without mirror": False,
            "unsorted column": False, "perturbed 1e-13": True, "perturbed 1e-11": False, "missing mirror": False}
    for name, ents in cases.items():
        A = _csc_from_entries(n, ents)
        r = _solve(A, 3, "sm")
        o = pkg.speigs(A, 3, "sm", lib=oracle, prefix="oracle_")
        print(name, r.rc, o.rc)
        assert (r.rc != -6) == want[name], name
        assert (o.rc != -6) == want[name], name
        if want[name]:
            assert r.rc == 0 and o.rc == 0 and np.allclose(r.lam, o.lam, rtol=1e-12, atol=0)


def test_diag_path_selection_order(hiplib, oracle):
    """Diagonal matrices: the k values of smallest / largest magnitude with
    equal magnitudes taken in index order, the unit eigenvectors of those
    positions, a missing diagonal entry read as 0 -- the same as the oracle."""
    d = np.array([3.0, -1.0, 1.0, 2.0, -1.0, 1.0, -3.0, 0.5])
    n = len(d)
    keep = np.arange(n) != 7
    Ap = np.concatenate([[0], np.cumsum(keep)]).astype(np.int64)
    Ai = np.arange(n, dtype=np.int64)[keep]
    A = (Ap, Ai, d[keep])
    dd = np.where(keep, d, 0.0)
    for which, k in (("sm", 5), ("lm", 4)):
        mag = np.abs(dd)
        order = np.argsort(mag if which == "sm" else -mag, kind="stable")[:k]
        r = _solve(A, k, which)
        o = pkg.speigs(A, k, which, lib=oracle, prefix="oracle_")
        assert r.rc == 0 and r.info["path"] == 0
        assert np.array_equal(r.lam, dd[order]) and np.array_equal(o.lam, r.lam)
        assert np.array_equal(r.V, np.eye(n)[:, order]) and np.array_equal(o.V, r.V)
