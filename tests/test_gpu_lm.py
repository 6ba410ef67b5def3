"""'lm' DIRECT path on the GPU (SELL SpMV + CGS2 over a device-resident basis)
against closed forms and the CPU oracle."""
import ctypes
import sys

import numpy as np
import pytest

import ref_cases as rc
from conftest import load_pkg

pkg = load_pkg()
gen = pkg.gen
pytestmark = pytest.mark.gpu


def _solve(A, k, which, opts=None, vectors=True, resid=True):
    print(f"[case] n={len(A[0]) - 1} k={k} which={which} opts={opts}", flush=True)
    return pkg.speigs(A, k, which, opts=opts, vectors=vectors, resid=resid)


def test_ref4_lm(hiplib):
    r = rc.check_lm(_solve)
    assert r.info["flag"] == 0


def test_spmv_vs_numpy(hiplib):
    for upper in (False, True):
        A = gen.weighted_lap3d(9, seed=3, upper=upper)
        n = len(A[0]) - 1
        op = ctypes.c_void_p()
        Ap, Ai, Ax = (np.ascontiguousarray(a) for a in A)
        hiplib.speigs_hip_op_create.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p]
        hiplib.speigs_hip_op_multA.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        hiplib.speigs_hip_op_destroy.argtypes = [ctypes.c_void_p]
        assert hiplib.speigs_hip_op_create(n, Ap.ctypes.data, Ai.ctypes.data, Ax.ctypes.data, 0, 0, 0.0,
                                           ctypes.byref(op), None) == 0
        X = np.asfortranarray(np.random.default_rng(1).standard_normal((n, 5)))
        Y = np.zeros((n, 5), order="F")
        assert hiplib.speigs_hip_op_multA(op, X.ctypes.data, Y.ctypes.data, 5) == 0
        ref = gen.spmv_upper(Ap, Ai, Ax, X)
        assert np.abs(Y - ref).max() <= 1e-13 * np.abs(ref).max()
        hiplib.speigs_hip_op_destroy(op)


def test_lm_config2_closed_form_and_oracle(hiplib, oracle):
    """46^3 Dirichlet, k=6 'lm': closed-form top values and the oracle, 1e-10 relative."""
    A = gen.lap3d(46, True)
    want = np.array([11.986601271739, 11.973222490802, 11.973222490802, 11.973222490802,
                     11.959843709865, 11.959843709865])
    r = _solve(A, 6, "lm")
    assert r.rc == 0 and r.info["flag"] == 0 and r.info["path"] == 3
    assert np.abs(r.lam - want).max() <= 1e-10 * 12
    ex = gen.exact_spectrum3(46, True)[::-1][:6]
    assert np.abs(r.lam - ex).max() <= 1e-10 * ex.max()
    R = gen.spmv_upper(*A, r.V) - r.V * r.lam
    assert np.linalg.norm(R, axis=0).max() <= 1e-14 * 12 * 10
    o = pkg.speigs(A, 6, "lm", lib=oracle, prefix="oracle_", vectors=False)
    assert np.abs(r.lam - o.lam).max() <= 1e-10 * 12
    print("nops gpu", r.info["nops"], "oracle", o.info["nops"], "t_iter", r.info["t_iter"], o.info["t_iter"])


def test_lm_weighted_vs_oracle(hiplib, oracle):
    A = gen.weighted_lap3d(20, seed=7, upper=False)
    r = _solve(A, 8, "lm")
    o = pkg.speigs(A, 8, "lm", lib=oracle, prefix="oracle_")
    assert r.rc == 0 and o.rc == 0
    assert np.abs(r.lam - o.lam).max() <= 1e-10 * np.abs(o.lam).max()
    assert r.resid.max() <= 1e-13 * r.info["anorm"] * 10
