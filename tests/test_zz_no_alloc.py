"""speigs_hip_op_apply_noalloc: the allocation-free variant of
speigs_hip_op_apply (its staging lives in a fixed device array instead of
buffers grown on demand).  Same kernels, so the same bits: checked against
speigs_hip_op_apply on two calls in a row, the second with a smaller input,
for the FP64 and the refined FP32 solve and for the SpMV operator; inputs
beyond the array's capacity and FP32_BLOCK operators are refused with
SPEIGS_ERR_ARG before any device work."""
import ctypes

import numpy as np
import pytest

from conftest import load_pkg

pkg = load_pkg()
gen = pkg.gen
pytestmark = pytest.mark.gpu

CAP = 4 << 20          # values of one half of the scratch array


def _bind(lib):
    lib.speigs_hip_op_create.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p]
    for f in (lib.speigs_hip_op_apply, lib.speigs_hip_op_apply_noalloc):
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        f.restype = ctypes.c_int
    lib.speigs_hip_op_destroy.argtypes = [ctypes.c_void_p]


def _create(lib, A, mode, precision):
    Ap, Ai, Ax = (np.ascontiguousarray(a) for a in A)
    op = ctypes.c_void_p()
    rc = lib.speigs_hip_op_create(len(Ap) - 1, Ap.ctypes.data, Ai.ctypes.data, Ax.ctypes.data, mode, precision,
                                  0.0, ctypes.byref(op), None)
    assert rc == 0, pkg.last_error()
    return op


def _apply(fn, op, B):
    X = np.zeros_like(B)
    rc = fn(op, B.ctypes.data, X.ctypes.data, B.shape[1])
    assert rc == 0, pkg.last_error()
    return X


@pytest.mark.parametrize("mode,precision", [(1, 0), (1, 1), (0, 0)])
def test_noalloc_matches_apply(hiplib, mode, precision):
    _bind(hiplib)
    rng = np.random.default_rng(17)
    for N, nvec in ((24, 3), (12, 1)):            # the second call: a smaller input
        A = gen.weighted_lap3d(N, seed=N, upper=True)
        n = len(A[0]) - 1
        print(f"[case] noalloc mode={mode} precision={precision} n={n} nvec={nvec}", flush=True)
        op = _create(hiplib, A, mode, precision)
        B = np.asfortranarray(rng.standard_normal((n, nvec)))
        X1 = _apply(hiplib.speigs_hip_op_apply_noalloc, op, B)
        X0 = _apply(hiplib.speigs_hip_op_apply, op, B)
        X2 = _apply(hiplib.speigs_hip_op_apply_noalloc, op, B)
        hiplib.speigs_hip_op_destroy(op)
        assert np.array_equal(X1, X0) and np.array_equal(X2, X0)
        AX = gen.spmv_upper(*A, X1)
        if mode == 1:
            assert np.linalg.norm(AX - B) <= 1e-11 * np.linalg.norm(B)
        else:
            assert np.abs(X1 - gen.spmv_upper(*A, B)).max() <= 1e-12 * np.abs(X1).max()


def test_noalloc_refusals(hiplib):
    _bind(hiplib)
    A = gen.weighted_lap3d(12, seed=3, upper=True)
    n = len(A[0]) - 1
    op = _create(hiplib, A, 0, 0)
    nvec = CAP // n + 1                           # n * nvec just beyond one half
    B = np.zeros((n, 1))
    # the size test comes first: no vector of that size is ever touched
    assert hiplib.speigs_hip_op_apply_noalloc(op, B.ctypes.data, B.ctypes.data, nvec) == -1
    assert "exceeds" in pkg.last_error()
    hiplib.speigs_hip_op_destroy(op)
    op = _create(hiplib, A, 1, 2)
    assert hiplib.speigs_hip_op_apply_noalloc(op, B.ctypes.data, B.ctypes.data, 1) == -1
    assert "FP32_BLOCK" in pkg.last_error()
    hiplib.speigs_hip_op_destroy(op)
