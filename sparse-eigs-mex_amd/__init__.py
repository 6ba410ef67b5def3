"""speigs_amd -- Python mirror of the speigs C API over the MI355X library.

`libspeigs_hip.so` (built from csrc/ by build.py) exports the reference's C
API unchanged (speigs, speigs_z, speigs_default_opts, speigs_errmsg;
reference include/speigs.h:70-99) plus the speigs_hip_* extension declared in
include/speigs_hip.h.  This module binds both with ctypes and mirrors the
reference's calling convention: CSC input (int64 Ap/Ai, float64 Ax), results
lambda / V / resid / info, negative return codes for errors.

The product path has no CPU fallback: if the shared library is missing,
`load()` raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int64, c_uint64, c_void_p
from dataclasses import dataclass
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_DIR = PKG_DIR / "lib"
LIB_NAME = "libspeigs_hip.so"

SM, LM = 0, 1
PATHS = {0: "DIAG", 1: "DENSE", 2: "SHIFTINV", 3: "DIRECT", 4: "BLOCKS", 5: "LU", 6: "BANDED"}
OK, ERR_ARG, ERR_MEM, ERR_FACTOR, ERR_LAPACK, ERR_INTERNAL, ERR_NOTSYM = 0, -1, -2, -3, -4, -5, -6


class Opts(ctypes.Structure):
    """speigs_opts (reference include/speigs.h:40-55)."""
    _fields_ = [("tol", c_double), ("maxit", c_int64), ("ncv", c_int64), ("dense_max", c_int),
                ("detect", c_int), ("seed", c_uint64), ("check_sym", c_int), ("band_max", c_int),
                ("ordering", c_int), ("shift0", c_double)]


class Info(ctypes.Structure):
    """speigs_info (reference include/speigs.h:57-68)."""
    _fields_ = [("flag", c_int), ("nconv", c_int64), ("restarts", c_int64), ("nops", c_int64),
                ("shift", c_double), ("anorm", c_double), ("path", c_int), ("ncomp", c_int64),
                ("t_analyze", c_double), ("t_factor", c_double), ("t_iter", c_double),
                ("t_op", c_double), ("t_ortho", c_double), ("t_resid", c_double), ("t_ritz", c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class HipOpts(ctypes.Structure):
    """speigs_hip_opts (include/speigs_hip.h)."""
    _fields_ = [("precision", c_int), ("ir_tol", c_double), ("ir_maxit", c_int),
                ("start", c_int), ("start_block", c_int), ("use_graph", c_int),
                ("device", c_int), ("verbose", c_int), ("start_vec", POINTER(c_double)),
                ("model", ctypes.c_void_p), ("reserved", c_int64 * 7)]


class HipInfo(ctypes.Structure):
    """speigs_hip_info (include/speigs_hip.h)."""
    _fields_ = [("nnz_L", c_int64), ("factor_flops", c_double), ("nsuper", c_int64),
                ("nlevels", c_int64), ("ir_steps", c_int64), ("solves", c_int64),
                ("precision_used", c_int), ("fallback", c_int),
                ("t_order", c_double), ("t_symbolic", c_double), ("t_upload", c_double),
                ("t_numeric", c_double), ("t_solve", c_double), ("t_spmv", c_double),
                ("bytes_solve", c_double), ("bytes_spmv", c_double),
                ("t_invert", c_double), ("perturbed_pivots", c_double), ("reserved", c_double * 6)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


_I64P = POINTER(c_int64)
_F64P = POINTER(c_double)

_lib = None


def lib_path() -> Path:
    return LIB_DIR / LIB_NAME


def load(path: str | os.PathLike | None = None) -> ctypes.CDLL:
    """Load the MI355X library; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else lib_path()
    if not p.exists():
        raise RuntimeError(f"speigs HIP library not built: {p} is missing (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(p), mode=ctypes.RTLD_GLOBAL)
    bind_api(lib, "")
    _bind_ext(lib)
    if path is None:
        _lib = lib
    return lib


def bind_api(lib: ctypes.CDLL, prefix: str) -> None:
    """Declare the four reference entry points (with a symbol prefix, so the
    same binding also serves the oracle's oracle_* copies in tests)."""
    f = getattr(lib, prefix + "speigs")
    f.restype = c_int
    f.argtypes = [c_int64, _I64P, _I64P, _F64P, c_int64, c_int, POINTER(Opts), _F64P, _F64P, _F64P,
                  POINTER(Info)]
    f = getattr(lib, prefix + "speigs_z")
    f.restype = c_int
    f.argtypes = [c_int64, _I64P, _I64P, _F64P, _F64P, c_int64, c_int, POINTER(Opts), _F64P, _F64P,
                  _F64P, _F64P, POINTER(Info)]
    f = getattr(lib, prefix + "default_opts" if prefix else "speigs_default_opts")
    f.restype = None
    f.argtypes = [POINTER(Opts)]
    if not prefix:
        lib.speigs_errmsg.restype = ctypes.c_char_p
        lib.speigs_errmsg.argtypes = [c_int]


def _bind_ext(lib: ctypes.CDLL) -> None:
    lib.speigs_hip_default_opts.restype = None
    lib.speigs_hip_default_opts.argtypes = [POINTER(HipOpts)]
    lib.speigs_ex.restype = c_int
    lib.speigs_ex.argtypes = [c_int64, _I64P, _I64P, _F64P, c_int64, c_int, POINTER(Opts),
                              POINTER(HipOpts), _F64P, _F64P, _F64P, POINTER(Info), POINTER(HipInfo)]
    lib.speigs_hip_device_count.restype = c_int
    lib.speigs_hip_device_count.argtypes = []
    lib.speigs_hip_last_error.restype = ctypes.c_char_p
    lib.speigs_hip_last_error.argtypes = []
    lib.speigs_hip_release_cache.restype = None
    lib.speigs_hip_release_cache.argtypes = []
    vp, u64, f32p = ctypes.c_void_p, ctypes.c_uint64, POINTER(ctypes.c_float)
    lib.speigs_hip_model_create.restype = c_int
    lib.speigs_hip_model_create.argtypes = [c_int, u64, POINTER(vp)]
    lib.speigs_hip_model_nparams.restype = c_int64
    lib.speigs_hip_model_nparams.argtypes = [vp]
    lib.speigs_hip_model_set_params.restype = c_int
    lib.speigs_hip_model_set_params.argtypes = [vp, f32p, c_int64]
    lib.speigs_hip_model_get_params.restype = c_int
    lib.speigs_hip_model_get_params.argtypes = [vp, f32p, c_int64]
    lib.speigs_hip_model_infer.restype = c_int
    lib.speigs_hip_model_infer.argtypes = [vp, c_int64, _I64P, _I64P, _F64P, u64, f32p]
    lib.speigs_hip_model_grad.restype = c_int
    lib.speigs_hip_model_grad.argtypes = [vp, c_int64, _I64P, _I64P, _F64P, _F64P, c_int64, u64, _F64P, f32p]
    lib.speigs_hip_model_train_step.restype = c_int
    lib.speigs_hip_model_train_step.argtypes = [vp, c_int64, _I64P, _I64P, _F64P, _F64P, c_int64, u64,
                                                ctypes.c_float, _F64P]
    lib.speigs_hip_model_destroy.restype = None
    lib.speigs_hip_model_destroy.argtypes = [vp]


class StartModel:
    """The learned-start network (speigs_hip_model_*): create, load/save
    parameters, infer the proposed block, loss/gradient and Adam steps."""

    WIDTH = 32

    def __init__(self, layers: int = 3, seed: int = 1, lib=None):
        self.lib = lib or load()
        h = ctypes.c_void_p()
        rc = self.lib.speigs_hip_model_create(layers, seed, ctypes.byref(h))
        if rc:
            raise RuntimeError(f"speigs_hip_model_create failed: {rc}")
        self.handle = h
        self.nparams = int(self.lib.speigs_hip_model_nparams(h))

    def close(self):
        if self.handle:
            self.lib.speigs_hip_model_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.nparams, np.float32)
        rc = self.lib.speigs_hip_model_get_params(self.handle, _ptr(p, ctypes.c_float), self.nparams)
        if rc:
            raise RuntimeError(f"get_params failed: {rc}")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        rc = self.lib.speigs_hip_model_set_params(self.handle, _ptr(p, ctypes.c_float), len(p))
        if rc:
            raise RuntimeError(f"set_params failed: {rc}")

    def save(self, path):
        self.get_params().tofile(str(path))

    def load(self, path):
        self.set_params(np.fromfile(str(path), np.float32))

    def infer(self, A, seed: int = 88172645463325252) -> np.ndarray:
        Ap, Ai, Ax = _csc(A)
        n = len(Ap) - 1
        X = np.zeros((n, self.WIDTH), np.float32)
        rc = self.lib.speigs_hip_model_infer(self.handle, n, _ptr(Ap, c_int64), _ptr(Ai, c_int64), _ptr(Ax, c_double),
                                             seed, _ptr(X, ctypes.c_float))
        if rc:
            raise RuntimeError(f"infer failed: {rc}")
        return X

    def grad(self, A, U, seed: int = 1):
        Ap, Ai, Ax = _csc(A)
        n = len(Ap) - 1
        U = np.asfortranarray(U, np.float64)
        loss = np.zeros(1)
        g = np.zeros(self.nparams, np.float32)
        rc = self.lib.speigs_hip_model_grad(self.handle, n, _ptr(Ap, c_int64), _ptr(Ai, c_int64), _ptr(Ax, c_double),
                                            U.ctypes.data_as(POINTER(c_double)), U.shape[1], seed,
                                            _ptr(loss, c_double), _ptr(g, ctypes.c_float))
        if rc:
            raise RuntimeError(f"grad failed: {rc}")
        return float(loss[0]), g

    def train_step(self, A, U, seed: int, lr: float) -> float:
        Ap, Ai, Ax = _csc(A)
        n = len(Ap) - 1
        U = np.asfortranarray(U, np.float64)
        loss = np.zeros(1)
        rc = self.lib.speigs_hip_model_train_step(self.handle, n, _ptr(Ap, c_int64), _ptr(Ai, c_int64),
                                                  _ptr(Ax, c_double), U.ctypes.data_as(POINTER(c_double)),
                                                  U.shape[1], seed, lr, _ptr(loss, c_double))
        if rc:
            raise RuntimeError(f"train_step failed: {rc}")
        return float(loss[0])


def _ptr(a, t):
    return a.ctypes.data_as(POINTER(t)) if a is not None else None


def _csc(A):
    if isinstance(A, tuple):
        Ap, Ai, Ax = A
    else:  # scipy.sparse
        A = A.tocsc()
        A.sort_indices()
        Ap, Ai, Ax = A.indptr, A.indices, A.data
    Ap = np.ascontiguousarray(Ap, dtype=np.int64)
    Ai = np.ascontiguousarray(Ai, dtype=np.int64)
    Ax = np.ascontiguousarray(Ax, dtype=np.float64)
    return Ap, Ai, Ax


@dataclass
class Result:
    rc: int
    lam: np.ndarray
    V: np.ndarray | None
    resid: np.ndarray | None
    info: dict
    hip_info: dict | None = None


def default_opts(lib=None, prefix: str = "") -> Opts:
    lib = lib or load()
    o = Opts()
    if prefix:
        getattr(lib, prefix + "default_opts")(ctypes.byref(o))
    else:
        lib.speigs_default_opts(ctypes.byref(o))
    return o


def default_hip_opts(lib=None) -> HipOpts:
    lib = lib or load()
    o = HipOpts()
    lib.speigs_hip_default_opts(ctypes.byref(o))
    return o


def speigs(A, k: int, which: str | int = "sm", opts: Opts | dict | None = None, vectors: bool = True,
           resid: bool = True, lib=None, prefix: str = "", hip_opts: HipOpts | dict | None = None) -> Result:
    """Call speigs() (or <prefix>speigs) on CSC input; mirrors the C contract."""
    if lib is None:
        lib = load()
    Ap, Ai, Ax = _csc(A)
    n = len(Ap) - 1
    w = SM if which in ("sm", SM) else LM
    o = default_opts(lib, prefix)
    if isinstance(opts, Opts):
        o = opts
    elif opts:
        for key, val in opts.items():
            setattr(o, key, val)
    kk = max(int(k), 1)
    lam = np.zeros(kk)
    V = np.zeros((kk, max(n, 1))) if vectors else None  # row t = eigenvector t (column-major n x k)
    res = np.zeros(kk) if resid else None
    info = Info()
    hinfo = None
    if hip_opts is not None and not prefix:
        ho = default_hip_opts(lib)
        if isinstance(hip_opts, HipOpts):
            ho = hip_opts
        else:
            for key, val in hip_opts.items():
                setattr(ho, key, val)
        hi = HipInfo()
        rc = lib.speigs_ex(n, _ptr(Ap, c_int64), _ptr(Ai, c_int64), _ptr(Ax, c_double), int(k), w,
                           ctypes.byref(o), ctypes.byref(ho), _ptr(lam, c_double), _ptr(V, c_double),
                           _ptr(res, c_double), ctypes.byref(info), ctypes.byref(hi))
        hinfo = hi.as_dict()
    else:
        fn = getattr(lib, prefix + "speigs")
        rc = fn(n, _ptr(Ap, c_int64), _ptr(Ai, c_int64), _ptr(Ax, c_double), int(k), w, ctypes.byref(o),
                _ptr(lam, c_double), _ptr(V, c_double), _ptr(res, c_double), ctypes.byref(info))
    # V[:, :n].T is the n x k column-major result as a Fortran-ordered view (no copy)
    return Result(rc, lam, None if V is None else V[:, :n].T, res, info.as_dict(), hinfo)


def speigs_z(A_p, A_i, Axr, Axi, k: int, which="sm", opts: dict | None = None, lib=None, prefix: str = ""):
    """Complex Hermitian entry (split real/imaginary values)."""
    if lib is None:
        lib = load()
    Ap, Ai, Axr = _csc((A_p, A_i, Axr))
    Axi = None if Axi is None else np.ascontiguousarray(Axi, dtype=np.float64)
    n = len(Ap) - 1
    o = default_opts(lib, prefix)
    for key, val in (opts or {}).items():
        setattr(o, key, val)
    lam = np.zeros(k)
    Vr = np.zeros((k, n))
    Vi = np.zeros((k, n))
    res = np.zeros(k)
    info = Info()
    fn = getattr(lib, prefix + "speigs_z")
    rc = fn(n, _ptr(Ap, c_int64), _ptr(Ai, c_int64), _ptr(Axr, c_double), _ptr(Axi, c_double), int(k),
            SM if which in ("sm", SM) else LM, ctypes.byref(o), _ptr(lam, c_double), _ptr(Vr, c_double),
            _ptr(Vi, c_double), _ptr(res, c_double), ctypes.byref(info))
    return Result(rc, lam, (Vr.T + 1j * Vi.T), res, info.as_dict())


def release_cache(lib=None) -> None:
    """Return the device memory the library keeps between calls to the HIP
    runtime (speigs_hip_release_cache)."""
    (lib or load()).speigs_hip_release_cache()


def last_error(lib=None) -> str:
    """Where this thread's most recent internal failure happened
    (speigs_hip_last_error): "file:line: step (code c)", or ""."""
    return (lib or load()).speigs_hip_last_error().decode()


def errmsg(code: int, lib=None) -> str:
    lib = lib or load()
    return lib.speigs_errmsg(int(code)).decode()


from . import gen  # noqa: E402  (workload generators; no library needed)
from .mmread import MMMatrix, mmread, mmwrite  # noqa: E402  (Matrix Market files; no library needed)
