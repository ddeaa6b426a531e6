"""CPU oracle for the batched LU path -- TEST INFRASTRUCTURE ONLY.

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` import it; the product library
(``rocsolver_amd``) never does, and never falls back to it.

Two independent CPU implementations live here:

* ``liboracle.so`` (``oracle.c``): a restatement of the reference algorithms
  (getf2 / laswp / getrs / gesv, plus LAPACK dsgesv refinement semantics),
  citing the reference file:line it follows.
* ``lapack``: a ctypes adapter over a host LAPACK (the one the reference
  clients call through ``cpu_getrf``/``cpu_getrs``/``cpu_gesv``), used to pin
  the restatement and as the CPU baseline.

All arrays are numpy, column-major per matrix: a batch of matrices is an
array of shape ``(bc, n_cols, ld)`` -- C order over (batch, column, row) --
which is exactly the strided-batched device layout ``A[b*strideA + i + j*lda]``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> str:
    """Compile oracle.c with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        i, d, p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
        c = ctypes.c_char
        for pre in ("d", "s"):
            getattr(L, f"orc_{pre}getf2").argtypes = [i, i, p, i, p, p]
            getattr(L, f"orc_{pre}getf2").restype = i
            getattr(L, f"orc_{pre}getf2_npvt").argtypes = [i, i, p, i]
            getattr(L, f"orc_{pre}getf2_npvt").restype = i
            getattr(L, f"orc_{pre}laswp").argtypes = [i, p, i, i, i, p, i]
            getattr(L, f"orc_{pre}laswp").restype = None
            getattr(L, f"orc_{pre}getrs").argtypes = [c, i, i, p, i, p, p, i]
            getattr(L, f"orc_{pre}getrs").restype = None
            getattr(L, f"orc_{pre}gesv").argtypes = [i, i, p, i, p, p, i]
            getattr(L, f"orc_{pre}gesv").restype = i
            getattr(L, f"orc_{pre}getri").argtypes = [i, p, i, p]
            getattr(L, f"orc_{pre}getri").restype = i
            getattr(L, f"orc_{pre}potf2").argtypes = [c, i, p, i]
            getattr(L, f"orc_{pre}potf2").restype = i
            getattr(L, f"orc_{pre}potrs").argtypes = [c, i, i, p, i, p, i]
            getattr(L, f"orc_{pre}potrs").restype = None
            getattr(L, f"orc_{pre}posv").argtypes = [c, i, i, p, i, p, i]
            getattr(L, f"orc_{pre}posv").restype = i
        L.orc_dsgesv.argtypes = [i, i, p, i, p, i, p, i, i, p]
        L.orc_dsgesv.restype = i
        L.orc_dresidual.argtypes = [c, i, i, p, i, p, i, p, i, p, i]
        L.orc_dresidual.restype = None
        L.orc_dnorm_fro.argtypes = [i, i, p, i]
        L.orc_dnorm_fro.restype = d
        L.orc_dnorm_inf.argtypes = [i, i, p, i]
        L.orc_dnorm_inf.restype = d
        L.orc_dlu_residual.argtypes = [i, i, p, i, p, i, p]
        L.orc_dlu_residual.restype = d
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _prefix(dtype) -> str:
    if dtype == np.float64:
        return "d"
    if dtype == np.float32:
        return "s"
    raise TypeError(f"oracle supports float32/float64, got {dtype}")


def _as_batch(A: np.ndarray) -> np.ndarray:
    if A.ndim == 2:
        A = A[None]
    assert A.flags.c_contiguous
    return A


def getf2(A: np.ndarray, m: int, with_gap: bool = False):
    """LU with partial pivoting of every matrix of a (bc, n, lda) batch, in place.

    Returns (ipiv[bc, min(m,n)] int32 1-based, info[bc] int32[, gap[bc, min(m,n)]]).
    """
    A = _as_batch(A)
    bc, n, lda = A.shape
    k = min(m, n)
    ipiv = np.zeros((bc, max(k, 1)), dtype=np.int32)
    info = np.zeros(bc, dtype=np.int32)
    gap = np.zeros((bc, max(k, 1)), dtype=np.float64) if with_gap else None
    f = getattr(lib(), f"orc_{_prefix(A.dtype)}getf2")
    for b in range(bc):
        info[b] = f(m, n, _ptr(A[b]), lda, _ptr(ipiv[b]), _ptr(gap[b]) if with_gap else None)
    ipiv = ipiv[:, :k]
    if with_gap:
        return ipiv, info, gap[:, :k]
    return ipiv, info


def getf2_npvt(A: np.ndarray, m: int):
    """LU without pivoting of every matrix of a (bc, n, lda) batch, in place.

    Returns info[bc] int32 (first zero diagonal pivot, 1-based, or 0)."""
    A = _as_batch(A)
    bc, n, lda = A.shape
    info = np.zeros(bc, dtype=np.int32)
    f = getattr(lib(), f"orc_{_prefix(A.dtype)}getf2_npvt")
    for b in range(bc):
        info[b] = f(m, n, _ptr(A[b]), lda)
    return info


def laswp(B: np.ndarray, k1: int, k2: int, ipiv: np.ndarray, incx: int) -> None:
    """Row interchanges on every matrix of a (bc, ncols, ldb) batch, in place."""
    B = _as_batch(B)
    bc, ncols, ldb = B.shape
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32).reshape(bc, -1)
    f = getattr(lib(), f"orc_{_prefix(B.dtype)}laswp")
    for b in range(bc):
        f(ncols, _ptr(B[b]), ldb, k1, k2, _ptr(ipiv[b]), incx)


def getrs(trans: str, LU: np.ndarray, n: int, ipiv: np.ndarray, B: np.ndarray) -> None:
    """Solve op(A) X = B for each batch entry, B (bc, nrhs, ldb) overwritten."""
    LU = _as_batch(LU)
    B = _as_batch(B)
    bc, _, lda = LU.shape
    _, nrhs, ldb = B.shape
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32).reshape(bc, -1)
    f = getattr(lib(), f"orc_{_prefix(LU.dtype)}getrs")
    t = trans.upper().encode()[:1]
    for b in range(bc):
        f(t, n, nrhs, _ptr(LU[b]), lda, _ptr(ipiv[b]), _ptr(B[b]), ldb)


def gesv(A: np.ndarray, n: int, B: np.ndarray):
    """getf2 + getrs per batch entry, in place; returns (ipiv, info)."""
    A = _as_batch(A)
    B = _as_batch(B)
    bc, _, lda = A.shape
    _, nrhs, ldb = B.shape
    ipiv = np.zeros((bc, max(n, 1)), dtype=np.int32)
    info = np.zeros(bc, dtype=np.int32)
    f = getattr(lib(), f"orc_{_prefix(A.dtype)}gesv")
    for b in range(bc):
        info[b] = f(n, nrhs, _ptr(A[b]), lda, _ptr(ipiv[b]), _ptr(B[b]), ldb)
    return ipiv[:, :n], info


def getri(LU: np.ndarray, n: int, ipiv):
    """Inverse from the getf2 output of every matrix of a (bc, n, lda) batch, in
    place; ipiv None for the no-pivoting variants.  Returns info[bc]."""
    LU = _as_batch(LU)
    bc, _, lda = LU.shape
    info = np.zeros(bc, dtype=np.int32)
    f = getattr(lib(), f"orc_{_prefix(LU.dtype)}getri")
    if ipiv is not None:
        ipiv = np.ascontiguousarray(ipiv, dtype=np.int32).reshape(bc, -1)
    for b in range(bc):
        info[b] = f(n, _ptr(LU[b]), lda, None if ipiv is None else _ptr(ipiv[b]))
    return info


def potf2(A: np.ndarray, n: int, uplo: str):
    """Cholesky factorisation (uplo 'L' or 'U') of every matrix of a batch, in
    place.  Returns info[bc]."""
    A = _as_batch(A)
    bc, _, lda = A.shape
    info = np.zeros(bc, dtype=np.int32)
    f = getattr(lib(), f"orc_{_prefix(A.dtype)}potf2")
    u = uplo.upper().encode()
    for b in range(bc):
        info[b] = f(u, n, _ptr(A[b]), lda)
    return info


def potrs(A: np.ndarray, n: int, B: np.ndarray, uplo: str) -> None:
    """Solve with the Cholesky factor of potf2, B (bc, nrhs, ldb) in place."""
    A, B = _as_batch(A), _as_batch(B)
    bc, _, lda = A.shape
    _, nrhs, ldb = B.shape
    f = getattr(lib(), f"orc_{_prefix(A.dtype)}potrs")
    u = uplo.upper().encode()
    for b in range(bc):
        f(u, n, nrhs, _ptr(A[b]), lda, _ptr(B[b]), ldb)


def posv(A: np.ndarray, n: int, B: np.ndarray, uplo: str):
    """potf2 + potrs per batch entry (B unchanged where info != 0) -> info."""
    A, B = _as_batch(A), _as_batch(B)
    bc, _, lda = A.shape
    _, nrhs, ldb = B.shape
    info = np.zeros(bc, dtype=np.int32)
    f = getattr(lib(), f"orc_{_prefix(A.dtype)}posv")
    u = uplo.upper().encode()
    for b in range(bc):
        info[b] = f(u, n, nrhs, _ptr(A[b]), lda, _ptr(B[b]), ldb)
    return info


def dsgesv(A: np.ndarray, n: int, B: np.ndarray, itermax: int = 30):
    """Mixed-precision refinement per batch entry (LAPACK dsgesv semantics).

    Returns (X (bc, nrhs, n), iter[bc], info[bc]); A and B are not modified."""
    A = _as_batch(A)
    B = _as_batch(B)
    bc, _, lda = A.shape
    _, nrhs, ldb = B.shape
    X = np.zeros((bc, nrhs, max(n, 1)), dtype=np.float64)
    it = np.zeros(bc, dtype=np.int32)
    info = np.zeros(bc, dtype=np.int32)
    one = ctypes.c_int()
    for b in range(bc):
        info[b] = lib().orc_dsgesv(n, nrhs, _ptr(A[b]), lda, _ptr(B[b]), ldb, _ptr(X[b]),
                                   max(n, 1), itermax, ctypes.byref(one))
        it[b] = one.value
    return X, it, info


def residual(A: np.ndarray, n: int, X: np.ndarray, B: np.ndarray, trans: str = "N") -> np.ndarray:
    """R = B - op(A) X per batch entry (long double accumulation)."""
    A = _as_batch(np.ascontiguousarray(A, dtype=np.float64))
    X = _as_batch(np.ascontiguousarray(X, dtype=np.float64))
    B = _as_batch(np.ascontiguousarray(B, dtype=np.float64))
    bc, _, lda = A.shape
    _, nrhs, ldx = X.shape
    ldb = B.shape[2]
    R = np.zeros((bc, nrhs, max(n, 1)), dtype=np.float64)
    t = trans.upper().encode()[:1]
    for b in range(bc):
        lib().orc_dresidual(t, n, nrhs, _ptr(A[b]), lda, _ptr(X[b]), ldx, _ptr(B[b]), ldb,
                            _ptr(R[b]), max(n, 1))
    return R


def norm_inf(A: np.ndarray, m: int, n: int) -> np.ndarray:
    A = _as_batch(np.ascontiguousarray(A, dtype=np.float64))
    return np.array([lib().orc_dnorm_inf(m, n, _ptr(A[b]), A.shape[2]) for b in range(A.shape[0])])


def norm_fro(A: np.ndarray, m: int, n: int) -> np.ndarray:
    A = _as_batch(np.ascontiguousarray(A, dtype=np.float64))
    return np.array([lib().orc_dnorm_fro(m, n, _ptr(A[b]), A.shape[2]) for b in range(A.shape[0])])


def lu_backward_error(A0: np.ndarray, LU: np.ndarray, ipiv: np.ndarray, m: int, n: int,
                      eps: float) -> np.ndarray:
    """||P A - L U||_F / (n ||A||_F eps) per batch entry."""
    A0 = _as_batch(np.ascontiguousarray(A0, dtype=np.float64))
    LU = _as_batch(np.ascontiguousarray(LU, dtype=np.float64))
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32).reshape(A0.shape[0], -1)
    out = np.zeros(A0.shape[0])
    for b in range(A0.shape[0]):
        r = lib().orc_dlu_residual(m, n, _ptr(A0[b]), A0.shape[2], _ptr(LU[b]), LU.shape[2],
                                   _ptr(ipiv[b]))
        an = lib().orc_dnorm_fro(m, n, _ptr(A0[b]), A0.shape[2])
        out[b] = r / (max(n, 1) * max(an, 1e-300) * eps)
    return out


def solve_backward_error(A: np.ndarray, n: int, X: np.ndarray, B: np.ndarray,
                         trans: str = "N") -> np.ndarray:
    """Normwise backward error max_c ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)."""
    R = residual(A, n, X, B, trans)
    An = norm_inf(A, n, n) if trans.upper() == "N" else norm_inf(np.swapaxes(A, 1, 2).copy(), n, n)
    X = _as_batch(np.asarray(X, dtype=np.float64))
    B = _as_batch(np.asarray(B, dtype=np.float64))
    xn = np.abs(X[:, :, :n]).max(axis=2)
    bn = np.abs(B[:, :, :n]).max(axis=2)
    rn = np.abs(R[:, :, :n]).max(axis=2)
    den = An[:, None] * xn + bn
    return (rn / np.where(den > 0, den, 1.0)).max(axis=1)
