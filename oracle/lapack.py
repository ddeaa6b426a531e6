"""Host LAPACK adapter (ctypes) -- TEST INFRASTRUCTURE ONLY.

The reference clients compute their goldens by calling Fortran LAPACK through
``cpu_getrf`` / ``cpu_getrs`` / ``cpu_gesv`` (clients/common/lapack_host_reference.cpp,
which forwards to dgetrf_/dgetrs_/dgesv_).  This module reaches the same LAPACK
entry points of whatever host LAPACK the machine has, searched in order:
liblapack.so.3, libopenblas.so.0, libflexiblas.so.3, then the OpenBLAS that
SciPy bundles (symbols prefixed ``scipy_``).  Used to pin the oracle's
restatement and as the CPU baseline.
"""
from __future__ import annotations

import ctypes
import glob
import os

import numpy as np

_lib = None
_prefix = ""
_name = None


def _candidates():
    yield "liblapack.so.3", ""
    yield "libopenblas.so.0", ""
    yield "libflexiblas.so.3", ""
    try:
        import scipy  # noqa: F401  (only to locate its bundled OpenBLAS)
        root = os.path.dirname(os.path.dirname(scipy.__file__))
        for p in sorted(glob.glob(os.path.join(root, "scipy.libs", "libscipy_openblas*.so"))):
            yield p, "scipy_"
    except Exception:
        pass


def load():
    """Return (ctypes lib, symbol prefix, description) or raise OSError."""
    global _lib, _prefix, _name
    if _lib is not None:
        return _lib, _prefix, _name
    for path, pre in _candidates():
        try:
            L = ctypes.CDLL(path)
            getattr(L, pre + "dgetrf_")
        except (OSError, AttributeError):
            continue
        _lib, _prefix = L, pre
        desc = os.path.basename(path)
        try:
            cfg = getattr(L, pre + "openblas_get_config")
            cfg.restype = ctypes.c_char_p
            desc += " (" + cfg().decode() + ")"
        except AttributeError:
            pass
        _name = desc
        return _lib, _prefix, _name
    raise OSError("no host LAPACK found")


def available() -> bool:
    try:
        load()
        return True
    except OSError:
        return False


def set_num_threads(n: int) -> None:
    L, pre, _ = load()
    for sym in (pre + "openblas_set_num_threads", "openblas_set_num_threads"):
        f = getattr(L, sym, None)
        if f is not None:
            f.argtypes = [ctypes.c_int]
            f(int(n))
            return


def _fn(name):
    L, pre, _ = load()
    return getattr(L, pre + name)


_I = ctypes.c_int
_P = ctypes.c_void_p


def _byref_int(v):
    return ctypes.byref(_I(int(v)))


def getrf(A: np.ndarray, m: int):
    """xGETRF on every matrix of a (bc, n, lda) batch, in place -> (ipiv, info)."""
    if A.ndim == 2:
        A = A[None]
    bc, n, lda = A.shape
    pre = "d" if A.dtype == np.float64 else "s"
    f = _fn(pre + "getrf_")
    k = min(m, n)
    ipiv = np.zeros((bc, max(k, 1)), dtype=np.int32)
    info = np.zeros(bc, dtype=np.int32)
    inf = _I()
    for b in range(bc):
        f(_byref_int(m), _byref_int(n), _P(A[b].ctypes.data), _byref_int(lda),
          _P(ipiv[b].ctypes.data), ctypes.byref(inf))
        info[b] = inf.value
    return ipiv[:, :k], info


def getf2(A: np.ndarray, m: int):
    """xGETF2 (unblocked LAPACK LU) per batch entry, in place -> (ipiv, info)."""
    if A.ndim == 2:
        A = A[None]
    bc, n, lda = A.shape
    pre = "d" if A.dtype == np.float64 else "s"
    f = _fn(pre + "getf2_")
    k = min(m, n)
    ipiv = np.zeros((bc, max(k, 1)), dtype=np.int32)
    info = np.zeros(bc, dtype=np.int32)
    inf = _I()
    for b in range(bc):
        f(_byref_int(m), _byref_int(n), _P(A[b].ctypes.data), _byref_int(lda),
          _P(ipiv[b].ctypes.data), ctypes.byref(inf))
        info[b] = inf.value
    return ipiv[:, :k], info


def getrs(trans: str, LU: np.ndarray, n: int, ipiv: np.ndarray, B: np.ndarray) -> None:
    if LU.ndim == 2:
        LU = LU[None]
    if B.ndim == 2:
        B = B[None]
    bc, _, lda = LU.shape
    _, nrhs, ldb = B.shape
    pre = "d" if LU.dtype == np.float64 else "s"
    f = _fn(pre + "getrs_")
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32).reshape(bc, -1)
    t = ctypes.c_char(trans.upper().encode()[:1])
    inf = _I()
    for b in range(bc):
        f(ctypes.byref(t), _byref_int(n), _byref_int(nrhs), _P(LU[b].ctypes.data),
          _byref_int(lda), _P(ipiv[b].ctypes.data), _P(B[b].ctypes.data), _byref_int(ldb),
          ctypes.byref(inf), ctypes.c_size_t(1))


def gesv(A: np.ndarray, n: int, B: np.ndarray):
    if A.ndim == 2:
        A = A[None]
    if B.ndim == 2:
        B = B[None]
    bc, _, lda = A.shape
    _, nrhs, ldb = B.shape
    pre = "d" if A.dtype == np.float64 else "s"
    f = _fn(pre + "gesv_")
    ipiv = np.zeros((bc, max(n, 1)), dtype=np.int32)
    info = np.zeros(bc, dtype=np.int32)
    inf = _I()
    for b in range(bc):
        f(_byref_int(n), _byref_int(nrhs), _P(A[b].ctypes.data), _byref_int(lda),
          _P(ipiv[b].ctypes.data), _P(B[b].ctypes.data), _byref_int(ldb), ctypes.byref(inf))
        info[b] = inf.value
    return ipiv[:, :n], info


def gesv_one(n: int, nrhs: int, a_ptr: int, lda: int, ipiv_ptr: int, b_ptr: int, ldb: int) -> int:
    """One dgesv_ call on raw pointers (used by the timed CPU-baseline loop)."""
    f = _fn("dgesv_")
    inf = _I()
    f(_byref_int(n), _byref_int(nrhs), _P(a_ptr), _byref_int(lda), _P(ipiv_ptr), _P(b_ptr),
      _byref_int(ldb), ctypes.byref(inf))
    return inf.value


def getri(LU: np.ndarray, n: int, ipiv: np.ndarray):
    """xGETRI on every matrix of a (bc, n, lda) batch of getrf factors, in place -> info."""
    if LU.ndim == 2:
        LU = LU[None]
    bc, _, lda = LU.shape
    pre = "d" if LU.dtype == np.float64 else "s"
    f = _fn(pre + "getri_")
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32).reshape(bc, -1)
    lwork = max(1, 64 * n)
    work = np.zeros(lwork, dtype=LU.dtype)
    info = np.zeros(bc, dtype=np.int32)
    inf = _I()
    for b in range(bc):
        f(_byref_int(n), _P(LU[b].ctypes.data), _byref_int(lda), _P(ipiv[b].ctypes.data),
          _P(work.ctypes.data), _byref_int(lwork), ctypes.byref(inf))
        info[b] = inf.value
    return info


def potrf(A: np.ndarray, n: int, uplo: str):
    """xPOTRF per batch entry, in place -> info."""
    if A.ndim == 2:
        A = A[None]
    bc, _, lda = A.shape
    pre = "d" if A.dtype == np.float64 else "s"
    f = _fn(pre + "potrf_")
    u = ctypes.c_char(uplo.upper().encode()[:1])
    info = np.zeros(bc, dtype=np.int32)
    inf = _I()
    for b in range(bc):
        f(ctypes.byref(u), _byref_int(n), _P(A[b].ctypes.data), _byref_int(lda), ctypes.byref(inf),
          ctypes.c_size_t(1))
        info[b] = inf.value
    return info


def potrs(A: np.ndarray, n: int, B: np.ndarray, uplo: str) -> None:
    """xPOTRS per batch entry, B (bc, nrhs, ldb) in place."""
    if A.ndim == 2:
        A = A[None]
    if B.ndim == 2:
        B = B[None]
    bc, _, lda = A.shape
    _, nrhs, ldb = B.shape
    pre = "d" if A.dtype == np.float64 else "s"
    f = _fn(pre + "potrs_")
    u = ctypes.c_char(uplo.upper().encode()[:1])
    inf = _I()
    for b in range(bc):
        f(ctypes.byref(u), _byref_int(n), _byref_int(nrhs), _P(A[b].ctypes.data), _byref_int(lda),
          _P(B[b].ctypes.data), _byref_int(ldb), ctypes.byref(inf), ctypes.c_size_t(1))


def dsgesv(A: np.ndarray, n: int, B: np.ndarray):
    """LAPACK DSGESV per batch entry -> (X, iter, info); A is overwritten only on fallback."""
    if A.ndim == 2:
        A = A[None]
    if B.ndim == 2:
        B = B[None]
    bc, _, lda = A.shape
    _, nrhs, ldb = B.shape
    f = _fn("dsgesv_")
    X = np.zeros((bc, nrhs, n), dtype=np.float64)
    it = np.zeros(bc, dtype=np.int32)
    info = np.zeros(bc, dtype=np.int32)
    work = np.zeros(n * nrhs, dtype=np.float64)
    swork = np.zeros(n * (n + nrhs), dtype=np.float32)
    ipiv = np.zeros(n, dtype=np.int32)
    itv, inf = _I(), _I()
    for b in range(bc):
        a = A[b].copy()
        f(_byref_int(n), _byref_int(nrhs), _P(a.ctypes.data), _byref_int(lda),
          _P(ipiv.ctypes.data), _P(B[b].ctypes.data), _byref_int(ldb), _P(X[b].ctypes.data),
          _byref_int(n), _P(work.ctypes.data), _P(swork.ctypes.data), ctypes.byref(itv),
          ctypes.byref(inf))
        it[b] = itv.value
        info[b] = inf.value
    return X, it, info
