"""rocsolver_amd -- MI355X-native batched LU factor-and-solve engine.

The product is the C-ABI shared library ``librocsolver_amd.so`` (declared in
``include/rocsolver_amd.h``): rocsolver_{s,d}{getrf,getf2,getrf_npvt,getrs,gesv}
in plain, ``_batched`` and ``_strided_batched`` forms, and rocsolver_{s,d}laswp,
with the rocsolver.h signatures and semantics.  This module is a thin Python
host mirror of that boundary: it loads the library (loudly failing when it is
missing -- there is no fallback path), wraps a ``rocblas_handle`` and offers
torch-tensor convenience calls with the reference's argument meaning.

Batch layout used by the helpers: a tensor of shape (bc, ncols, ld), C
contiguous, i.e. matrix b is column-major at ``base + b*ncols*ld`` -- exactly
the strided-batched device layout with strideA = ncols*ld.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librocsolver_amd.so")
ROCBLAS_PATH = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "librocblas.so")

# rocblas_status values (rocblas-types.h)
STATUS = {
    0: "success", 1: "invalid_handle", 2: "not_implemented", 3: "invalid_pointer",
    4: "invalid_size", 5: "memory_error", 6: "internal_error", 7: "perf_degraded",
    8: "size_query_mismatch", 9: "size_increased", 10: "size_unchanged", 11: "invalid_value",
    12: "continue",
}
OP_NONE, OP_TRANSPOSE, OP_CONJ_TRANSPOSE = 111, 112, 113

_lib = None
_rocblas = None


class RocsolverError(RuntimeError):
    def __init__(self, name: str, status: int):
        super().__init__(f"{name} returned rocblas_status {status} ({STATUS.get(status, '?')})")
        self.status = status


def lib() -> ctypes.CDLL:
    """The engine's shared library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with `python -m rocsolver_amd.build` "
                          "(there is no fallback implementation)")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
        _install_tuning(_lib)
    return _lib


TUNING_PATH = os.path.join(HERE, "tuning", "vbatched_gfx950.json")


def _install_tuning(L: ctypes.CDLL) -> None:
    """Install the autotuned variable-size schedule when one was saved."""
    if not os.path.exists(TUNING_PATH):
        return
    import json
    with open(TUNING_PATH) as f:
        vals = [int(v) for v in json.load(f)["table"]]
    buf = (ctypes.c_int * len(vals))(*vals)
    st = L.rocsolver_amd_vbatched_set_tuning(ctypes.cast(buf, ctypes.c_void_p), len(vals))
    if st != 0:
        raise RocsolverError(f"invalid schedule table in {TUNING_PATH}", st)


def rocblas() -> ctypes.CDLL:
    global _rocblas
    if _rocblas is None:
        _rocblas = ctypes.CDLL(ROCBLAS_PATH)
        _rocblas.rocblas_create_handle.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _rocblas.rocblas_destroy_handle.argtypes = [ctypes.c_void_p]
        _rocblas.rocblas_set_stream.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _rocblas.rocblas_start_device_memory_size_query.argtypes = [ctypes.c_void_p]
        _rocblas.rocblas_stop_device_memory_size_query.argtypes = [
            ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
    return _rocblas


_I, _P, _S = ctypes.c_int, ctypes.c_void_p, ctypes.c_int64

# argument lists, in declaration order (include/rocsolver_amd.h)
SIGNATURES = {
    "getrf": [_P, _I, _I, _P, _I, _P, _P],
    "getrf_batched": [_P, _I, _I, _P, _I, _P, _S, _P, _I],
    "getrf_strided_batched": [_P, _I, _I, _P, _I, _S, _P, _S, _P, _I],
    "getf2": [_P, _I, _I, _P, _I, _P, _P],
    "getf2_batched": [_P, _I, _I, _P, _I, _P, _S, _P, _I],
    "getf2_strided_batched": [_P, _I, _I, _P, _I, _S, _P, _S, _P, _I],
    "getrf_npvt": [_P, _I, _I, _P, _I, _P],
    "getrf_npvt_batched": [_P, _I, _I, _P, _I, _P, _I],
    "getrf_npvt_strided_batched": [_P, _I, _I, _P, _I, _S, _P, _I],
    "getrs": [_P, _I, _I, _I, _P, _I, _P, _P, _I],
    "getrs_batched": [_P, _I, _I, _I, _P, _I, _P, _S, _P, _I, _I],
    "getrs_strided_batched": [_P, _I, _I, _I, _P, _I, _S, _P, _S, _P, _I, _S, _I],
    "gesv": [_P, _I, _I, _P, _I, _P, _P, _I, _P],
    "gesv_batched": [_P, _I, _I, _P, _I, _P, _S, _P, _I, _P, _I],
    "gesv_strided_batched": [_P, _I, _I, _P, _I, _S, _P, _S, _P, _I, _S, _P, _I],
    "laswp": [_P, _I, _P, _I, _I, _I, _P, _I],
    "potf2": [_P, _I, _I, _P, _I, _P],
    "potf2_batched": [_P, _I, _I, _P, _I, _P, _I],
    "potf2_strided_batched": [_P, _I, _I, _P, _I, _S, _P, _I],
    "potrf": [_P, _I, _I, _P, _I, _P],
    "potrf_batched": [_P, _I, _I, _P, _I, _P, _I],
    "potrf_strided_batched": [_P, _I, _I, _P, _I, _S, _P, _I],
    "potrs": [_P, _I, _I, _I, _P, _I, _P, _I],
    "potrs_batched": [_P, _I, _I, _I, _P, _I, _P, _I, _I],
    "potrs_strided_batched": [_P, _I, _I, _I, _P, _I, _S, _P, _I, _S, _I],
    "posv": [_P, _I, _I, _I, _P, _I, _P, _I, _P],
    "posv_batched": [_P, _I, _I, _I, _P, _I, _P, _I, _P, _I],
    "posv_strided_batched": [_P, _I, _I, _I, _P, _I, _S, _P, _I, _S, _P, _I],
    "getri": [_P, _I, _P, _I, _P, _P],
    "getri_batched": [_P, _I, _P, _I, _P, _S, _P, _I],
    "getri_strided_batched": [_P, _I, _P, _I, _S, _P, _S, _P, _I],
    "getri_npvt": [_P, _I, _P, _I, _P],
    "getri_npvt_batched": [_P, _I, _P, _I, _P, _I],
    "getri_npvt_strided_batched": [_P, _I, _P, _I, _S, _P, _I],
    "getri_outofplace": [_P, _I, _P, _I, _P, _P, _I, _P],
    "getri_outofplace_batched": [_P, _I, _P, _I, _P, _S, _P, _I, _P, _I],
    "getri_outofplace_strided_batched": [_P, _I, _P, _I, _S, _P, _S, _P, _I, _S, _P, _I],
    "getri_npvt_outofplace": [_P, _I, _P, _I, _P, _I, _P],
    "getri_npvt_outofplace_batched": [_P, _I, _P, _I, _P, _I, _P, _I],
    "getri_npvt_outofplace_strided_batched": [_P, _I, _P, _I, _S, _P, _I, _S, _P, _I],
}


EXTENSIONS = {
    "rocsolver_amd_dsgesv_strided_batched": [_P, _I, _I, _P, _I, _S, _P, _I, _S, _P, _I, _S, _P,
                                             _P, _P, _I],
    "rocsolver_amd_dsgetrf_strided_batched": [_P, _I, _P, _I, _S, _P, _I, _S, _P, _S, _P, _I],
    "rocsolver_amd_sgesv_adjoint_strided_batched": [_P, _I, _I, _P, _I, _S, _P, _S, _P, _I, _S,
                                                    _P, _I, _S, _P, _I, _S, _P, _I, _S, _I],
    "rocsolver_amd_dgesv_adjoint_strided_batched": [_P, _I, _I, _P, _I, _S, _P, _S, _P, _I, _S,
                                                    _P, _I, _S, _P, _I, _S, _P, _I, _S, _I],
    "rocsolver_amd_sgetrf_vbatched": [_P, _P, _P, _P, _P, _P, _I],
    "rocsolver_amd_dgetrf_vbatched": [_P, _P, _P, _P, _P, _P, _I],
    "rocsolver_amd_sgetrs_vbatched": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _I],
    "rocsolver_amd_dgetrs_vbatched": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _I],
    "rocsolver_amd_sgesv_vbatched": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I],
    "rocsolver_amd_dgesv_vbatched": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I],
    "rocsolver_amd_dgesv_vbatched_noalloc": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I],
    "rocsolver_amd_dxgesv_strided_batched": [_P, _I, _I, _I, _P, _I, _S, _P, _I, _S, _P, _I, _S,
                                             _P, _P, _P, _I],
    "rocsolver_amd_vbatched_set_tuning": [_P, _I],
    "rocsolver_log_begin": [],
    "rocsolver_log_end": [],
    "rocsolver_log_set_layer_mode": [ctypes.c_uint32],
    "rocsolver_log_set_max_levels": [_I],
    "rocsolver_log_restore_defaults": [],
    "rocsolver_log_write_profile": [],
    "rocsolver_log_flush_profile": [],
    "rocsolver_amd_vbatched_get_tuning": [_P, _I],
}


def _declare(L: ctypes.CDLL) -> None:
    for base, args in SIGNATURES.items():
        for pre in ("s", "d"):
            f = getattr(L, f"rocsolver_{pre}{base}")
            f.argtypes = args
            f.restype = ctypes.c_int
    for name, args in EXTENSIONS.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.rocsolver_get_version_string.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    L.rocsolver_get_version_string_size.argtypes = [ctypes.POINTER(ctypes.c_size_t)]


def exported_symbols() -> list[str]:
    out = [f"rocsolver_{p}{b}" for b in SIGNATURES for p in ("s", "d")]
    return out + list(EXTENSIONS) + ["rocsolver_get_version_string",
                                     "rocsolver_get_version_string_size"]


def version() -> str:
    n = ctypes.c_size_t()
    lib().rocsolver_get_version_string_size(ctypes.byref(n))
    buf = ctypes.create_string_buffer(n.value)
    lib().rocsolver_get_version_string(buf, n.value)
    return buf.value.decode()


def call(name: str, *args) -> int:
    """Call rocsolver_<name> with raw arguments; returns the status."""
    return getattr(lib(), name)(*args)


def check(name: str, *args) -> None:
    st = call(name, *args)
    if st != 0:
        raise RocsolverError(name, st)


class Handle:
    """A rocblas_handle bound to torch's current HIP stream."""

    def __init__(self, stream=None):
        import torch
        h = ctypes.c_void_p()
        st = rocblas().rocblas_create_handle(ctypes.byref(h))
        if st != 0:
            raise RocsolverError("rocblas_create_handle", st)
        self.value = h.value
        s = stream if stream is not None else torch.cuda.current_stream()
        rocblas().rocblas_set_stream(self.value, ctypes.c_void_p(s.cuda_stream))

    def set_stream(self, stream) -> None:
        rocblas().rocblas_set_stream(self.value, ctypes.c_void_p(stream.cuda_stream))

    def close(self) -> None:
        if self.value:
            rocblas().rocblas_destroy_handle(self.value)
            self.value = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _prefix(t) -> str:
    import torch
    if t.dtype == torch.float64:
        return "d"
    if t.dtype == torch.float32:
        return "s"
    raise TypeError(f"float32/float64 only, got {t.dtype}")


def pointer_array(t):
    """Device array of per-matrix pointers for the _batched forms."""
    import torch
    bc = t.shape[0]
    stride = t.stride(0) * t.element_size()
    base = t.data_ptr()
    ptrs = torch.arange(bc, dtype=torch.int64) * stride + base
    return ptrs.to(t.device)


# ------------------------------------------------------------ torch helpers
def getrf(handle: Handle, A, m: int | None = None, pivot: bool = True, batched: bool = False):
    """LU of every matrix of A (bc, n, lda) in place -> (ipiv[bc, min(m,n)], info[bc])."""
    import torch
    bc, n, lda = A.shape
    m = lda if m is None else m
    k = min(m, n)
    dev = A.device
    ipiv = torch.zeros((bc, max(k, 1)), dtype=torch.int32, device=dev)
    info = torch.zeros(bc, dtype=torch.int32, device=dev)
    p = _prefix(A)
    if batched:
        ptrs = pointer_array(A)
        if pivot:
            check(f"rocsolver_{p}getrf_batched", handle.value, m, n, ptrs.data_ptr(), lda,
                  ipiv.data_ptr(), ipiv.shape[1], info.data_ptr(), bc)
        else:
            check(f"rocsolver_{p}getrf_npvt_batched", handle.value, m, n, ptrs.data_ptr(), lda,
                  info.data_ptr(), bc)
    else:
        if pivot:
            check(f"rocsolver_{p}getrf_strided_batched", handle.value, m, n, A.data_ptr(), lda,
                  n * lda, ipiv.data_ptr(), ipiv.shape[1], info.data_ptr(), bc)
        else:
            check(f"rocsolver_{p}getrf_npvt_strided_batched", handle.value, m, n, A.data_ptr(),
                  lda, n * lda, info.data_ptr(), bc)
    return ipiv[:, :k], info


def getrs(handle: Handle, LU, ipiv, B, trans: str = "N", batched: bool = False):
    """Solve op(A) X = B in place for B (bc, nrhs, ldb); LU (bc, n, lda), ipiv (bc, >=n)."""
    bc, n, lda = LU.shape
    _, nrhs, ldb = B.shape
    op = {"N": OP_NONE, "T": OP_TRANSPOSE, "C": OP_CONJ_TRANSPOSE}[trans.upper()]
    p = _prefix(LU)
    sp = ipiv.stride(0)
    if batched:
        pa, pb = pointer_array(LU), pointer_array(B)
        check(f"rocsolver_{p}getrs_batched", handle.value, op, n, nrhs, pa.data_ptr(), lda,
              ipiv.data_ptr(), sp, pb.data_ptr(), ldb, bc)
    else:
        check(f"rocsolver_{p}getrs_strided_batched", handle.value, op, n, nrhs, LU.data_ptr(),
              lda, n * lda, ipiv.data_ptr(), sp, B.data_ptr(), ldb, nrhs * ldb, bc)


def getri(handle: Handle, LU, ipiv=None, C=None, form: str = "strided"):
    """Inverse from getrf factors LU (bc, n, lda): in place, or into C (bc, n,
    ldc) when given; ipiv None selects the no-pivoting (npvt) entry points.
    form: "strided", "batched" (pointer arrays) or "plain" (bc must be 1).
    Returns info (bc,) int32 on the device."""
    import torch
    bc, n, lda = LU.shape
    p = _prefix(LU)
    info = torch.zeros(bc, dtype=torch.int32, device=LU.device)
    npvt = ipiv is None
    name = f"rocsolver_{p}getri" + ("_npvt" if npvt else "") + ("_outofplace" if C is not None else "")
    piv = [] if npvt else [ipiv.data_ptr()]
    if form == "plain":
        assert bc == 1
        args = [LU.data_ptr(), lda] + piv
        if C is not None:
            args += [C.data_ptr(), C.shape[2]]
        check(name, handle.value, n, *args, info.data_ptr())
    elif form == "batched":
        pa = pointer_array(LU)
        args = [pa.data_ptr(), lda] + ([] if npvt else [ipiv.data_ptr(), ipiv.stride(0)])
        if C is not None:
            pc = pointer_array(C)
            args += [pc.data_ptr(), C.shape[2]]
        check(name + "_batched", handle.value, n, *args, info.data_ptr(), bc)
    else:
        args = [LU.data_ptr(), lda, n * lda] + ([] if npvt else [ipiv.data_ptr(), ipiv.stride(0)])
        if C is not None:
            args += [C.data_ptr(), C.shape[2], C.stride(0)]
        check(name + "_strided_batched", handle.value, n, *args, info.data_ptr(), bc)
    return info


FILL_UPPER, FILL_LOWER = 121, 122  # rocblas_fill


def potrf(handle: Handle, A, uplo: str = "L", form: str = "strided", unblocked: bool = False):
    """Cholesky factorisation of the uplo triangle of every A (bc, n, lda) in
    place (potrf, or potf2 with unblocked=True).  form: "strided", "batched" or
    "plain" (bc must be 1).  Returns info (bc,) int32 on the device."""
    import torch
    bc, n, lda = A.shape
    p = _prefix(A)
    u = FILL_UPPER if uplo.upper() == "U" else FILL_LOWER
    info = torch.zeros(bc, dtype=torch.int32, device=A.device)
    name = f"rocsolver_{p}" + ("potf2" if unblocked else "potrf")
    if form == "plain":
        assert bc == 1
        check(name, handle.value, u, n, A.data_ptr(), lda, info.data_ptr())
    elif form == "batched":
        pa = pointer_array(A)
        check(name + "_batched", handle.value, u, n, pa.data_ptr(), lda, info.data_ptr(), bc)
    else:
        check(name + "_strided_batched", handle.value, u, n, A.data_ptr(), lda, n * lda,
              info.data_ptr(), bc)
    return info


def potrs(handle: Handle, F, B, uplo: str = "L", form: str = "strided", posv: bool = False):
    """Solve A X = B in place for B (bc, nrhs, ldb) with the Cholesky factor F
    (potrs), or factor F first (posv: returns info)."""
    import torch
    bc, n, lda = F.shape
    _, nrhs, ldb = B.shape
    p = _prefix(F)
    u = FILL_UPPER if uplo.upper() == "U" else FILL_LOWER
    name = f"rocsolver_{p}" + ("posv" if posv else "potrs")
    info = torch.zeros(bc, dtype=torch.int32, device=F.device) if posv else None
    tail = [info.data_ptr()] if posv else []
    if form == "plain":
        assert bc == 1
        check(name, handle.value, u, n, nrhs, F.data_ptr(), lda, B.data_ptr(), ldb, *tail)
    elif form == "batched":
        pa, pb = pointer_array(F), pointer_array(B)
        check(name + "_batched", handle.value, u, n, nrhs, pa.data_ptr(), lda, pb.data_ptr(), ldb,
              *tail, bc)
    else:
        check(name + "_strided_batched", handle.value, u, n, nrhs, F.data_ptr(), lda, n * lda,
              B.data_ptr(), ldb, nrhs * ldb, *tail, bc)
    return info


def gesv(handle: Handle, A, B, batched: bool = False):
    """Solve A X = B in place (A -> LU, B -> X) -> (ipiv, info)."""
    import torch
    bc, n, lda = A.shape
    _, nrhs, ldb = B.shape
    dev = A.device
    ipiv = torch.zeros((bc, max(n, 1)), dtype=torch.int32, device=dev)
    info = torch.zeros(bc, dtype=torch.int32, device=dev)
    p = _prefix(A)
    if batched:
        pa, pb = pointer_array(A), pointer_array(B)
        check(f"rocsolver_{p}gesv_batched", handle.value, n, nrhs, pa.data_ptr(), lda,
              ipiv.data_ptr(), ipiv.shape[1], pb.data_ptr(), ldb, info.data_ptr(), bc)
    else:
        check(f"rocsolver_{p}gesv_strided_batched", handle.value, n, nrhs, A.data_ptr(), lda,
              n * lda, ipiv.data_ptr(), ipiv.shape[1], B.data_ptr(), ldb, nrhs * ldb,
              info.data_ptr(), bc)
    return ipiv[:, :n], info


FACTOR = {"fp32": 0, "fp16": 1, "bf16": 2}


def dsgesv(handle: Handle, A, B, want_berr: bool = True, factor: str = "fp32"):
    """Mixed-precision solve of every system (FP32 LU + FP64 refinement; with
    factor "fp16"/"bf16" the equilibrated LU uses half-width matrix-core
    operands).

    A (bc, n, lda) and B (bc, nrhs, ldb) are float64 and left unchanged.
    Returns (X (bc, nrhs, n), iter[bc], info[bc], berr[bc] or None)."""
    import torch
    bc, n, lda = A.shape
    _, nrhs, ldb = B.shape
    dev = A.device
    X = torch.zeros((bc, nrhs, n), dtype=torch.float64, device=dev)
    it = torch.zeros(bc, dtype=torch.int32, device=dev)
    info = torch.zeros(bc, dtype=torch.int32, device=dev)
    berr = torch.zeros(bc, dtype=torch.float64, device=dev) if want_berr else None
    if factor == "fp32":
        check("rocsolver_amd_dsgesv_strided_batched", handle.value, n, nrhs, A.data_ptr(), lda,
              n * lda, B.data_ptr(), ldb, nrhs * ldb, X.data_ptr(), n, nrhs * n, it.data_ptr(),
              info.data_ptr(), berr.data_ptr() if want_berr else None, bc)
    else:
        check("rocsolver_amd_dxgesv_strided_batched", handle.value, FACTOR[factor], n, nrhs,
              A.data_ptr(), lda, n * lda, B.data_ptr(), ldb, nrhs * ldb, X.data_ptr(), n,
              nrhs * n, it.data_ptr(), info.data_ptr(),
              berr.data_ptr() if want_berr else None, bc)
    return X, it, info, berr


# ------------------------------------------------- variable-size batches
VB_TUNING_LEN = 13


def vbatched_get_tuning() -> list[int]:
    buf = (ctypes.c_int * VB_TUNING_LEN)()
    check("rocsolver_amd_vbatched_get_tuning", ctypes.cast(buf, ctypes.c_void_p), VB_TUNING_LEN)
    return list(buf)


def vbatched_set_tuning(params) -> None:
    """Install a schedule table (team[6], pack[6], workgroups per CU)."""
    vals = [int(v) for v in params]
    buf = (ctypes.c_int * len(vals))(*vals)
    check("rocsolver_amd_vbatched_set_tuning", ctypes.cast(buf, ctypes.c_void_p), len(vals))


class VBatch:
    """Matrices of different orders packed in one device buffer each for A,
    B and ipiv, with the device size and pointer arrays the vbatched routines
    take.  Matrix b is column-major with lda = ldb = n[b]: A_of(b) is a view
    of shape (n, n) holding A^T (row j = column j), B_of(b) (nrhs, n)."""

    def __init__(self, n, nrhs, dtype, device="cuda"):
        import numpy as np
        import torch
        self.n = np.asarray(n, dtype=np.int64)
        self.nrhs = np.asarray(nrhs, dtype=np.int64)
        self.bc = len(self.n)
        self.dtype, self.device = dtype, device
        esz = torch.empty((), dtype=dtype).element_size()
        na = self.n * self.n
        nb = self.n * self.nrhs
        self.offA = np.concatenate([[0], np.cumsum(na)[:-1]]).astype(np.int64)
        self.offB = np.concatenate([[0], np.cumsum(nb)[:-1]]).astype(np.int64)
        self.offP = np.concatenate([[0], np.cumsum(self.n)[:-1]]).astype(np.int64)
        self.A = torch.zeros(int(na.sum()) + 1, dtype=dtype, device=device)
        self.B = torch.zeros(int(nb.sum()) + 1, dtype=dtype, device=device)
        self.ipiv = torch.zeros(int(self.n.sum()) + 1, dtype=torch.int32, device=device)
        self.info = torch.zeros(self.bc, dtype=torch.int32, device=device)
        i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(device)
        i64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.int64)).to(device)
        self.n_dev = i32(self.n)
        self.nrhs_dev = i32(self.nrhs)
        self.ld_dev = i32(np.maximum(self.n, 1))
        self.pA = i64(self.A.data_ptr() + self.offA * esz)
        self.pB = i64(self.B.data_ptr() + self.offB * esz)
        self.pP = i64(self.ipiv.data_ptr() + self.offP * 4)

    def A_of(self, b: int):
        k, o = int(self.n[b]), int(self.offA[b])
        return self.A[o:o + k * k].view(k, k)

    def B_of(self, b: int):
        k, r, o = int(self.n[b]), int(self.nrhs[b]), int(self.offB[b])
        return self.B[o:o + k * r].view(r, k)

    def ipiv_of(self, b: int):
        k, o = int(self.n[b]), int(self.offP[b])
        return self.ipiv[o:o + k]

    def _p(self) -> str:
        import torch
        return "d" if self.dtype == torch.float64 else "s"

    def getrf(self, handle: "Handle") -> None:
        check(f"rocsolver_amd_{self._p()}getrf_vbatched", handle.value, self.n_dev.data_ptr(),
              self.pA.data_ptr(), self.ld_dev.data_ptr(), self.pP.data_ptr(),
              self.info.data_ptr(), self.bc)

    def getrs(self, handle: "Handle", trans: str = "N") -> None:
        op = {"N": OP_NONE, "T": OP_TRANSPOSE, "C": OP_CONJ_TRANSPOSE}[trans.upper()]
        check(f"rocsolver_amd_{self._p()}getrs_vbatched", handle.value, op, self.n_dev.data_ptr(),
              self.nrhs_dev.data_ptr(), self.pA.data_ptr(), self.ld_dev.data_ptr(),
              self.pP.data_ptr(), self.pB.data_ptr(), self.ld_dev.data_ptr(),
              self.info.data_ptr(), self.bc)

    def gesv(self, handle: "Handle", noalloc: bool = False) -> None:
        """noalloc: the FP64 entry point whose work queue is the library's
        static device array (no allocation; batches up to 65536)."""
        name = f"rocsolver_amd_{self._p()}gesv_vbatched"
        if noalloc:
            if self._p() != "d":
                raise ValueError("the no-allocation gesv is FP64 only")
            name += "_noalloc"
        check(name, handle.value, self.n_dev.data_ptr(),
              self.nrhs_dev.data_ptr(), self.pA.data_ptr(), self.ld_dev.data_ptr(),
              self.pP.data_ptr(), self.pB.data_ptr(), self.ld_dev.data_ptr(),
              self.info.data_ptr(), self.bc)
