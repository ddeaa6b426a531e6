"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares (dlopen RTLD_NOW must resolve everything, as
clients/extras/test_dlopen_main.cpp:11-14 checks for the reference), and the
argument checks that return before the handle is first used give the
reference's status codes in the reference's order."""
import ctypes
import os
import re

import pytest

import rocsolver_amd as ra

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "rocsolver_amd.h")
DUMMY = 0x1000  # non-null handle value; these checks return before it is dereferenced


def header_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\brocblas_status\s+(rocsolver_\w+)\s*\(", text)))


def test_library_loads_with_rtld_now():
    ctypes.CDLL(ra.LIB_PATH, mode=os.RTLD_NOW)


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(ra.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 34
    for s in syms:
        assert hasattr(L, s), s
    assert set(ra.exported_symbols()) <= set(syms)


def test_version_string():
    v = ra.version()
    assert v.startswith("3.23.0")
    n = ctypes.c_size_t()
    assert ra.lib().rocsolver_get_version_string_size(ctypes.byref(n)) == 0
    assert n.value == len(v) + 1
    assert ra.lib().rocsolver_get_version_string(None, 10) == 3
    buf = ctypes.create_string_buffer(2)
    assert ra.lib().rocsolver_get_version_string(buf, 2) == 4


@pytest.mark.parametrize("p", ["s", "d"])
def test_null_handle(p):
    assert ra.call(f"rocsolver_{p}getrf_strided_batched", None, 4, 4, DUMMY, 4, 16, DUMMY, 4,
                   DUMMY, 1) == 1
    assert ra.call(f"rocsolver_{p}getrs_strided_batched", None, ra.OP_NONE, 4, 1, DUMMY, 4, 16,
                   DUMMY, 4, DUMMY, 4, 4, 1) == 1
    assert ra.call(f"rocsolver_{p}gesv_strided_batched", None, 4, 1, DUMMY, 4, 16, DUMMY, 4,
                   DUMMY, 4, 4, DUMMY, 1) == 1
    assert ra.call(f"rocsolver_{p}getrf_batched", None, 4, 4, DUMMY, 4, DUMMY, 4, DUMMY, 1) == 1
    assert ra.call(f"rocsolver_{p}gesv_batched", None, 4, 1, DUMMY, 4, DUMMY, 4, DUMMY, 4,
                   DUMMY, 1) == 1
    assert ra.call(f"rocsolver_{p}laswp", None, 4, DUMMY, 4, 1, 4, DUMMY, 1) == 1


@pytest.mark.parametrize("p", ["s", "d"])
@pytest.mark.parametrize("m,n,lda,bc", [(-1, 4, 4, 1), (4, -1, 4, 1), (20, 4, 5, 1),
                                        (4, 4, 4, -1)])
def test_getrf_invalid_size(p, m, n, lda, bc):
    # roclapack_getf2.hpp:502-503: m<0 || n<0 || lda<m || bc<0
    for name in ("getrf_strided_batched", "getf2_strided_batched"):
        assert ra.call(f"rocsolver_{p}{name}", DUMMY, m, n, DUMMY, lda, 16, DUMMY, 4, DUMMY,
                       bc) == 4
    assert ra.call(f"rocsolver_{p}getrf_npvt_strided_batched", DUMMY, m, n, DUMMY, lda, 16,
                   DUMMY, bc) == 4


@pytest.mark.parametrize("p", ["s", "d"])
def test_getrs_value_before_size(p):
    # roclapack_getrs.hpp:32-38: invalid trans is reported before invalid sizes
    assert ra.call(f"rocsolver_{p}getrs_strided_batched", DUMMY, 999, -1, 1, DUMMY, 4, 16, DUMMY,
                   4, DUMMY, 4, 4, 1) == 11
    for n, nrhs, lda, ldb, bc in [(-1, 1, 4, 4, 1), (4, -1, 4, 4, 1), (10, 1, 2, 10, 1),
                                  (10, 1, 10, 2, 1), (4, 1, 4, 4, -1)]:
        assert ra.call(f"rocsolver_{p}getrs_strided_batched", DUMMY, ra.OP_NONE, n, nrhs, DUMMY,
                       lda, 16, DUMMY, 4, DUMMY, ldb, 4, bc) == 4


@pytest.mark.parametrize("p", ["s", "d"])
def test_gesv_and_laswp_invalid_size(p):
    for n, nrhs, lda, ldb, bc in [(-1, 1, 4, 4, 1), (4, -1, 4, 4, 1), (10, 1, 2, 10, 1),
                                  (10, 1, 10, 2, 1), (4, 1, 4, 4, -1)]:
        assert ra.call(f"rocsolver_{p}gesv_strided_batched", DUMMY, n, nrhs, DUMMY, lda, 16,
                       DUMMY, 4, DUMMY, ldb, 4, DUMMY, bc) == 4
    # rocauxiliary_laswp.hpp:82-83: n<0 || lda<1 || !incx || k1<1 || k2<1 || k2<k1
    for n, lda, k1, k2, incx in [(-1, 4, 1, 2, 1), (4, 0, 1, 2, 1), (4, 4, 1, 2, 0),
                                 (4, 4, 0, 2, 1), (4, 4, 1, 0, 1), (4, 4, 3, 2, 1)]:
        assert ra.call(f"rocsolver_{p}laswp", DUMMY, n, DUMMY, lda, k1, k2, DUMMY, incx) == 4


@pytest.mark.parametrize("p", ["s", "d"])
def test_vbatched_argument_checks(p):
    # null handle first, then trans (getrs), then batch_count < 0
    assert ra.call(f"rocsolver_amd_{p}getrf_vbatched", None, DUMMY, DUMMY, DUMMY, DUMMY, DUMMY,
                   1) == 1
    assert ra.call(f"rocsolver_amd_{p}gesv_vbatched", None, DUMMY, DUMMY, DUMMY, DUMMY, DUMMY,
                   DUMMY, DUMMY, DUMMY, 1) == 1
    # getrs: an invalid trans is reported before the sizes; a valid one passes
    # on to the size check (rocblas_operation_none = 111)
    assert ra.call(f"rocsolver_amd_{p}getrs_vbatched", DUMMY, 999, DUMMY, DUMMY, DUMMY, DUMMY,
                   DUMMY, DUMMY, DUMMY, None, -1) == 11
    assert ra.call(f"rocsolver_amd_{p}getrs_vbatched", DUMMY, 111, DUMMY, DUMMY, DUMMY, DUMMY,
                   DUMMY, DUMMY, DUMMY, None, -1) == 4
    assert ra.call(f"rocsolver_amd_{p}getrf_vbatched", DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, DUMMY,
                   -1) == 4
    assert ra.call(f"rocsolver_amd_{p}gesv_vbatched", DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, DUMMY,
                   DUMMY, DUMMY, DUMMY, -1) == 4


def test_vbatched_tuning_table():
    default = ra.vbatched_get_tuning()
    assert len(default) == ra.VB_TUNING_LEN
    try:
        new = [64, 32, 64, 64, 64, 64, 1, 3, 8, 64, 2, 1, 2]
        ra.vbatched_set_tuning(new)
        assert ra.vbatched_get_tuning() == new
        bad = [
            [8, 16, 32, 32, 64, 64, 2, 2, 2, 2, 2, 2, 1],   # team 8 is not a team size
            [16, 16, 16, 32, 64, 64, 2, 2, 2, 2, 2, 2, 1],  # 16 lanes cannot hold order 24
            [16, 16, 32, 32, 64, 64, 0, 2, 2, 2, 2, 2, 1],  # pack 0
            [16, 16, 32, 32, 64, 64, 2, 2, 2, 2, 2, 65, 1],  # pack 65
            [16, 16, 32, 32, 64, 64, 2, 2, 2, 2, 2, 2, 5],  # 5 workgroups per CU
        ]
        for t in bad:
            with pytest.raises(ra.RocsolverError):
                ra.vbatched_set_tuning(t)
            assert ra.vbatched_get_tuning() == new
        with pytest.raises(ra.RocsolverError):
            ra.vbatched_set_tuning(new[:-1])
    finally:
        ra.vbatched_set_tuning(default)


def test_dxgesv_invalid_factor():
    # the factor is checked right after the handle
    assert ra.call("rocsolver_amd_dxgesv_strided_batched", None, 1, 4, 1, DUMMY, 4, 16, DUMMY,
                   4, 4, DUMMY, 4, 4, DUMMY, DUMMY, None, 1) == 1
    for f in (-1, 3):
        assert ra.call("rocsolver_amd_dxgesv_strided_batched", DUMMY, f, 4, 1, DUMMY, 4, 16,
                       DUMMY, 4, 4, DUMMY, 4, 4, DUMMY, DUMMY, None, 1) == 11


@pytest.mark.parametrize("p", ["s", "d"])
def test_getri_argument_checks(p):
    """getri argCheck (roclapack_getri.hpp:216-242, roclapack_getri_outofplace.hpp:
    40-67): null handle first, then n < 0 || lda < n (|| ldc < n) || bc < 0."""
    f = f"rocsolver_{p}getri_strided_batched"
    assert ra.call(f, None, 4, DUMMY, 4, 16, DUMMY, 4, DUMMY, 1) == 1
    for n, lda, bc in [(-1, 4, 1), (4, 3, 1), (4, 4, -1)]:
        assert ra.call(f, DUMMY, n, DUMMY, lda, 16, DUMMY, 4, DUMMY, bc) == 4
        assert ra.call(f"rocsolver_{p}getri_npvt_strided_batched", DUMMY, n, DUMMY, lda, 16,
                       DUMMY, bc) == 4
    g = f"rocsolver_{p}getri_outofplace_strided_batched"
    assert ra.call(g, DUMMY, 4, DUMMY, 4, 16, DUMMY, 4, DUMMY, 3, 16, DUMMY, 1) == 4  # ldc < n
    assert ra.call(f"rocsolver_{p}getri_npvt_outofplace_batched", DUMMY, 4, DUMMY, 4, DUMMY, 2,
                   DUMMY, 1) == 4


@pytest.mark.parametrize("p", ["s", "d"])
def test_cholesky_argument_checks(p):
    """potf2/potrf (roclapack_potf2.hpp:95-121), potrs (roclapack_potrs.hpp:20-46),
    posv (roclapack_posv.hpp:20-48): null handle, then uplo value, then sizes."""
    L = 122  # rocblas_fill_lower
    for name in ("potf2_strided_batched", "potrf_strided_batched"):
        f = f"rocsolver_{p}{name}"
        assert ra.call(f, None, L, 4, DUMMY, 4, 16, DUMMY, 1) == 1
        assert ra.call(f, DUMMY, 123, -1, DUMMY, 4, 16, DUMMY, 1) == 11  # value before size
        for n, lda, bc in [(-1, 4, 1), (4, 3, 1), (4, 4, -1)]:
            assert ra.call(f, DUMMY, L, n, DUMMY, lda, 16, DUMMY, bc) == 4
    for name, extra in (("potrs_strided_batched", []), ("posv_strided_batched", [DUMMY])):
        f = f"rocsolver_{p}{name}"
        assert ra.call(f, DUMMY, 0, 4, 1, DUMMY, 4, 16, DUMMY, 4, 4, *extra, 1) == 11
        for n, nrhs, lda, ldb, bc in [(-1, 1, 4, 4, 1), (4, -1, 4, 4, 1), (10, 1, 2, 10, 1),
                                      (10, 1, 10, 2, 1), (4, 1, 4, 4, -1)]:
            assert ra.call(f, DUMMY, L, n, nrhs, DUMMY, lda, 16, DUMMY, ldb, 4, *extra, bc) == 4
