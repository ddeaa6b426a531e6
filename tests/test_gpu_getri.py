"""GPU checks of the inverse from the LU factors (rocsolver_{s,d}getri*,
getri_outofplace*, and their npvt forms) against the oracle's getri.

Inputs follow the reference getri tests (testing_getri.hpp:117-157): the
clients' integers reshaped to a well-conditioned matrix, rows reversed, the
host LU, then zeros planted on the diagonal of U in batch entries bc/4, bc/2
and bc-1.  Checks: info exactly (the first zero diagonal entry of U), and
for info = 0 the inverse within the reference threshold n*eps in relative
Frobenius norm (testing_getri.hpp:224-233,465), plus A X = I.
"""
import numpy as np
import pytest

import oracle
import rocsolver_amd as ra
from rocsolver_amd import workloads as wl

pytestmark = pytest.mark.gpu

EPS = {np.float64: np.finfo(np.float64).eps, np.float32: np.finfo(np.float32).eps}


def rel_fro(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def factors(n, lda, bc, dtype, pivot=True, plant=True):
    A = wl.reference_getri(n, lda, bc, dtype=dtype)
    LU = A.copy()
    if pivot:
        ipiv, info = oracle.getf2(LU, n)
    else:
        for b in range(bc):  # diagonally dominant: no interchange needed
            LU[b, :n, :n] += np.eye(n, dtype=dtype) * 4 * n
        A = LU.copy()
        info = oracle.getf2_npvt(LU, n)
        ipiv = None
    assert (info == 0).all()
    if plant:
        wl.reference_getri_singular(LU, n)
    return A, LU, ipiv


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,lda", [(1, 1), (5, 5), (32, 32), (50, 50), (64, 64), (65, 70),
                                   (100, 150), (192, 192), (300, 300), (520, 520)])
@pytest.mark.parametrize("place", ["inplace", "outofplace"])
def test_getri_strided(torch_gpu, handle, dtype, n, lda, place):
    torch = torch_gpu
    bc = 5
    print(f"getri {place} n={n} lda={lda} {dtype.__name__}", flush=True)
    A, LU, ipiv = factors(n, lda, bc, dtype)
    Xo = LU.copy()
    info_o = oracle.getri(Xo, n, ipiv)
    LUd = torch.from_numpy(LU.copy()).cuda()
    ipd = torch.from_numpy(np.ascontiguousarray(ipiv)).cuda()
    if place == "inplace":
        info = ra.getri(handle, LUd, ipd)
        X = LUd
    else:
        C = torch.full((bc, n, n + 3), 7.0, dtype=LUd.dtype, device="cuda")
        info = ra.getri(handle, LUd, ipd, C=C)
        np.testing.assert_array_equal(LUd.cpu().numpy(), LU)  # the factors are read only
        X = C
    torch.cuda.synchronize()
    info = info.cpu().numpy()
    X = X.cpu().numpy()
    np.testing.assert_array_equal(info, info_o)
    assert (info != 0).sum() == 3
    for b in range(bc):
        if info[b] == 0:
            assert rel_fro(X[b, :n, :n], Xo[b, :n, :n]) <= n * EPS[dtype], b
            M = A[b, :n, :n].T.astype(np.float64)
            res = np.abs(M @ X[b, :n, :n].T.astype(np.float64) - np.eye(n)).max()
            assert res <= 100 * n * EPS[dtype]
    if place == "inplace" and lda > n:
        np.testing.assert_array_equal(X[:, :n, n:], LU[:, :n, n:])  # padding untouched
    if place == "outofplace":
        np.testing.assert_array_equal(X[:, :, n:], 7.0)  # ldc padding untouched


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [8, 40, 96, 200])
@pytest.mark.parametrize("form", ["batched", "plain"])
@pytest.mark.parametrize("place", ["inplace", "outofplace"])
@pytest.mark.parametrize("pivot", [True, False])
def test_getri_forms(torch_gpu, handle, dtype, n, form, place, pivot):
    """Pointer-array and single-matrix entry points, with and without pivots."""
    torch = torch_gpu
    bc = 1 if form == "plain" else 4
    print(f"getri {form} {place} pivot={pivot} n={n} {dtype.__name__}", flush=True)
    A, LU, ipiv = factors(n, n, bc, dtype, pivot, plant=form != "plain")
    Xo = LU.copy()
    info_o = oracle.getri(Xo, n, ipiv)
    LUd = torch.from_numpy(LU.copy()).cuda()
    ipd = None if ipiv is None else torch.from_numpy(np.ascontiguousarray(ipiv)).cuda()
    C = torch.zeros((bc, n, n), dtype=LUd.dtype, device="cuda") if place == "outofplace" else None
    info = ra.getri(handle, LUd, ipd, C=C, form=form)
    torch.cuda.synchronize()
    X = (LUd if C is None else C).cpu().numpy()
    info = info.cpu().numpy()
    np.testing.assert_array_equal(info, info_o)
    for b in range(bc):
        if info[b] == 0:
            assert rel_fro(X[b], Xo[b]) <= n * EPS[dtype]


def test_getri_quick_returns_and_workspace_query(torch_gpu, handle):
    torch = torch_gpu
    info = torch.full((3,), 9, dtype=torch.int32, device="cuda")
    A = torch.zeros((3, 1, 1), dtype=torch.float64, device="cuda")
    # n = 0: info reset, nothing else
    ra.check("rocsolver_dgetri_strided_batched", handle.value, 0, A.data_ptr(), 1, 1, None, 1,
             info.data_ptr(), 3)
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [0, 0, 0]
    # size query (pointers unchecked): the in-place form asks for n*n*bc
    # elements, the out-of-place one for none
    import ctypes
    rb = ra.rocblas()
    h = handle.value
    size = ctypes.c_size_t()
    assert rb.rocblas_start_device_memory_size_query(h) == 0
    st = ra.call("rocsolver_dgetri_strided_batched", h, 40, None, 40, 1600, None, 40, None, 7)
    assert st in (9, 10)
    assert rb.rocblas_stop_device_memory_size_query(h, ctypes.byref(size)) == 0
    assert size.value >= 40 * 40 * 7 * 8
    assert rb.rocblas_start_device_memory_size_query(h) == 0
    st = ra.call("rocsolver_dgetri_outofplace_strided_batched", h, 40, None, 40, 1600, None, 40,
                 None, 40, 1600, None, 7)
    assert st in (9, 10)
    assert rb.rocblas_stop_device_memory_size_query(h, ctypes.byref(size)) == 0
    assert size.value == 0
