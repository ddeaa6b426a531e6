"""GPU checks of the Cholesky family (rocsolver_{s,d}{potf2,potrf,potrs,posv}
in single, batched and strided-batched forms) against the oracle.

Inputs follow the reference tests (testing_potf2_potrf.hpp:116-140,
testing_potrs.hpp:124-135, testing_posv.hpp): the clients' integers with
each diagonal entry a replaced by 400 a^2, and in batch entries bc/4, bc/2,
bc-1 three zeroed diagonal entries (non-positive-definite).  The factor's
order of operations is the oracle's, so factors and info are compared
exactly; the other triangle must come back untouched.  Solves are compared
with the oracle within n*eps (the reference's threshold) and checked
through A X = B.
"""
import numpy as np
import pytest

import oracle
import rocsolver_amd as ra
from rocsolver_amd import workloads as wl

pytestmark = pytest.mark.gpu

EPS = {np.float64: np.finfo(np.float64).eps, np.float32: np.finfo(np.float32).eps}
SIZES = [(1, 1), (10, 10), (20, 30), (32, 32), (50, 50), (64, 64), (65, 70), (100, 100),
         (192, 192), (300, 310), (520, 520)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,lda", SIZES)
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_potrf_strided_equals_oracle(torch_gpu, handle, dtype, n, lda, uplo):
    torch = torch_gpu
    bc = 5
    print(f"potrf {uplo} n={n} lda={lda} {dtype.__name__}", flush=True)
    A = wl.reference_potrf(n, lda, bc, singular=True, dtype=dtype)
    Ao = A.copy()
    info_o = oracle.potf2(Ao, n, uplo)
    Ad = torch.from_numpy(A.copy()).cuda()
    info = ra.potrf(handle, Ad, uplo)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(info.cpu().numpy(), info_o)
    assert (info_o != 0).sum() == 3
    # the whole array: factor bit-identical, other triangle and padding untouched
    np.testing.assert_array_equal(Ad.cpu().numpy(), Ao)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [16, 70, 256])
@pytest.mark.parametrize("form", ["batched", "plain"])
@pytest.mark.parametrize("unblocked", [False, True])
def test_potrf_forms(torch_gpu, handle, dtype, n, form, unblocked):
    torch = torch_gpu
    bc = 1 if form == "plain" else 4
    A = wl.reference_potrf(n, n, bc, singular=form != "plain", dtype=dtype)
    Ao = A.copy()
    info_o = oracle.potf2(Ao, n, "U")
    Ad = torch.from_numpy(A.copy()).cuda()
    info = ra.potrf(handle, Ad, "U", form=form, unblocked=unblocked)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(info.cpu().numpy(), info_o)
    np.testing.assert_array_equal(Ad.cpu().numpy(), Ao)


def sym_from(A, n, uplo):
    M = A[:n, :n].T.astype(np.float64)
    if uplo == "L":
        return np.tril(M) + np.tril(M, -1).T
    return np.triu(M) + np.triu(M, 1).T


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,lda,ldb,nrhs", [(10, 10, 10, 2), (20, 20, 20, 20), (30, 50, 30, 5),
                                            (64, 64, 64, 9), (70, 70, 100, 1), (192, 192, 192, 8),
                                            (600, 600, 600, 3)])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_potrs_and_posv(torch_gpu, handle, dtype, n, lda, ldb, nrhs, uplo):
    torch = torch_gpu
    bc = 4
    print(f"potrs/posv {uplo} n={n} nrhs={nrhs} {dtype.__name__}", flush=True)
    A = wl.reference_potrf(n, lda, bc, singular=True, dtype=dtype)
    B = np.random.default_rng(n + nrhs).uniform(1, 10, (bc, nrhs, ldb)).astype(dtype)
    F = A.copy()
    info_o = oracle.potf2(F, n, uplo)
    # potrs with the (oracle) factor of every entry that is positive definite
    Xo = B.copy()
    for b in range(bc):
        if info_o[b] == 0:
            oracle.potrs(F[b:b + 1], n, Xo[b:b + 1], uplo)
    Fd = torch.from_numpy(F.copy()).cuda()
    Bd = torch.from_numpy(B.copy()).cuda()
    ra.potrs(handle, Fd, Bd, uplo)
    torch.cuda.synchronize()
    X = Bd.cpu().numpy()
    for b in range(bc):
        if info_o[b] != 0:
            continue
        err = np.abs(X[b, :, :n] - Xo[b, :, :n]).max() / np.abs(Xo[b, :, :n]).max()
        assert err <= n * EPS[dtype], (b, err)
        S = sym_from(A[b], n, uplo)
        res = np.abs(S @ X[b, :, :n].T.astype(np.float64) - B[b, :, :n].T).max()
        assert res <= 100 * n * EPS[dtype] * np.abs(S).max() * np.abs(X[b, :, :n]).max()
    np.testing.assert_array_equal(X[:, :, n:], B[:, :, n:])  # ldb padding untouched
    # posv: factor + solve; B unchanged where info != 0
    Ad = torch.from_numpy(A.copy()).cuda()
    Bd = torch.from_numpy(B.copy()).cuda()
    info = ra.potrs(handle, Ad, Bd, uplo, posv=True)
    torch.cuda.synchronize()
    info = info.cpu().numpy()
    np.testing.assert_array_equal(info, info_o)
    Xp = Bd.cpu().numpy()
    np.testing.assert_array_equal(Ad.cpu().numpy(), F)
    for b in range(bc):
        if info[b] != 0:
            np.testing.assert_array_equal(Xp[b], B[b])
        else:
            np.testing.assert_array_equal(Xp[b], X[b])


@pytest.mark.parametrize("form", ["batched", "plain"])
def test_potrs_posv_forms(torch_gpu, handle, form):
    torch = torch_gpu
    n, nrhs = 40, 3
    bc = 1 if form == "plain" else 4
    A = wl.reference_potrf(n, n, bc, singular=False)
    B = np.random.default_rng(3).uniform(1, 10, (bc, nrhs, n))
    F = A.copy()
    oracle.potf2(F, n, "L")
    Xo = B.copy()
    oracle.potrs(F, n, Xo, "L")
    Fd, Bd = torch.from_numpy(F.copy()).cuda(), torch.from_numpy(B.copy()).cuda()
    ra.potrs(handle, Fd, Bd, "L", form=form)
    torch.cuda.synchronize()
    assert np.abs(Bd.cpu().numpy() - Xo).max() <= n * EPS[np.float64] * np.abs(Xo).max()
    Ad, Bd = torch.from_numpy(A.copy()).cuda(), torch.from_numpy(B.copy()).cuda()
    info = ra.potrs(handle, Ad, Bd, "L", form=form, posv=True)
    torch.cuda.synchronize()
    assert (info.cpu().numpy() == 0).all()
    np.testing.assert_array_equal(Ad.cpu().numpy(), F)
    assert np.abs(Bd.cpu().numpy() - Xo).max() <= n * EPS[np.float64] * np.abs(Xo).max()
