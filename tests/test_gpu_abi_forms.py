"""GPU parity, through the C ABI, of the entry points beyond the strided
getrf/getrs/gesv: rocsolver_{s,d}getf2{,_batched,_strided_batched}, the
plain (non-batched) getrf/getrs/gesv, rocsolver_{s,d}getrf_npvt in all three
forms, and the pointer-array rocsolver_{s,d}getrs_batched.

getf2 and getrf share one contract (reference testing_getf2_getrf.hpp:
237-255): ipiv and info exactly equal to the host result, LU within
min(m,n)*eps; the register-resident sizes (m, n <= 64) are bit-identical to
the oracle.  getrf_npvt is compared with the oracle's restatement of the
reference no-pivot kernel (oracle.getf2_npvt), bit for bit at m, n <= 64,
including a zero diagonal pivot above non-zero entries.
"""
import numpy as np
import pytest

import oracle
import rocsolver_amd as ra
from rocsolver_amd import workloads as wl

pytestmark = pytest.mark.gpu
EPS = {np.float64: np.finfo(np.float64).eps, np.float32: np.finfo(np.float32).eps}
PRE = {np.float64: "d", np.float32: "s"}


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def rel_fro(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


@pytest.mark.parametrize("form", ["plain", "batched", "strided"])
@pytest.mark.parametrize("m,n,lda", [(32, 32, 32), (50, 40, 50), (20, 64, 30), (100, 100, 100),
                                     (130, 70, 130)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_getf2_forms(torch_gpu, handle, form, m, n, lda, dtype):
    torch = torch_gpu
    print(f"getf2 {form} m={m} n={n} lda={lda} {dtype.__name__}", flush=True)
    bc = 1 if form == "plain" else 4
    A0 = wl.reference_getrf(m, n, lda, bc, singular=(bc > 1), dtype=dtype)
    Ao = A0.copy()
    ip_o, in_o = oracle.getf2(Ao, m)
    Ad = dev(torch, A0)
    k = min(m, n)
    ip = torch.zeros((bc, k), dtype=torch.int32, device="cuda")
    info = torch.full((bc,), -9, dtype=torch.int32, device="cuda")
    p = PRE[dtype]
    h = handle.value
    if form == "plain":
        ra.check(f"rocsolver_{p}getf2", h, m, n, Ad.data_ptr(), lda, ip.data_ptr(),
                 info.data_ptr())
    elif form == "batched":
        ptrs = ra.pointer_array(Ad)
        ra.check(f"rocsolver_{p}getf2_batched", h, m, n, ptrs.data_ptr(), lda, ip.data_ptr(), k,
                 info.data_ptr(), bc)
    else:
        ra.check(f"rocsolver_{p}getf2_strided_batched", h, m, n, Ad.data_ptr(), lda, n * lda,
                 ip.data_ptr(), k, info.data_ptr(), bc)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(host(info), in_o)
    np.testing.assert_array_equal(host(ip), ip_o)
    Ag = host(Ad)
    if m <= 64 and n <= 64:
        np.testing.assert_array_equal(Ag, Ao)
    for b in range(bc):
        assert rel_fro(Ag[b, :n, :m], Ao[b, :n, :m]) <= min(m, n) * EPS[dtype]
    np.testing.assert_array_equal(Ag[:, :, m:], A0[:, :, m:])


@pytest.mark.parametrize("n,nrhs", [(24, 3), (64, 1), (96, 5), (300, 2)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_plain_getrf_getrs_gesv(torch_gpu, handle, n, nrhs, dtype):
    """rocsolver_{s,d}getrf / getrs / gesv (one matrix, roclapack_getrs.cpp:8-72)."""
    torch = torch_gpu
    print(f"plain getrf/getrs/gesv n={n} nrhs={nrhs} {dtype.__name__}", flush=True)
    A, B = wl.reference_gesv(n, nrhs, n, n, 1, False, dtype)
    p, h = PRE[dtype], handle.value
    # getrf + getrs (N and T)
    Ao = A.copy()
    ip_o, in_o = oracle.getf2(Ao, n)
    Ad = dev(torch, A)
    ip = torch.zeros((1, n), dtype=torch.int32, device="cuda")
    info = torch.full((1,), -9, dtype=torch.int32, device="cuda")
    ra.check(f"rocsolver_{p}getrf", h, n, n, Ad.data_ptr(), n, ip.data_ptr(), info.data_ptr())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(host(ip), ip_o)
    np.testing.assert_array_equal(host(info), in_o)
    if n <= 64:
        np.testing.assert_array_equal(host(Ad), Ao)
    assert rel_fro(host(Ad), Ao) <= n * EPS[dtype]
    for trans, op in (("N", ra.OP_NONE), ("T", ra.OP_TRANSPOSE)):
        Xo = B.copy()
        oracle.getrs(trans, Ao, n, ip_o, Xo)
        Bd = dev(torch, B)
        ra.check(f"rocsolver_{p}getrs", h, op, n, nrhs, Ad.data_ptr(), n, ip.data_ptr(),
                 Bd.data_ptr(), n)
        torch.cuda.synchronize()
        Xg = host(Bd)
        if n <= 64:
            np.testing.assert_array_equal(Xg, Xo)
        assert np.abs(Xg - Xo).max() / np.abs(Xo).max() <= n * EPS[dtype]
    # gesv
    Ao2, Xo = A.copy(), B.copy()
    ip_o2, in_o2 = oracle.gesv(Ao2, n, Xo)
    Ad, Bd = dev(torch, A), dev(torch, B)
    info.fill_(-9)
    ra.check(f"rocsolver_{p}gesv", h, n, nrhs, Ad.data_ptr(), n, ip.data_ptr(), Bd.data_ptr(), n,
             info.data_ptr())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(host(info), in_o2)
    np.testing.assert_array_equal(host(ip), ip_o2)
    if n <= 64:
        np.testing.assert_array_equal(host(Bd), Xo)
    assert np.abs(host(Bd) - Xo).max() / np.abs(Xo).max() <= n * EPS[dtype]


def npvt_inputs(m, n, lda, bc, dtype, seed):
    """Diagonally dominant matrices (as the reference npvt test builds them,
    testing_getf2_getrf_npvt.hpp:113-149), one with whole zero columns, and one
    whose first diagonal pivot is zero above non-zero entries."""
    g = np.random.default_rng(seed)
    A = g.integers(1, 11, size=(bc, n, lda)).astype(dtype)
    for b in range(bc):
        for j in range(n):
            A[b, j, :m] -= 4
            if j < m:
                A[b, j, j] += 404
    if bc > 1:
        for j in (n // 4, n // 2, n - 1):
            A[1, j, :m] = 0  # singular: info = first zero column + 1
    if bc > 2:
        A[2, 0, 0] = 0  # zero pivot with non-zeros below it
    return A


@pytest.mark.parametrize("form", ["plain", "batched", "strided"])
@pytest.mark.parametrize("m,n,lda", [(1, 1, 1), (8, 8, 8), (32, 32, 40), (50, 30, 50), (30, 50, 30),
                                     (64, 64, 64), (100, 100, 100), (200, 150, 210)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_getrf_npvt_forms(torch_gpu, handle, form, m, n, lda, dtype):
    torch = torch_gpu
    print(f"getrf_npvt {form} m={m} n={n} lda={lda} {dtype.__name__}", flush=True)
    bc = 1 if form == "plain" else 4
    A0 = npvt_inputs(m, n, lda, bc, dtype, seed=m * 31 + n)
    Ao = A0.copy()
    in_o = oracle.getf2_npvt(Ao, m)
    Ad = dev(torch, A0)
    info = torch.full((bc,), -9, dtype=torch.int32, device="cuda")
    p, h = PRE[dtype], handle.value
    if form == "plain":
        ra.check(f"rocsolver_{p}getrf_npvt", h, m, n, Ad.data_ptr(), lda, info.data_ptr())
    elif form == "batched":
        ptrs = ra.pointer_array(Ad)
        ra.check(f"rocsolver_{p}getrf_npvt_batched", h, m, n, ptrs.data_ptr(), lda,
                 info.data_ptr(), bc)
    else:
        ra.check(f"rocsolver_{p}getrf_npvt_strided_batched", h, m, n, Ad.data_ptr(), lda,
                 n * lda, info.data_ptr(), bc)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(host(info), in_o)
    Ag = host(Ad)
    if m <= 64 and n <= 64:
        np.testing.assert_array_equal(Ag, Ao)
    else:
        for b in range(bc):
            assert rel_fro(Ag[b, :n, :m], Ao[b, :n, :m]) <= 10 * min(m, n) * EPS[dtype]
    np.testing.assert_array_equal(Ag[:, :, m:], A0[:, :, m:])


@pytest.mark.parametrize("n,nrhs,lda,ldb", [(20, 3, 20, 25), (64, 1, 64, 64), (100, 4, 110, 100),
                                            (600, 2, 600, 600)])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_getrs_pointer_array(torch_gpu, handle, n, nrhs, lda, ldb, trans, dtype):
    """rocsolver_{s,d}getrs_batched: A and B through device pointer arrays."""
    torch = torch_gpu
    print(f"getrs_batched n={n} nrhs={nrhs} {trans} {dtype.__name__}", flush=True)
    A, B = wl.reference_getrs(n, nrhs, lda, ldb, 3, dtype)
    ipiv, _ = oracle.getf2(A, n)
    Xo = B.copy()
    oracle.getrs(trans, A, n, ipiv, Xo)
    Ad, Bd, ipd = dev(torch, A), dev(torch, B), dev(torch, ipiv)
    ra.getrs(handle, Ad, ipd, Bd, trans=trans, batched=True)
    torch.cuda.synchronize()
    Xg = host(Bd)
    if n <= 64:
        np.testing.assert_array_equal(Xg, Xo)
    for b in range(3):
        d = np.abs(Xg[b, :, :n] - Xo[b, :, :n]).max() / np.abs(Xo[b, :, :n]).max()
        assert d <= n * EPS[dtype]
    np.testing.assert_array_equal(Xg[:, :, n:], B[:, :, n:])
