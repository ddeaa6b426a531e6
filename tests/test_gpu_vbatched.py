"""GPU checks of the variable-size batched routines
(rocsolver_amd_{s,d}{getrf,getrs,gesv}_vbatched).

Every matrix runs through the device code of the fixed-size path for its
order.  Orders up to 64 (the register kernels) and every solve with given
factors must equal, bit for bit, what the fixed-size strided routine gives
that matrix alone.  For larger orders the factorisation differs in one
place: the fixed-size kernel forms its block solves U = L_tt^-1 C as
inv(L_tt) C on the matrix cores (the inverses kept in workspace), the
persistent kernel, which has no per-matrix workspace, by substitution.
There the pivots and info must still be equal, the factors agree to a few
ulps of the data, and both solutions meet the reference's backward-error
threshold.  The fixed-size routines are themselves pinned against the
oracle (test_gpu_lu.py); the oracle checks here cover the pivots (exact)
and the solution (tolerance) directly as well.
"""
import numpy as np
import pytest

import oracle
import rocsolver_amd as ra

pytestmark = pytest.mark.gpu

# every bucket edge and the large-path boundaries
ORDERS = [1, 2, 5, 8, 9, 13, 16, 17, 20, 24, 25, 31, 32, 33, 40, 48, 49, 55, 64, 65, 70, 100,
          127, 128, 129, 200, 255, 256, 257, 300, 384, 511, 512]


def make(seed, orders, dtype):
    g = np.random.default_rng(seed)
    n = np.array(orders, dtype=np.int64)
    nrhs = g.integers(1, 7, size=len(n))  # up to 6: more than one 4-column pass
    As = [g.uniform(-1, 1, (k, k)) + k * np.eye(k) * (g.uniform() < 0.7) for k in n]
    Bs = [g.uniform(-1, 1, (r, k)) for k, r in zip(n, nrhs)]
    return n, nrhs, [a.astype(dtype) for a in As], [b.astype(dtype) for b in Bs]


def load(torch, n, nrhs, As, Bs, dtype):
    vb = ra.VBatch(n, nrhs, dtype=getattr(torch, dtype))
    for b in range(len(n)):
        vb.A_of(b).copy_(torch.from_numpy(As[b]))
        vb.B_of(b).copy_(torch.from_numpy(Bs[b]))
    return vb


SMALL_MAX = 64  # register-kernel orders: bit-identical to the fixed-size path


def assert_same_factorisation(k, LUv, LUf, ipv, ipf, a, dtype, what):
    """Factors of one matrix a from the persistent kernel (LUv, ipv) against
    the fixed-size routine's (LUf, ipf): bit for bit for the register-kernel
    orders; else, where the oracle finds every pivot unambiguous (a rounding
    difference cannot flip it), equal pivots and factors within 64 k ulps of
    max |LU| (elsewhere the backward errors of the solutions decide)."""
    if k <= SMALL_MAX:
        np.testing.assert_array_equal(ipv, ipf, err_msg=what)
        np.testing.assert_array_equal(LUv, LUf, err_msg=what)
        return
    eps = np.finfo(dtype).eps
    _, _, gap = oracle.getf2(a[None].copy(), k, with_gap=True)
    if (gap > 1e3 * k * eps).all():
        np.testing.assert_array_equal(ipv, ipf, err_msg=what)
        err = np.abs(LUv.astype(np.float64) - LUf).max()
        assert err <= 64 * k * eps * max(np.abs(LUf).max(), 1.0), (what, err)


def assert_same_solution(k, Xv, Xf, a, bm, dtype, what, trans="N"):
    """Solutions of one system: bit for bit for the register-kernel orders,
    else both within the reference's backward-error threshold."""
    if k <= SMALL_MAX:
        np.testing.assert_array_equal(Xv, Xf, err_msg=what)
        return
    for X in (Xv, Xf):
        be = oracle.solve_backward_error(a[None].astype(np.float64), k,
                                         X[None].astype(np.float64),
                                         bm[None].astype(np.float64), trans).max()
        assert be <= 10 * k * np.finfo(dtype).eps, (what, be)


def fixed_gesv(torch, handle, a, bm):
    A = torch.from_numpy(a[None].copy()).cuda()
    B = torch.from_numpy(bm[None].copy()).cuda()
    ipiv, info = ra.gesv(handle, A, B)
    torch.cuda.synchronize()
    return A[0].cpu().numpy(), B[0].cpu().numpy(), ipiv[0].cpu().numpy(), int(info[0])


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_gesv_vbatched_equals_fixed_size(torch_gpu, handle, dtype):
    torch = torch_gpu
    order = list(np.random.default_rng(3).permutation(ORDERS * 2))
    n, nrhs, As, Bs = make(11, order, dtype)
    print(f"gesv_vbatched {dtype} bc={len(n)}", flush=True)
    vb = load(torch, n, nrhs, As, Bs, dtype)
    vb.gesv(handle)
    torch.cuda.synchronize()
    info = vb.info.cpu().numpy()
    for b in range(len(n)):
        LU, X, ipiv, inf = fixed_gesv(torch, handle, As[b], Bs[b])
        k = int(n[b])
        assert_same_factorisation(k, vb.A_of(b).cpu().numpy(), LU, vb.ipiv_of(b).cpu().numpy(),
                                  ipiv, As[b], dtype, f"n={k}")
        assert_same_solution(k, vb.B_of(b).cpu().numpy(), X, As[b], Bs[b], dtype, f"n={k}")
        assert info[b] == inf
        # and against the oracle directly
        k = int(n[b])
        Ao = As[b][None].copy()
        ipo, info_o, gap = oracle.getf2(Ao, k, with_gap=True)
        if (gap > 1e3 * k * np.finfo(As[b].dtype).eps).all():  # unambiguous pivots
            np.testing.assert_array_equal(ipiv, ipo[0])
        assert info[b] == info_o[0]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_getrs_vbatched_equals_fixed_size(torch_gpu, handle, dtype, trans):
    """getrs_vbatched on factors made by the fixed-size getrf of each matrix:
    every solution equals the fixed-size getrs of that matrix bit for bit."""
    torch = torch_gpu
    n, nrhs, As, Bs = make(14, ORDERS, dtype)
    print(f"getrs_vbatched {dtype} {trans} bc={len(n)}", flush=True)
    vb = load(torch, n, nrhs, As, Bs, dtype)
    fixed = []
    for b in range(len(n)):
        A = torch.from_numpy(As[b][None].copy()).cuda()
        ipiv, _ = ra.getrf(handle, A)
        vb.A_of(b).copy_(A[0])
        vb.ipiv_of(b).copy_(ipiv[0])
        B = torch.from_numpy(Bs[b][None].copy()).cuda()
        ra.getrs(handle, A, ipiv, B, trans)
        fixed.append(B)
    vb.getrs(handle, trans)
    torch.cuda.synchronize()
    assert np.all(vb.info.cpu().numpy() == 0)  # getrs writes 0 for every processed matrix
    for b in range(len(n)):
        np.testing.assert_array_equal(vb.B_of(b).cpu().numpy(), fixed[b][0].cpu().numpy(),
                                      err_msg=f"n={n[b]}")


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_getrf_getrs_vbatched_equal_fixed_size(torch_gpu, handle, dtype, trans):
    torch = torch_gpu
    n, nrhs, As, Bs = make(12, ORDERS, dtype)
    print(f"getrf/getrs_vbatched {dtype} {trans} bc={len(n)}", flush=True)
    vb = load(torch, n, nrhs, As, Bs, dtype)
    vb.getrf(handle)
    vb.getrs(handle, trans)
    torch.cuda.synchronize()
    info = vb.info.cpu().numpy()
    assert np.all(info == 0)  # getrs writes 0 for every processed matrix
    for b in range(len(n)):
        A = torch.from_numpy(As[b][None].copy()).cuda()
        B = torch.from_numpy(Bs[b][None].copy()).cuda()
        ipiv, _ = ra.getrf(handle, A)
        ra.getrs(handle, A, ipiv, B, trans)
        torch.cuda.synchronize()
        k = int(n[b])
        assert_same_factorisation(k, vb.A_of(b).cpu().numpy(), A[0].cpu().numpy(),
                                  vb.ipiv_of(b).cpu().numpy(), ipiv[0].cpu().numpy(), As[b],
                                  dtype, f"n={k}")
        assert_same_solution(k, vb.B_of(b).cpu().numpy(), B[0].cpu().numpy(), As[b], Bs[b],
                             dtype, f"n={k} {trans}", trans)


def test_gesv_vbatched_solution_accuracy(torch_gpu, handle):
    torch = torch_gpu
    n, nrhs, As, Bs = make(13, ORDERS, "float64")
    vb = load(torch, n, nrhs, As, Bs, "float64")
    vb.gesv(handle)
    torch.cuda.synchronize()
    for b in range(len(n)):
        k = int(n[b])
        X = vb.B_of(b).cpu().numpy()
        Xo = Bs[b][None].copy()
        _, info_o = oracle.gesv(As[b][None].copy(), k, Xo)
        if info_o[0] == 0:
            err = np.abs(X - Xo[0]).max() / max(np.abs(Xo[0]).max(), 1e-300)
            assert err <= 1e3 * k * np.finfo(np.float64).eps, (k, err)


def test_vbatched_invalid_and_empty_entries(torch_gpu, handle):
    torch = torch_gpu
    n = np.array([16, 0, 100, 40, 300, 8])
    nrhs = np.array([2, 1, 3, 0, 1, 1])
    vb = ra.VBatch(n, nrhs, dtype=torch.float64)
    g = np.random.default_rng(5)
    for b in range(len(n)):
        k = int(n[b])
        vb.A_of(b).copy_(torch.from_numpy(g.uniform(-1, 1, (k, k)) + k * np.eye(k)))
        vb.B_of(b).copy_(torch.from_numpy(g.uniform(-1, 1, (int(nrhs[b]), k))))
    A0, B0 = vb.A.clone(), vb.B.clone()
    # invalid sizes: order 600 (> 512) and lda < n, set through the device arrays
    vb.n_dev[2] = 600
    vb.ld_dev[4] = 10
    vb.gesv(handle)
    torch.cuda.synchronize()
    info = vb.info.cpu().numpy()
    assert info[2] == -1 and info[4] == -1
    assert info[1] == 0 and info[3] == 0  # n = 0 and nrhs = 0: nothing to do
    # skipped matrices are untouched
    for b in (2, 3, 4):
        o, k = int(vb.offA[b]), int(n[b])
        np.testing.assert_array_equal(vb.A[o:o + k * k].cpu().numpy(),
                                      A0[o:o + k * k].cpu().numpy())
    # the others are solved
    for b in (0, 5):
        assert info[b] == 0
        k, r = int(n[b]), int(nrhs[b])
        a = A0[int(vb.offA[b]):int(vb.offA[b]) + k * k].view(k, k).cpu().numpy()
        rhs = B0[int(vb.offB[b]):int(vb.offB[b]) + k * r].view(r, k).cpu().numpy()
        X = vb.B_of(b).cpu().numpy()
        assert oracle.solve_backward_error(a[None], k, X[None], rhs[None]).max() \
            <= 10 * k * np.finfo(np.float64).eps


def test_vbatched_singular_leaves_b(torch_gpu, handle):
    torch = torch_gpu
    n = np.array([20, 150, 20, 150])
    nrhs = np.array([2, 2, 2, 2])
    g = np.random.default_rng(8)
    As = [g.uniform(-1, 1, (k, k)) + k * np.eye(k) for k in n]
    As[1][7, :] = 0.0  # zero column 8 (row 7 of A^T)
    As[2][3, :] = 0.0
    Bs = [g.uniform(-1, 1, (r, k)) for k, r in zip(n, nrhs)]
    vb = load(torch, n, nrhs, As, Bs, "float64")
    vb.gesv(handle)
    torch.cuda.synchronize()
    info = vb.info.cpu().numpy()
    assert info[1] == 8 and info[2] == 4
    np.testing.assert_array_equal(vb.B_of(1).cpu().numpy(), Bs[1])
    np.testing.assert_array_equal(vb.B_of(2).cpu().numpy(), Bs[2])
    assert info[0] == 0 and info[3] == 0


def test_vbatched_tuning_does_not_change_results(torch_gpu, handle):
    torch = torch_gpu
    n, nrhs, As, Bs = make(14, ORDERS, "float64")
    default = ra.vbatched_get_tuning()
    out = []
    try:
        for table in (default, [64, 64, 64, 64, 64, 64, 1, 1, 1, 1, 1, 1, 2],
                      [32, 16, 32, 64, 64, 64, 8, 3, 5, 1, 2, 7, 1]):
            ra.vbatched_set_tuning(table)
            vb = load(torch, n, nrhs, As, Bs, "float64")
            vb.gesv(handle)
            torch.cuda.synchronize()
            out.append((vb.A.cpu().numpy(), vb.B.cpu().numpy(), vb.ipiv.cpu().numpy()))
    finally:
        ra.vbatched_set_tuning(default)
    for o in out[1:]:
        for x, y in zip(out[0], o):
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("op", ["gesv", "getrf"])
def test_vbatched_graph_replays_equal_eager(torch_gpu, op):
    """The plan + persistent-kernel sequence captured in a graph (workspace
    allocated by an eager call first) and replayed several times, with eager
    work and synchronisations between replays: every replay equals the eager
    result bit for bit."""
    torch = torch_gpu
    n, nrhs, As, Bs = make(21, ORDERS, "float64")
    vb = load(torch, n, nrhs, As, Bs, "float64")
    A0, B0 = vb.A.clone(), vb.B.clone()
    h = ra.Handle()
    try:
        run = (lambda: vb.gesv(h)) if op == "gesv" else (lambda: vb.getrf(h))
        run()
        torch.cuda.synchronize()
        ref = (vb.A.clone(), vb.B.clone(), vb.ipiv.clone(), vb.info.clone())
        g = torch.cuda.CUDAGraph()
        vb.A.copy_(A0)
        vb.B.copy_(B0)
        torch.cuda.synchronize()
        with torch.cuda.graph(g):
            h.set_stream(torch.cuda.current_stream())
            run()
        h.set_stream(torch.cuda.current_stream())
        for rep in range(4):
            vb.A.copy_(A0)
            vb.B.copy_(B0)
            vb.ipiv.zero_()
            vb.info.fill_(-7)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            for x, y in zip((vb.A, vb.B, vb.ipiv, vb.info), ref):
                assert torch.equal(x, y), f"replay {rep}"
            vb.A.mul_(1.0)  # eager work between replays
    finally:
        h.close()
