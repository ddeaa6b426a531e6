"""GPU checks of the mixed-precision solve (rocsolver_amd_dsgesv_strided_batched).

Pinned against the oracle's LAPACK-dsgesv restatement (itself pinned against
LAPACK DSGESV in test_oracle.py): the same class of outcome per matrix
(converged by refinement vs FP64 fallback), FP64-level backward error
||b - A x|| / (||A|| ||x|| + ||b||) <= 10 n eps64 (the target accuracy),
and forward error against the FP64 solution within kappa * n * eps64.
"""
import numpy as np
import pytest

import oracle
import rocsolver_amd as ra
from rocsolver_amd import workloads as wl

pytestmark = pytest.mark.gpu
EPS64 = np.finfo(np.float64).eps


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


# batches of 256 and more are solved in two staggered halves on two streams
@pytest.mark.parametrize("n,nrhs,bc", [(96, 16, 6), (128, 4, 5), (256, 8, 6), (200, 1, 4), (512, 2, 3),
                                       (70, 3, 3), (96, 3, 301), (1, 1, 4), (5, 2, 3), (17, 3, 4),
                                       (32, 8, 6), (64, 4, 5)])
def test_dsgesv_against_oracle(torch_gpu, handle, n, nrhs, bc):
    print(f"dsgesv n={n} nrhs={nrhs} bc={bc}", flush=True)
    A, B, kappa = wl.config3(bc=bc, n=n, nrhs=nrhs)
    Xo, it_o, info_o = oracle.dsgesv(A, n, B)
    Ad, Bd = dev(torch_gpu, A), dev(torch_gpu, B)
    X, it, info, berr = ra.dsgesv(handle, Ad, Bd)
    torch_gpu.cuda.synchronize()
    X, it, info, berr = X.cpu().numpy(), it.cpu().numpy(), info.cpu().numpy(), berr.cpu().numpy()
    np.testing.assert_array_equal(Ad.cpu().numpy(), A)  # A untouched
    np.testing.assert_array_equal(Bd.cpu().numpy(), B)
    np.testing.assert_array_equal(info, info_o)
    assert np.all((it >= 0) == (it_o >= 0))  # same outcome class per matrix
    assert np.all(np.abs(it[it >= 0] - it_o[it >= 0]) <= 2)
    be = oracle.solve_backward_error(A, n, X, B)
    assert be.max() <= 10 * n * EPS64, be
    for b in range(bc):
        fe = np.abs(X[b] - Xo[b]).max() / np.abs(Xo[b]).max()
        assert fe <= kappa[b] * n * EPS64 * 10
        if it[b] >= 0:
            assert 0 <= berr[b] <= np.sqrt(n) * EPS64 * 1.0001


@pytest.mark.parametrize("factor", ["fp32", "fp16"])
@pytest.mark.parametrize("n,nrhs,bc", [(513, 2, 3), (640, 3, 2)])
def test_dsgesv_beyond_fused_orders_uses_fp64(torch_gpu, handle, factor, n, nrhs, bc):
    """n > 512: every system is solved by the FP64 LU, iter = -1 (LAPACK
    DSGESV's code for a fall-back to full precision for implementation-specific
    reasons), berr = -1, info as the FP64 LU's (the oracle's fall-back LU
    finds the same first zero pivot in the singular entry)."""
    print(f"dsgesv beyond fused n={n} nrhs={nrhs} bc={bc} factor={factor}", flush=True)
    g = np.random.default_rng(n)
    A = g.standard_normal((bc, n, n))
    A[1][20, :] = 0.0  # column 20 zero -> first zero pivot at 21
    B = g.standard_normal((bc, nrhs, n))
    Xo, it_o, info_o = oracle.dsgesv(A, n, B)
    Ad, Bd = dev(torch_gpu, A), dev(torch_gpu, B)
    X, it, info, berr = ra.dsgesv(handle, Ad, Bd, factor=factor)
    torch_gpu.cuda.synchronize()
    X, it, info, berr = X.cpu().numpy(), it.cpu().numpy(), info.cpu().numpy(), berr.cpu().numpy()
    np.testing.assert_array_equal(Ad.cpu().numpy(), A)
    np.testing.assert_array_equal(Bd.cpu().numpy(), B)
    np.testing.assert_array_equal(it, -1)
    np.testing.assert_array_equal(berr, -1.0)
    np.testing.assert_array_equal(info, info_o)
    assert info[1] == 21 and info[0] == 0
    ok = info == 0
    be = oracle.solve_backward_error(A[ok], n, X[ok], B[ok])
    assert be.max() <= 10 * n * EPS64, be
    for b in np.flatnonzero(ok):
        kappa = np.linalg.cond(A[b])
        fe = np.abs(X[b] - Xo[b]).max() / np.abs(Xo[b]).max()
        assert fe <= 10 * kappa * n * EPS64, (b, fe, kappa)


def test_dsgesv_fallback_paths(torch_gpu, handle):
    g = np.random.default_rng(7)
    n, nrhs, bc = 80, 2, 4
    A = g.standard_normal((bc, n, n)) + 6 * np.eye(n)[None]
    # entry 1: kappa ~ 1e12 -> no convergence -> FP64 path (ITER = -31)
    U, _ = np.linalg.qr(g.standard_normal((n, n)))
    V, _ = np.linalg.qr(g.standard_normal((n, n)))
    A[1] = ((U * np.logspace(0, -12, n)) @ V.T).T
    # entry 2: singular (zero column) -> FP32 LU zero pivot (-3), FP64 info > 0
    A[2][10, :] = 0.0
    # entry 3: an entry beyond the FP32 range -> ITER = -2
    A[3][0, 0] = 1e39
    B = g.standard_normal((bc, nrhs, n))
    Xo, it_o, info_o = oracle.dsgesv(A, n, B)
    X, it, info, berr = ra.dsgesv(handle, dev(torch_gpu, A), dev(torch_gpu, B))
    torch_gpu.cuda.synchronize()
    it, info, X = it.cpu().numpy(), info.cpu().numpy(), X.cpu().numpy()
    np.testing.assert_array_equal(it[1:], [-31, -3, -2])
    np.testing.assert_array_equal(it_o[1:], [-31, -3, -2])
    assert it[0] >= 0
    np.testing.assert_array_equal(info, info_o)
    assert info[2] == 11
    ok = info == 0
    be = oracle.solve_backward_error(A[ok], n, X[ok], B[ok])
    assert be.max() <= 10 * n * EPS64


@pytest.mark.parametrize("factor", ["fp16", "bf16"])
@pytest.mark.parametrize("n,nrhs,bc", [(96, 4, 6), (256, 8, 4), (300, 3, 2), (80, 2, 260),
                                       (40, 2, 4)])
def test_half_width_factor_refines_to_fp64(torch_gpu, handle, factor, n, nrhs, bc):
    """Equilibrated FP16/BF16-operand LU + FP64 refinement: well-conditioned
    systems converge by refinement to FP64 backward error; the result agrees
    with the oracle's FP64 solution within kappa * n * eps64."""
    print(f"dxgesv {factor} n={n} nrhs={nrhs} bc={bc}", flush=True)
    g = np.random.default_rng(n + nrhs)
    # kappa ~ 10 with different row and column scales (2^-10 .. 2^10), which
    # the equilibration removes exactly
    A = g.standard_normal((bc, n, n)) + 2.0 * np.sqrt(n) * np.eye(n)[None]
    dc = np.exp2(g.integers(-10, 11, size=(bc, n, 1)))  # column scales (A is stored transposed)
    dr = np.exp2(g.integers(-10, 11, size=(bc, 1, n)))  # row scales
    A *= dc * dr
    B = g.standard_normal((bc, nrhs, n))
    Xo = B.copy()
    _, info_o = oracle.gesv(A.copy(), n, Xo)
    X, it, info, berr = ra.dsgesv(handle, dev(torch_gpu, A), dev(torch_gpu, B), factor=factor)
    torch_gpu.cuda.synchronize()
    X, it, info = X.cpu().numpy(), it.cpu().numpy(), info.cpu().numpy()
    assert np.all(info == 0) and np.all(info_o == 0)
    assert np.all(it >= 1), it  # refined (a half-width LU alone is not FP64 accurate)
    assert np.all(it <= 20), it
    be = oracle.solve_backward_error(A, n, X, B)
    assert be.max() <= 10 * n * EPS64, be
    # (no forward-error check here: with rows and columns scaled by 2^+-10 a
    # normwise-small residual -- what refinement and FP64 GEPP both deliver --
    # does not determine x to better than kappa(A) * eps ~ 2^40 * eps; the
    # well-scaled case below checks the forward error)


def test_half_width_factor_fallback_and_fp32_identity(torch_gpu, handle):
    g = np.random.default_rng(21)
    n, nrhs, bc = 128, 2, 3
    A = g.standard_normal((bc, n, n)) + 3 * np.sqrt(n) * np.eye(n)[None]
    U, _ = np.linalg.qr(g.standard_normal((n, n)))
    V, _ = np.linalg.qr(g.standard_normal((n, n)))
    A[1] = ((U * np.logspace(0, -9, n)) @ V.T).T  # kappa 1e9: FP16 refinement cannot converge
    A[2] *= 1e300  # beyond FP32 range: factor 0 gives -2, the scaled factors do not
    B = g.standard_normal((bc, nrhs, n))
    Ad, Bd = dev(torch_gpu, A), dev(torch_gpu, B)
    X16, it16, info16, _ = ra.dsgesv(handle, Ad, Bd, factor="fp16")
    X32, it32, info32, b32 = ra.dsgesv(handle, Ad, Bd, factor="fp32")
    torch_gpu.cuda.synchronize()
    it16, it32 = it16.cpu().numpy(), it32.cpu().numpy()
    assert it16[0] >= 1 and it16[1] == -31 and it16[2] >= 1
    assert it32[2] == -2
    assert np.all(info16.cpu().numpy() == 0)
    be = oracle.solve_backward_error(A, n, X16.cpu().numpy(), B)
    assert be.max() <= 10 * n * EPS64, be
    # well-scaled, well-conditioned entry 0: forward error against the FP64 oracle
    Xo = B[:1].copy()
    oracle.gesv(A[:1].copy(), n, Xo)
    fe = np.abs(X16.cpu().numpy()[0] - Xo[0]).max() / np.abs(Xo[0]).max()
    assert fe <= 1e3 * n * EPS64, fe
    # factor 0 through the general entry point is the dsgesv entry point
    X0 = torch_gpu.zeros_like(X32)
    it0 = torch_gpu.zeros(bc, dtype=torch_gpu.int32, device="cuda")
    info0 = torch_gpu.zeros(bc, dtype=torch_gpu.int32, device="cuda")
    b0 = torch_gpu.zeros(bc, dtype=torch_gpu.float64, device="cuda")
    ra.check("rocsolver_amd_dxgesv_strided_batched", handle.value, 0, n, nrhs, Ad.data_ptr(), n,
             n * n, Bd.data_ptr(), n, n * nrhs, X0.data_ptr(), n, n * nrhs, it0.data_ptr(),
             info0.data_ptr(), b0.data_ptr(), bc)
    torch_gpu.cuda.synchronize()
    assert torch_gpu.equal(X0, X32)
    np.testing.assert_array_equal(it0.cpu().numpy(), it32)


def test_dsgesv_graph_capture_replays(torch_gpu):
    """The mixed solve of a batch that takes the two-stream path, captured in a
    graph on the handle's stream: every replay (with eager work and
    synchronisations between replays) equals the eager call bit for bit."""
    torch = torch_gpu
    n, nrhs, bc = 96, 2, 320
    A, B, _ = wl.config3(bc=bc, n=n, nrhs=nrhs, seed=123)
    s = torch.cuda.Stream()
    h = ra.Handle(s)
    try:
        Ad, Bd = dev(torch, A), dev(torch, B)
        X = torch.zeros((bc, nrhs, n), dtype=torch.float64, device="cuda")
        it = torch.zeros(bc, dtype=torch.int32, device="cuda")
        info = torch.zeros(bc, dtype=torch.int32, device="cuda")

        def call():
            ra.check("rocsolver_amd_dsgesv_strided_batched", h.value, n, nrhs, Ad.data_ptr(), n,
                     n * n, Bd.data_ptr(), n, n * nrhs, X.data_ptr(), n, n * nrhs,
                     it.data_ptr(), info.data_ptr(), 0, bc)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            call()  # eager: sizes the handle's workspace, reference result
        torch.cuda.synchronize()
        X0, it0, info0 = X.clone(), it.clone(), info.clone()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            call()
        torch.cuda.synchronize()
        for rep in range(3):
            X.fill_(float(rep + 1))  # eager work between replays
            it.fill_(-99)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(X, X0), f"replay {rep}"
            assert torch.equal(it, it0) and torch.equal(info, info0), f"replay {rep}"
        be = oracle.solve_backward_error(A, n, X0.cpu().numpy(), B)
        assert be.max() <= 10 * n * EPS64
    finally:
        h.close()


def gpu_backward_error(torch, A, X, B):
    """max over right-hand sides of ||b - A x||_inf / (||A||_inf ||x||_inf +
    ||b||_inf) per system, in FP64 on the GPU.  A (bc, n, n) column-major per
    matrix (A[b, j, i] = a_ij), X and B (bc, nrhs, n)."""
    AX = torch.matmul(X, A)  # (A x_c)_i = sum_j A[b, j, i] X[b, c, j]
    rn = (B - AX).abs().amax(dim=2)
    an = A.abs().sum(dim=1).amax(dim=1)
    den = an[:, None] * X.abs().amax(dim=2) + B.abs().amax(dim=2)
    return (rn / den).amax(dim=1)


def test_dsgesv_metric_batch(torch_gpu, handle):
    """The benchmark configuration itself: n=256, batch 4096, 8 right-hand
    sides, kappa log-uniform in [1e2, 1e6] (the batch is large enough for the
    two-stream split).  Every system reaches FP64 backward error <= 10 n eps64
    (checked on the GPU for all 4096); info exactly and the outcome class per
    system as the oracle's dsgesv on every 64th system."""
    torch = torch_gpu
    n, nrhs, bc = 256, 8, 4096
    print(f"dsgesv metric batch n={n} nrhs={nrhs} bc={bc}", flush=True)
    A, B, kappa = wl.config3_torch(bc=bc, n=n, nrhs=nrhs)
    X, it, info, berr = ra.dsgesv(handle, A, B)
    torch.cuda.synchronize()
    be = gpu_backward_error(torch, A, X, B)
    assert float(be.max()) <= 10 * n * EPS64, float(be.max())
    assert (info == 0).all()
    itc = it.cpu().numpy()
    idx = np.arange(0, bc, 64)
    Ah = A[idx].cpu().numpy()
    Bh = B[idx].cpu().numpy()
    Xo, it_o, info_o = oracle.dsgesv(Ah, n, Bh)
    np.testing.assert_array_equal(info.cpu().numpy()[idx], info_o)
    np.testing.assert_array_equal(itc[idx] >= 0, it_o >= 0)
    conv = it_o >= 0
    assert np.all(np.abs(itc[idx][conv] - it_o[conv]) <= 2)
    # the per-system FP64 forward error against the oracle's solution
    Xs = X[idx].cpu().numpy()
    for k, b in enumerate(idx):
        fe = np.abs(Xs[k] - Xo[k]).max() / np.abs(Xo[k]).max()
        assert fe <= kappa[b].item() * n * EPS64 * 10
    print(f"iterations: {np.bincount(itc[itc >= 0]).tolist()}, fallbacks {(itc < 0).sum()}, "
          f"max backward error {float(be.max()):.2e}", flush=True)


def test_dsgesv_capture_and_eager_threads(torch_gpu):
    """One thread captures a two-stream mixed solve in a graph while another
    thread runs an eager two-stream solve on its own handle and stream: the
    library's internal streams are per caller stream, so the eager call runs
    (and is correct) and the graph replays the captured call exactly."""
    import threading
    torch = torch_gpu
    n, nrhs, bc = 96, 2, 300
    A1, B1, _ = wl.config3(bc=bc, n=n, nrhs=nrhs, seed=41)
    A2, B2, _ = wl.config3(bc=bc, n=n, nrhs=nrhs, seed=42)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    ha, hb = ra.Handle(sa), ra.Handle(sb)
    try:
        bufs = []
        for A, B in ((A1, B1), (A2, B2)):
            bufs.append((dev(torch, A), dev(torch, B),
                         torch.zeros((bc, nrhs, n), dtype=torch.float64, device="cuda"),
                         torch.zeros(bc, dtype=torch.int32, device="cuda"),
                         torch.zeros(bc, dtype=torch.int32, device="cuda")))

        def call(h, k):
            Ad, Bd, X, it, info = bufs[k]
            return ra.call("rocsolver_amd_dsgesv_strided_batched", h.value, n, nrhs, Ad.data_ptr(),
                           n, n * n, Bd.data_ptr(), n, n * nrhs, X.data_ptr(), n, n * nrhs,
                           it.data_ptr(), info.data_ptr(), 0, bc)
        # eager references (also size both handles' workspaces)
        with torch.cuda.stream(sa):
            assert call(ha, 0) == 0
        with torch.cuda.stream(sb):
            assert call(hb, 1) == 0
        torch.cuda.synchronize()
        ref = [(b[2].clone(), b[3].clone()) for b in bufs]
        for b in bufs:
            b[2].zero_()
        torch.cuda.synchronize()
        inside = threading.Barrier(2)
        status = {}

        def capturer():
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=sa, capture_error_mode="thread_local"):
                status["a"] = call(ha, 0)
                inside.wait()  # the other thread runs while this capture is open
                inside.wait()
            status["graph"] = g

        def eager():
            inside.wait()
            with torch.cuda.stream(sb):
                status["b"] = call(hb, 1)
            sb.synchronize()
            inside.wait()
        ta, tb = threading.Thread(target=capturer), threading.Thread(target=eager)
        ta.start()
        tb.start()
        ta.join(60)
        tb.join(60)
        assert status.get("a") == 0 and status.get("b") == 0, status
        torch.cuda.synchronize()
        assert torch.equal(bufs[1][2], ref[1][0]) and torch.equal(bufs[1][3], ref[1][1])
        assert not torch.equal(bufs[0][2], ref[0][0])  # captured, not run
        status["graph"].replay()
        torch.cuda.synchronize()
        assert torch.equal(bufs[0][2], ref[0][0]) and torch.equal(bufs[0][3], ref[0][1])
        assert torch.equal(bufs[1][2], ref[1][0])  # the eager result was not touched
    finally:
        ha.close()
        hb.close()


@pytest.mark.parametrize("nrhs", [16, 12])
def test_dsgesv_wide_rhs_iter(torch_gpu, handle, nrhs):
    """More right-hand sides than one refinement group (8 columns): every
    group refines until its own columns pass LAPACK's stopping test, and ITER
    is the largest group count.  Pinned exactly: the solution of each group
    equals, bit for bit, that of a call with the group's columns alone, and
    ITER is the maximum of those calls' ITER.  Against the oracle (LAPACK
    semantics: all columns refined together) the outcome class matches and
    ITER stays within the FP32-rounding margin of the other tests."""
    torch = torch_gpu
    n, bc = 128, 24
    A, B, kappa = wl.config3(bc=bc, n=n, nrhs=nrhs, seed=77)
    Ad = dev(torch, A)
    X, it, info, berr = ra.dsgesv(handle, Ad, dev(torch, B))
    torch.cuda.synchronize()
    it, X = it.cpu().numpy(), X.cpu().numpy()
    its = []
    for c0 in range(0, nrhs, 8):
        Bg = np.ascontiguousarray(B[:, c0:c0 + 8, :])
        Xg, itg, _, _ = ra.dsgesv(handle, Ad, dev(torch, Bg))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(X[:, c0:c0 + 8, :], Xg.cpu().numpy())
        its.append(itg.cpu().numpy())
    np.testing.assert_array_equal(it, np.max(its, axis=0))
    Xo, it_o, info_o = oracle.dsgesv(A, n, B)
    assert np.all((it >= 0) == (it_o >= 0))
    assert np.all(np.abs(it - it_o) <= 2)
    be = oracle.solve_backward_error(A, n, X, B)
    assert be.max() <= 10 * n * EPS64


@pytest.mark.parametrize("factor", ["fp16", "bf16"])
def test_half_width_factors_on_metric_workload(torch_gpu, handle, factor):
    """The metric workload's matrices (A = U diag(s) V^T, kappa log-uniform in
    [1e2, 1e6], 8 RHS) with FP16 / BF16 factors: per matrix the outcome
    (refined vs FP64 fallback) is reported against the oracle's dsgesv (FP32
    factors): a system refined with half-width factors is one the FP32 path
    also refines; every system reaches FP64 backward error <= 10 n eps64 and
    the oracle's solution within 10 kappa n eps64, refined or not."""
    torch = torch_gpu
    n, nrhs, bc = 256, 8, 48
    A, B, kappa = wl.config3(bc=bc, n=n, nrhs=nrhs, seed=2026)
    Xo, it_o, info_o = oracle.dsgesv(A, n, B)
    X, it, info, berr = ra.dsgesv(handle, dev(torch, A), dev(torch, B), factor=factor)
    torch.cuda.synchronize()
    X, it, info = X.cpu().numpy(), it.cpu().numpy(), info.cpu().numpy()
    refined = it >= 0
    print(f"{factor}: refined {refined.sum()} / {bc} (oracle FP32: {(it_o >= 0).sum()}), "
          f"fallbacks {(it < 0).sum()}, iterations {sorted(set(it[refined].tolist()))}", flush=True)
    assert np.all(it_o[refined] >= 0)
    assert np.all((it < 0) <= (it == -31))  # the only fallback cause here: no convergence
    np.testing.assert_array_equal(info, info_o)
    be = oracle.solve_backward_error(A, n, X, B)
    assert be.max() <= 10 * n * EPS64
    for b in range(bc):
        fe = np.abs(X[b] - Xo[b]).max() / np.abs(Xo[b]).max()
        assert fe <= 10 * kappa[b] * n * EPS64
