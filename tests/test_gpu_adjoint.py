"""GPU checks of the adjoint of X = A^-1 B (rocsolver_amd_{s,d}gesv_adjoint_strided_batched
through rocsolver_amd.autograd.BatchedSolve).

Pinned by the closed form of the gradient computed independently in FP64 on
the CPU (torch.linalg.solve's autograd; torch is not the reference, it is an
independent FP64 implementation of the same formula) and by central finite
differences (h = 1e-6, FP64) of a scalar loss on sampled entries.
"""
import numpy as np
import pytest

from rocsolver_amd import workloads as wl

pytestmark = pytest.mark.gpu


def reference_grads(torch, A, B, dX):
    """FP64 CPU gradient of <X, dX> in the library's column-major batch layout."""
    M = A.detach().cpu().double().transpose(1, 2).contiguous().requires_grad_(True)
    Bm = B.detach().cpu().double().transpose(1, 2).contiguous().requires_grad_(True)
    X = torch.linalg.solve(M, Bm)
    (X * dX.detach().cpu().double().transpose(1, 2)).sum().backward()
    return M.grad.transpose(1, 2), Bm.grad.transpose(1, 2), X.transpose(1, 2)


@pytest.mark.parametrize("n,nrhs,bc", [(96, 16, 8), (32, 4, 8), (200, 3, 3), (300, 2, 2)])
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_adjoint_matches_fp64_reference(torch_gpu, n, nrhs, bc, dtype):
    torch = torch_gpu
    from rocsolver_amd.autograd import solve
    print(f"adjoint n={n} nrhs={nrhs} bc={bc} {dtype}", flush=True)
    A0, B0, dX0 = wl.config5(bc=bc, n=n, nrhs=nrhs)
    dt = getattr(torch, dtype)
    A = torch.from_numpy(A0).to("cuda", dt).requires_grad_(True)
    B = torch.from_numpy(B0).to("cuda", dt).requires_grad_(True)
    dX = torch.from_numpy(dX0).to("cuda", dt)
    X = solve(A, B)
    X.backward(dX)
    dA_ref, dB_ref, X_ref = reference_grads(torch, A, B, dX)
    # forward-error scale of a backward-stable solve: kappa_2(A) * n * eps
    kappa = np.linalg.cond(np.swapaxes(A0, 1, 2)).max()
    tol = 10 * kappa * n * float(np.finfo(getattr(np, dtype)).eps)
    for got, ref in ((X, X_ref), (B.grad, dB_ref), (A.grad, dA_ref)):
        g = got.detach().cpu().double()
        err = (g - ref).abs().max() / ref.abs().max()
        assert float(err) < tol, (float(err), tol)


def test_adjoint_central_differences(torch_gpu):
    torch = torch_gpu
    from rocsolver_amd.autograd import solve
    n, nrhs, bc = 96, 16, 2
    A0, B0, dX0 = wl.config5(bc=bc, n=n, nrhs=nrhs)
    A = torch.from_numpy(A0).cuda().requires_grad_(True)
    B = torch.from_numpy(B0).cuda().requires_grad_(True)
    W = torch.from_numpy(dX0).cuda()
    loss = lambda a, b: float((solve(a, b) * W).sum())
    X = solve(A, B)
    (X * W).sum().backward()
    h = 1e-6
    g = np.random.default_rng(1)
    for _ in range(6):
        b, j, i = int(g.integers(bc)), int(g.integers(n)), int(g.integers(n))
        Ap, Am = A.detach().clone(), A.detach().clone()
        Ap[b, j, i] += h
        Am[b, j, i] -= h
        fd = (loss(Ap, B.detach()) - loss(Am, B.detach())) / (2 * h)
        assert abs(fd - float(A.grad[b, j, i])) <= 1e-6 * max(1.0, abs(fd))
        c = int(g.integers(nrhs))
        Bp, Bm = B.detach().clone(), B.detach().clone()
        Bp[b, c, i] += h
        Bm[b, c, i] -= h
        fd = (loss(A.detach(), Bp) - loss(A.detach(), Bm)) / (2 * h)
        assert abs(fd - float(B.grad[b, c, i])) <= 1e-6 * max(1.0, abs(fd))
