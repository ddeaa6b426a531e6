"""torch.autograd wrapper of the batched solve X = A^-1 B.

Forward: rocsolver_{s,d}gesv_strided_batched on copies of A and B (the LU,
pivots and X are kept for the backward pass).  Backward: the hand-written
adjoint rocsolver_amd_{s,d}gesv_adjoint_strided_batched,
    dB = A^-T dX  (transposed getrs reusing the stored LU and pivots)
    dA = -dB X^T  (batched outer product).
Tensors use the library's batch layout: A (bc, n, n) and B (bc, nrhs, n),
each matrix column-major (A[b, j, i] is entry (i, j) of matrix b).
"""
from __future__ import annotations

import torch

import rocsolver_amd as ra

_handles: dict = {}


def _handle() -> ra.Handle:
    key = (torch.cuda.current_device(), torch.cuda.current_stream().cuda_stream)
    h = _handles.get(key)
    if h is None:
        h = _handles[key] = ra.Handle()
    return h


class BatchedSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, B):
        if A.dtype != B.dtype or A.dtype not in (torch.float32, torch.float64):
            raise TypeError("A and B must both be float32 or both float64")
        h = _handle()
        LU = A.detach().clone().contiguous()
        X = B.detach().clone().contiguous()
        ipiv, info = ra.gesv(h, LU, X)
        ctx.save_for_backward(LU, ipiv, X)
        ctx.info = info
        return X

    @staticmethod
    def backward(ctx, dX):
        LU, ipiv, X = ctx.saved_tensors
        bc, n, lda = LU.shape
        _, nrhs, ldx = X.shape
        dX = dX.contiguous()
        dB = torch.empty_like(X)
        dA = torch.empty_like(LU)
        p = "d" if LU.dtype == torch.float64 else "s"
        ra.check(f"rocsolver_amd_{p}gesv_adjoint_strided_batched", _handle().value, n, nrhs,
                 LU.data_ptr(), lda, n * lda, ipiv.data_ptr(), ipiv.stride(0), X.data_ptr(), ldx,
                 nrhs * ldx, dX.data_ptr(), ldx, nrhs * ldx, dB.data_ptr(), ldx, nrhs * ldx,
                 dA.data_ptr(), lda, n * lda, bc)
        return dA, dB


def solve(A, B):
    """X = A^-1 B for every system, differentiable in A and B."""
    return BatchedSolve.apply(A, B)
