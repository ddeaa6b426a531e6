"""GPU check of rocsolver_amd_dgesv_vbatched_noalloc, the variable-size FP64
gesv whose work queue is the library's static device array (no allocation).

Its results must equal, bit for bit, those of rocsolver_amd_dgesv_vbatched on
the same batch (same kernels; only the queue's memory differs), on two calls
in a row, the second with a smaller batch (the array keeps the first call's
queue, so stale items or counters would show up there).  Batches above the
array's limit are refused on the host before anything runs.
"""
import ctypes

import numpy as np
import pytest

import rocsolver_amd as ra

pytestmark = pytest.mark.gpu

ORDERS = [1, 5, 8, 9, 16, 17, 24, 31, 32, 33, 48, 64, 65, 100, 128, 129, 200, 256, 257, 384,
          512]


def make(seed, bc):
    g = np.random.default_rng(seed)
    n = g.choice(ORDERS, size=bc).astype(np.int64)
    nrhs = g.integers(1, 5, size=bc)
    As = [g.uniform(-1, 1, (k, k)) + k * np.eye(k) * (g.uniform() < 0.8) for k in n]
    for b in range(0, bc, 9):  # singular entries: a zero column (info > 0, B left unchanged)
        k = int(n[b])
        As[b][:, g.integers(k)] = 0.0
    Bs = [g.uniform(-1, 1, (r, k)) for k, r in zip(n, nrhs)]
    return n, nrhs, As, Bs


def solve(torch, handle, n, nrhs, As, Bs, noalloc):
    vb = ra.VBatch(n, nrhs, dtype=torch.float64)
    for b in range(len(n)):
        vb.A_of(b).copy_(torch.from_numpy(As[b]))
        vb.B_of(b).copy_(torch.from_numpy(Bs[b]))
    vb.gesv(handle, noalloc=noalloc)
    torch.cuda.synchronize()
    return (vb.A.cpu().numpy(), vb.B.cpu().numpy(), vb.ipiv.cpu().numpy(),
            vb.info.cpu().numpy())


def test_noalloc_gesv_equals_vbatched(torch_gpu, handle):
    torch = torch_gpu
    for seed, bc in ((21, 700), (22, 90)):  # second call smaller than the first
        n, nrhs, As, Bs = make(seed, bc)
        print(f"dgesv_vbatched_noalloc bc={bc} seed={seed}", flush=True)
        ref = solve(torch, handle, n, nrhs, As, Bs, noalloc=False)
        got = solve(torch, handle, n, nrhs, As, Bs, noalloc=True)
        for name, r, g in zip(("A", "B", "ipiv", "info"), ref, got):
            np.testing.assert_array_equal(g, r, err_msg=f"{name} bc={bc}")
        assert (ref[3] != 0).any(), "the batch holds singular matrices"


def test_noalloc_limits(torch_gpu, handle):
    L = ra.lib()
    f = L.rocsolver_amd_dgesv_vbatched_noalloc
    one = ctypes.c_void_p(8)  # never dereferenced: the size check comes first
    assert f(handle.value, one, one, one, one, one, one, one, one, 65537) == 4  # invalid_size
    assert f(handle.value, one, one, one, one, one, one, one, one, -1) == 4
    assert f(handle.value, None, None, None, None, None, None, None, None, 0) == 0
    assert f(handle.value, None, None, None, None, None, None, None, None, 3) == 3  # pointers
