"""Seeded synthetic inputs for the batched LU path.

Two families:

* ``reference_*``: the matrices the reference clients test with.
  ``std::mt19937`` seeded 69069 (clients/rocblascommon/clients_utility.cpp:19),
  integers from ``std::uniform_int_distribution<int>(1, 10)``
  (clients/rocblascommon/rocblas_random.hpp:106-109) over each whole lda*n
  buffer with the seed reset per array (rocblas_vector.hpp:38-58); then +400
  on the diagonal and -4 off it, rows reversed for getrf, and optional zero
  columns in batch entries bc/4, bc/2, bc-1
  (clients/include/testing_getf2_getrf.hpp:129-178, testing_getrs.hpp:137-151,
  testing_gesv.hpp:150-176).  The integer draw restates libstdc++ (GCC >= 11):
  Lemire's multiply-shift with rejection over the raw 32-bit MT19937 stream;
  numpy's legacy-seeded MT19937 produces that same stream.
* ``config_*``: the five benchmark workloads (uniform/normal entries from a
  numpy PCG64 stream of a fixed seed).

Batches are numpy arrays of shape (bc, ncols, ld): column-major per matrix.
"""
from __future__ import annotations

import numpy as np

REFERENCE_SEED = 69069
BASE_SEED = 20261202


# -------------------------------------------------------------- reference
def _mt_uniform_int_1_10(count: int, bitgen: np.random.MT19937) -> np.ndarray:
    """libstdc++ uniform_int_distribution<int>(1,10) over a 32-bit MT19937.

    Each output is (x * 10) >> 32 + 1 for the next raw draw x whose low 32 bits
    of x * 10 are not below 2^32 mod 10; a rejected draw is skipped, so the
    outputs are the accepted subsequence of the raw stream."""
    rng = 10
    thr = np.uint64((1 << 32) % rng)
    parts, got = [], 0
    while got < count:
        raw = bitgen.random_raw(count - got).astype(np.uint64)
        prod = raw * np.uint64(rng)
        keep = (prod & np.uint64(0xFFFFFFFF)) >= thr
        vals = (prod[keep] >> np.uint64(32)).astype(np.int64) + 1
        parts.append(vals)
        got += vals.size
    return np.concatenate(parts)[:count]


def mt_int_stream(count: int) -> np.ndarray:
    """The first `count` draws of the reference clients' random_generator<T>()."""
    bg = np.random.MT19937(0)
    bg._legacy_seeding(REFERENCE_SEED)
    return _mt_uniform_int_1_10(count, bg)


def _reference_fill(bc: int, ncols: int, ld: int, dtype) -> np.ndarray:
    vals = mt_int_stream(bc * ncols * ld)
    return vals.astype(dtype).reshape(bc, ncols, ld)


def reference_getrf(m: int, n: int, lda: int, bc: int, singular: bool = False,
                    dtype=np.float64) -> np.ndarray:
    """Input of the reference getf2/getrf tests (testing_getf2_getrf.hpp:129-178)."""
    A = _reference_fill(bc, n, lda, dtype)
    for b in range(bc):
        M = A[b]  # M[j, i] = A(i, j)
        for j in range(n):
            for i in range(m):
                M[j, i] += 400 if i == j else -4
        rows = M[:n, :m].copy()
        M[:n, :m] = rows[:, ::-1]
        if singular and (b == bc // 4 or b == bc // 2 or b == bc - 1):
            for j0 in (n // 4 + b, n // 2 + b, n - 1 + b):
                j = j0 - (j0 // n) * n
                M[j, :m] = 0
    return A


def reference_gesv(n: int, nrhs: int, lda: int, ldb: int, bc: int, singular: bool = False,
                   dtype=np.float64):
    """A and B of the reference gesv tests (testing_gesv.hpp:150-176)."""
    A = _reference_fill(bc, n, lda, dtype)
    B = _reference_fill(bc, nrhs, ldb, dtype)
    for b in range(bc):
        M = A[b]
        for j in range(n):
            for i in range(n):
                M[j, i] += 400 if i == j else -4
        if singular and (b == bc // 4 or b == bc // 2 or b == bc - 1):
            for j0 in (n // 4 + b, n // 2 + b, n - 1 + b):
                j = j0 - (j0 // n) * n
                M[j, :n] = 0
    return A, B


def reference_getrs(n: int, nrhs: int, lda: int, ldb: int, bc: int, dtype=np.float64):
    """A (before its host LU) and B of the reference getrs tests (testing_getrs.hpp:137-151)."""
    A, B = reference_gesv(n, nrhs, lda, ldb, bc, False, dtype)
    return A, B


def _reference_zero_diagonal_positions(n: int, b: int):
    """Diagonal positions the reference tests zero in their singular batch
    entries (n/4 + b, n/2 + b, n - 1 + b, each mod n)."""
    return [j0 - (j0 // n) * n for j0 in (n // 4 + b, n // 2 + b, n - 1 + b)]


def reference_getri(n: int, lda: int, bc: int, dtype=np.float64) -> np.ndarray:
    """Input of the reference getri tests before their host LU
    (testing_getri.hpp:117-140): diagonal v/10 + 10, off-diagonal (v - 4)/10,
    rows reversed.  The singular entries are made after the LU by zeroing
    diagonal entries of U (reference_getri_singular)."""
    A = _reference_fill(bc, n, lda, dtype)
    for b in range(bc):
        M = A[b]  # M[j, i] = A(i, j)
        for j in range(n):
            for i in range(n):
                M[j, i] = M[j, i] / 10.0 + 10 if i == j else (M[j, i] - 4) / 10.0
        rows = M[:n, :n].copy()
        M[:n, :n] = rows[:, ::-1]
    return A


def reference_getri_singular(LU: np.ndarray, n: int) -> None:
    """Zero the diagonal entries of U that the reference getri tests zero in
    batch entries bc/4, bc/2 and bc-1 (testing_getri.hpp:142-157), in place."""
    bc = LU.shape[0]
    for b in range(bc):
        if b == bc // 4 or b == bc // 2 or b == bc - 1:
            for i in _reference_zero_diagonal_positions(n, b):
                LU[b, i, i] = 0


def reference_potrf(n: int, lda: int, bc: int, singular: bool = False,
                    dtype=np.float64) -> np.ndarray:
    """Input of the reference potf2/potrf tests (testing_potf2_potrf.hpp:116-140):
    the clients' integers with each diagonal entry a replaced by a*a*400; in
    the singular entries (bc/4, bc/2, bc-1) three diagonal entries are 0."""
    A = _reference_fill(bc, n, lda, dtype)
    for b in range(bc):
        for i in range(n):
            A[b, i, i] = A[b, i, i] * A[b, i, i] * 400
        if singular and (b == bc // 4 or b == bc // 2 or b == bc - 1):
            for i in _reference_zero_diagonal_positions(n, b):
                A[b, i, i] = 0
    return A


# ----------------------------------------------------------------- configs
def rng(config_index: int, seed: int = BASE_SEED) -> np.random.Generator:
    return np.random.default_rng(seed + config_index)


def config1(bc: int = 1024, n: int = 32, seed: int = BASE_SEED):
    """FP64, 32x32, lda=32, U(-1,1); one RHS U(-1,1)."""
    g = rng(1, seed)
    A = g.uniform(-1.0, 1.0, size=(bc, n, n))
    B = g.uniform(-1.0, 1.0, size=(bc, 1, n))
    return A, B


def config2(bc: int = 16384, n: int = 128, nrhs: int = 4, seed: int = BASE_SEED):
    """FP64, 128x128, N(0,1) + 0.1*n on the diagonal; 4 RHS N(0,1)."""
    g = rng(2, seed)
    A = g.standard_normal(size=(bc, n, n))
    idx = np.arange(n)
    A[:, idx, idx] += 0.1 * n
    B = g.standard_normal(size=(bc, nrhs, n))
    return A, B


def config3_kappa(bc: int, seed: int = BASE_SEED) -> np.ndarray:
    g = rng(3, seed)
    return np.exp(g.uniform(np.log(1e2), np.log(1e6), size=bc))


def config3(bc: int = 4096, n: int = 256, nrhs: int = 8, seed: int = BASE_SEED):
    """A = U diag(s) V^T, U, V from QR of N(0,1), s log-spaced in [1, kappa],
    kappa log-uniform in [1e2, 1e6] per matrix; 8 RHS N(0,1).  CPU numpy build
    (use config3_torch for full-size batches)."""
    g = rng(3, seed)
    kappa = np.exp(g.uniform(np.log(1e2), np.log(1e6), size=bc))
    G1 = g.standard_normal(size=(bc, n, n))
    G2 = g.standard_normal(size=(bc, n, n))
    U, _ = np.linalg.qr(G1)
    V, _ = np.linalg.qr(G2)
    t = np.linspace(0.0, 1.0, n)
    s = np.exp(np.log(kappa)[:, None] * t[None, :])[:, ::-1]  # kappa ... 1
    M = (U * s[:, None, :]) @ np.swapaxes(V, 1, 2)  # row-major M[b, i, j]
    A = np.ascontiguousarray(np.swapaxes(M, 1, 2))  # column-major per matrix
    B = g.standard_normal(size=(bc, nrhs, n))
    return A, B, kappa


def config3_torch(bc: int = 4096, n: int = 256, nrhs: int = 8, seed: int = BASE_SEED,
                  device="cuda"):
    """config3 generated with torch on `device` (for full-size batches)."""
    import torch
    gen = torch.Generator(device=device)
    gen.manual_seed(seed + 3)
    kappa = torch.exp(torch.empty(bc, dtype=torch.float64, device=device).uniform_(
        np.log(1e2), np.log(1e6), generator=gen))
    out = torch.empty((bc, n, n), dtype=torch.float64, device=device)
    chunk = 64  # torch's batched QR fails to allocate its workspace for large batches
    t = torch.linspace(0.0, 1.0, n, dtype=torch.float64, device=device)

    def qr_q(G):
        try:
            return torch.linalg.qr(G)[0]
        except RuntimeError:
            return torch.linalg.qr(G.cpu())[0].to(device)

    for b0 in range(0, bc, chunk):
        b1 = min(bc, b0 + chunk)
        G1 = torch.randn((b1 - b0, n, n), dtype=torch.float64, device=device, generator=gen)
        G2 = torch.randn((b1 - b0, n, n), dtype=torch.float64, device=device, generator=gen)
        U = qr_q(G1)
        V = qr_q(G2)
        s = torch.exp(torch.log(kappa[b0:b1])[:, None] * t[None, :]).flip(1)
        M = torch.matmul(U * s[:, None, :], V.transpose(1, 2).contiguous())
        out[b0:b1] = M.transpose(1, 2)
    B = torch.randn((bc, nrhs, n), dtype=torch.float64, device=device, generator=gen)
    return out, B, kappa


def config4_sizes(count: int = 10000, seed: int = BASE_SEED):
    g = rng(4, seed)
    n = np.rint(np.exp(g.uniform(np.log(8), np.log(512), size=count))).astype(np.int64)
    nrhs = g.integers(1, 5, size=count)
    return n, nrhs


def config4(count: int = 10000, seed: int = BASE_SEED):
    """Variable-size: n log-uniform integers in [8, 512], U(-1,1) + n*I, 1..4 RHS."""
    n, nrhs = config4_sizes(count, seed)
    g = rng(40, seed)
    As, Bs = [], []
    for k in range(count):
        a = g.uniform(-1.0, 1.0, size=(n[k], n[k]))
        a[np.arange(n[k]), np.arange(n[k])] += n[k]
        As.append(a)
        Bs.append(g.uniform(-1.0, 1.0, size=(nrhs[k], n[k])))
    return n, nrhs, As, Bs


def config5(bc: int = 2048, n: int = 96, nrhs: int = 16, seed: int = BASE_SEED):
    """Adjoint workload: A N(0,1)+10 I, B (16 RHS) and dX N(0,1)."""
    g = rng(5, seed)
    A = g.standard_normal(size=(bc, n, n))
    idx = np.arange(n)
    A[:, idx, idx] += 10.0
    B = g.standard_normal(size=(bc, nrhs, n))
    dX = g.standard_normal(size=(bc, nrhs, n))
    return A, B, dX
