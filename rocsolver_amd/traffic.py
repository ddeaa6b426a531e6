"""HBM traffic of the dominant kernels, from hardware counters.

Run under rocprofv3 once per counter (FETCH_SIZE and WRITE_SIZE do not fit in
one pass together with anything else):

  rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir> -o fetch -- python3 -m rocsolver_amd.traffic [getrf32|c2]
  rocprofv3 --pmc WRITE_SIZE --output-format csv -d <dir> -o write -- python3 -m rocsolver_amd.traffic [getrf32|c2]
  python3 -m rocsolver_amd.traffic --parse <dir>/.../fetch_counter_collection.csv \\
                                           <dir>/.../write_counter_collection.csv

The program first copies a 1 GiB device buffer through an elementwise kernel
(1 GiB read, 1 GiB written), which calibrates what the counters report per byte,
then launches, once each, on the metric workload's shapes (n = 256, batch
4096, 8 right-hand sides, FP64 input): the FP32 LU of the FP64 batch
(rocsolver_amd_dsgetrf_strided_batched) and the whole mixed-precision solve
(rocsolver_amd_dsgesv_strided_batched), then the plain FP64
rocsolver_dgesv_strided_batched of the same systems; with the argument
getrf32, instead,
one plain FP64 rocsolver_dgetrf_strided_batched at n = 32 (lda 32, stride
1024), batch 131072; with c2, one plain FP64 dgetrf and one dgetrs (4
right-hand sides) of 16384 matrices of order 128 (N(0,1) + 12.8 I).  The parser prints the counter values per
dispatch scaled by the copy's factor, per kernel and grid size.
"""
from __future__ import annotations

import csv
import sys

COPY_BYTES = 1 << 30
N, BC, NRHS = 256, 4096, 8
N32, BC32 = 32, 131072
N128, BC128 = 128, 16384


def run(workload: str = "mixed") -> None:
    import torch

    import rocsolver_amd as ra
    torch.cuda.set_device(0)
    src = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.mul(src, 1.0, out=dst)  # elementwise kernel: reads 1 GiB, writes 1 GiB
    torch.cuda.synchronize()
    print("copy of 1 GiB done", flush=True)
    del src, dst
    if workload == "c2":
        n, bc, nrhs = N128, BC128, 4
        g = torch.Generator(device="cuda").manual_seed(20261204)
        A = torch.randn((bc, n, n), dtype=torch.float64, device="cuda", generator=g)
        A.diagonal(dim1=1, dim2=2).add_(0.1 * n)
        B = torch.randn((bc, nrhs, n), dtype=torch.float64, device="cuda", generator=g)
        ipiv = torch.zeros((bc, n), dtype=torch.int32, device="cuda")
        info = torch.zeros(bc, dtype=torch.int32, device="cuda")
        h = ra.Handle()
        ra.check("rocsolver_dgetrf_strided_batched", h.value, n, n, A.data_ptr(), n, n * n,
                 ipiv.data_ptr(), n, info.data_ptr(), bc)
        ra.check("rocsolver_dgetrs_strided_batched", h.value, ra.OP_NONE, n, nrhs, A.data_ptr(),
                 n, n * n, ipiv.data_ptr(), n, B.data_ptr(), n, n * nrhs, bc)
        torch.cuda.synchronize()
        print("dgetrf + dgetrs n=128 done", flush=True)
        return
    if workload == "getrf32":
        g = torch.Generator(device="cuda").manual_seed(11)
        A = torch.rand((BC32, N32, N32), dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        ipiv = torch.zeros((BC32, N32), dtype=torch.int32, device="cuda")
        info = torch.zeros(BC32, dtype=torch.int32, device="cuda")
        h = ra.Handle()
        ra.check("rocsolver_dgetrf_strided_batched", h.value, N32, N32, A.data_ptr(), N32,
                 N32 * N32, ipiv.data_ptr(), N32, info.data_ptr(), BC32)
        torch.cuda.synchronize()
        print("dgetrf n=32 done", flush=True)
        return
    g = torch.Generator(device="cuda").manual_seed(7)
    A = torch.randn((BC, N, N), dtype=torch.float64, device="cuda", generator=g)
    A += 4.0 * N ** 0.5 * torch.eye(N, dtype=torch.float64, device="cuda")
    B = torch.randn((BC, NRHS, N), dtype=torch.float64, device="cuda", generator=g)
    LU = torch.empty((BC, N, N), dtype=torch.float32, device="cuda")
    ipiv = torch.zeros((BC, N), dtype=torch.int32, device="cuda")
    info = torch.zeros(BC, dtype=torch.int32, device="cuda")
    h = ra.Handle()
    ra.check("rocsolver_amd_dsgetrf_strided_batched", h.value, N, A.data_ptr(), N, N * N,
             LU.data_ptr(), N, N * N, ipiv.data_ptr(), N, info.data_ptr(), BC)
    torch.cuda.synchronize()
    print("dsgetrf done", flush=True)
    del LU
    ra.dsgesv(h, A, B)
    torch.cuda.synchronize()
    print("dsgesv done", flush=True)
    # the plain FP64 solve the mixed one is compared with, on copies
    Aw, Bw = A.clone(), B.clone()
    ra.check("rocsolver_dgesv_strided_batched", h.value, N, NRHS, Aw.data_ptr(), N, N * N,
             ipiv.data_ptr(), N, Bw.data_ptr(), N, N * NRHS, info.data_ptr(), BC)
    torch.cuda.synchronize()
    print("dgesv done", flush=True)


def _rows(path: str):
    with open(path) as f:
        return list(csv.DictReader(f))


def _grid(r) -> int:
    g = r.get("Grid_Size") or r.get("Grid_Size_X") or r.get("Grid-Size") or "0"
    return int(float(g))


def parse(fetch_csv: str, write_csv: str) -> dict:
    """Per kernel name and grid size: dispatch count and mean FETCH_SIZE and
    WRITE_SIZE per dispatch in bytes, scaled so that the 1 GiB copy reads and
    writes exactly 1 GiB.  The grid size tells a launch over the whole batch
    from one over part of it (the mixed solve factors and refines the two
    halves of a large batch on two streams)."""
    out = {}
    for key, path in (("fetch", fetch_csv), ("write", write_csv)):
        per = {}
        for r in _rows(path):
            name = r.get("Kernel_Name") or r.get("Kernel-Name") or r.get("KernelName")
            val = float(r.get("Counter_Value") or r.get("Counter-Value") or r.get("CounterValue"))
            per.setdefault((name, _grid(r)), []).append(val * 1024.0)  # counters report KiB
        out[key] = per
    # the calibration copy: the elementwise multiply by 1.0 of the 1 GiB buffer
    copy = [k for k in out["fetch"] if "MulFunctor" in k[0]]
    big = max(copy, key=lambda k: max(out["fetch"][k]))
    f_scale = COPY_BYTES / max(out["fetch"][big])
    w_scale = COPY_BYTES / max(out["write"][big])
    res = {"copy_kernel": big[0], "fetch_scale": f_scale, "write_scale": w_scale, "kernels": []}
    for (name, grid), fe in sorted(out["fetch"].items()):
        if "rsa::" not in name:
            continue
        wr = out["write"].get((name, grid), [0.0])
        res["kernels"].append({
            "kernel": name.split("(")[0],
            "grid": grid,
            "dispatches": len(fe),
            "fetch_bytes": sum(fe) / len(fe) * f_scale,
            "write_bytes": sum(wr) / len(wr) * w_scale,
        })
    return res


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--parse":
        import json
        print(json.dumps(parse(sys.argv[2], sys.argv[3]), indent=1))
    else:
        run(sys.argv[1] if len(sys.argv) > 1 else "mixed")
