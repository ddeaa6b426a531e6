"""CPU regression test for the work-queue loop of the persistent
variable-size kernel (rocsolver_amd/csrc/vbatched.hpp, vb_work_kernel).

The kernel loops: read the claimed item from LDS, barrier, exit when the
queue is empty, process the item, thread 0 claims the next item with an
atomic add, barrier.  An earlier form claimed at the top of the loop; for
the getrf instantiation the compiler then merged the divergent
`threadIdx.x == 0` store that ends a large matrix with the claim's branch
and built a nested loop whose inner header held the barriers but not the
atomic.  Lanes 1..63 of wavefront 0 and all other wavefronts re-entered the
barriers without lane 0's claim, read the old item and factored it again
forever (a GPU hang).

This test compiles the three double-precision instantiations to gfx950
assembly (device only) and checks, from the compiler's loop annotations,
that the queue atomic lies in a depth-1 loop whose header block -- before
its first branch -- holds an s_barrier, i.e. every pass over the loop's
barriers goes through the claim.
"""
import concurrent.futures as cf
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "rocsolver_amd", "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")


def _asm(src: str, out: str) -> str:
    from rocsolver_amd import build
    flags = [f for f in build.COMMON if f not in ("-fPIC",)]
    cmd = [HIPCC, *flags, "--cuda-device-only", "-S", src, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    with open(out) as f:
        return f.read()


def _check_queue_loop(asm: str, name: str) -> None:
    lines = asm.splitlines()
    # innermost loop of every basic block, from the annotations
    loop_of = {}
    cur = None
    header_line = {}
    for i, l in enumerate(lines):
        m = re.match(r"^(\.LBB\d+_\d+|; %bb\.\d+):", l)
        if m:
            t = re.search(r"This (?:Inner )?Loop Header: Depth=(\d+)", l)
            h = re.search(r"Header=(BB\d+_\d+) Depth=(\d+)", l)
            if t:
                cur = (m.group(1).lstrip(".L"), int(t.group(1)))
                header_line[cur[0]] = i
            elif h:
                cur = (h.group(1), int(h.group(2)))
            else:
                cur = None
        loop_of[i] = cur
    atomics = [i for i, l in enumerate(lines)
               if l.strip().startswith("global_atomic_add") and loop_of[i] is not None]
    assert len(atomics) == 1, f"{name}: expected one in-loop queue atomic, found {len(atomics)}"
    hdr, depth = loop_of[atomics[0]]
    assert depth == 1, f"{name}: the queue claim sits in a depth-{depth} loop"
    # header block: from the header label to its first branch
    i = header_line[hdr] + 1
    saw_barrier = False
    while i < len(lines):
        s = lines[i].strip()
        if s == "s_barrier":
            saw_barrier = True
        if s.startswith("s_cbranch") or s.startswith("s_branch") or s.startswith("s_setpc"):
            break
        i += 1
    assert saw_barrier, f"{name}: the queue loop's header block has no barrier before its exit test"


@pytest.mark.timeout(900)
def test_queue_claim_and_barriers_share_one_loop(tmp_path):
    ops = ["getrf", "getrs", "gesv"]
    srcs = {op: os.path.join(CSRC, f"vbatched_{op}_double.hip") for op in ops}
    with cf.ThreadPoolExecutor(len(ops)) as ex:
        futs = {op: ex.submit(_asm, srcs[op], str(tmp_path / f"{op}.s")) for op in ops}
        asms = {op: f.result() for op, f in futs.items()}
    for op in ops:
        _check_queue_loop(asms[op], op)
