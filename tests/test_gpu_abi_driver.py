"""The compiled C++ caller of the C ABI (tests/abi/xkrt_abi_driver.cpp) on the
GPU: a domain fed in batches, a two-domain pool and a coherency tree, driven
through include/xkrt_hip.h with no Python in between.  Its edges equal the
oracle's on the same streams (the driver generates them in the same order as
xkaapi_amd/streams)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import khp
from xkaapi_amd import streams as S

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "abi", "xkrt_abi_driver")


def oracle_set(st):
    r, nr = st.rects()
    e = khp.resolve(st.task, st.mode, r, nr)
    return set(map(tuple, np.asarray(e, np.int64).tolist()))


def program():
    """The driver's mixed task-domain program (tests/abi/xkrt_abi_driver.cpp, program())."""
    rows = []  # task, mode, kind, addr, m, n, ld, s
    base = 1 << 20
    for t in range(160):
        k = t % 64
        i, j = k % 8, (k * 3 + t // 64) % 8
        if t % 5 == 4:
            rows.append((t, 1 if t % 2 else 2, 1, base + (4096 * t) % (512 << 10), 5000, 0, 0, 0))
            continue
        if t % 3 == 0:
            a = (t, 3, 0, base + (j * 32 * 512 + i * 64) * 4, 64, 32, 512, 4)
        else:
            a = (t, 3, 0, base + (j * 32 * 256 + i * 32) * 8, 32, 32, 256, 8)
        rows.append(a)
        if t % 7 == 0:
            rows.append((t, 1, 1, a[3], 256, 32, a[6], a[7]))
    r = np.array(rows, np.uint64)
    return [r[:, 0].astype(np.uint32), r[:, 1].astype(np.uint8), r[:, 2].astype(np.uint8), r[:, 3], r[:, 4],
            r[:, 5], r[:, 6], r[:, 7]]


def test_compiled_driver_matches_oracle():
    assert os.path.exists(EXE), "tests/abi/xkrt_abi_driver is not built: run __graft_entry__.build()"
    print("running", EXE, flush=True)
    sys.stdout.flush()
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = r.stdout.splitlines()
    assert lines[-1] == "ok"
    chol = {(int(p), int(q)) for k, p, q in (l.split() for l in lines if l.startswith("chol "))}
    assert chol == oracle_set(S.cholesky(4096, 256))
    dom = {l.split()[1]: int(l.split()[2]) for l in lines if l.startswith("domain ")}
    pool = {}
    for l in lines:
        if l.startswith("pool "):
            _, d, p, q = l.split()
            pool.setdefault(int(d), set()).add((int(p), int(q)))
    assert pool[dom["cholesky"]] == oracle_set(S.cholesky(1024, 128))
    assert pool[dom["lu"]] == oracle_set(S.lu(1024, 128))
    assert "who_owns_equal 1" in lines
    prog = {(int(p), int(q)) for _, p, q in (l.split() for l in lines if l.startswith("prog "))}
    want = {tuple(r) for r in np.asarray(khp.resolve_program(*program()), np.int64).tolist()}
    assert len(want) > 50 and prog == want
