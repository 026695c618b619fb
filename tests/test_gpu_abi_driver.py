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
