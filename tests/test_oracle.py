"""Pins the oracle: the C restatement (oracle/khp_oracle.c) against the golden
fixtures written by the reference's own KHP engine, against the reference
engine itself when it is built, and against brute-force per-point semantics."""
import os

import numpy as np
import pytest

from oracle import khp, pointwise, refkhp
from xkaapi_amd import streams as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RESOLVER_FIXTURES = ["cholesky_t2_edges.npz", "cholesky_4096_nb256.npz", "lu_1024_nb128.npz",
                     "heat_1024_t128_s4.npz", "random2d_4000.npz", "wrap_mixed.npz"]


def load_stream(z):
    n = len(z["task"])
    dev = z["dev"] if "dev" in z.files else np.ones(n, np.uint8)
    slot = z["slot"] if "slot" in z.files else np.zeros(n, np.uint16)
    task = z["task"] if "task" in z.files else np.arange(n, dtype=np.uint32)
    return S.Stream(task, slot, z["mode"], dev, z["addr"], z["ld"], z["m"], z["n"], z["s"])


def edge_rows(st, e):
    e = np.asarray(e, np.int64).reshape(-1, 2)
    rows = np.stack([st.task[e[:, 0]].astype(np.int64), st.slot[e[:, 0]].astype(np.int64),
                     st.task[e[:, 1]].astype(np.int64), st.slot[e[:, 1]].astype(np.int64)], axis=1)
    return rows[np.lexsort(rows.T[::-1])] if len(rows) else rows.reshape(0, 4)


def tiny_stream(n, ld, cols, s, maxm, maxn, seed, modes=(1, 2, 3, 4, 5, 6, 8, 9), per_task=3, wrap=False):
    rng = np.random.default_rng(seed)
    k = rng.integers(1, per_task + 1, n)
    task = np.repeat(np.arange(n, dtype=np.uint32), k)[:n]
    slot = np.zeros(n, np.uint16)
    for i in range(1, n):
        slot[i] = slot[i - 1] + 1 if task[i] == task[i - 1] else 0
    mode = rng.choice(np.array(modes, np.uint8), n)
    m = rng.integers(1, maxm + 1, n).astype(np.uint64)
    nn = rng.integers(1, maxn + 1, n).astype(np.uint64)
    row = rng.integers(0, ld, n).astype(np.uint64)
    if not wrap:
        m = np.minimum(m, ld - row)
    col = rng.integers(0, cols, n).astype(np.uint64)
    return S.Stream(task, slot, mode, rng.integers(0, 2, n).astype(np.uint8), (col * ld + row) * s,
                    np.full(n, ld, np.uint64), m, nn, np.full(n, s, np.uint64))


def test_cholesky_t2_known_answer():
    # POTRF(0) -> TRSM(1) slot 0, TRSM(1) slot 1 -> SYRK(2) slot 0, SYRK(2) slot 1 -> POTRF(3) slot 0
    st = S.cholesky(512, 256)
    r, nr = st.rects()
    got = edge_rows(st, khp.resolve(st.task, st.mode, r, nr))
    assert got.tolist() == [[0, 0, 1, 0], [1, 1, 2, 0], [2, 1, 3, 0]]


@pytest.mark.parametrize("name", RESOLVER_FIXTURES)
def test_oracle_matches_reference_golden(name):
    z = np.load(os.path.join(GOLDEN, name))
    st = load_stream(z)
    r, nr = st.rects()
    e, f = khp.resolve(st.task, st.mode, r, nr, want_frags=True)
    assert np.array_equal(edge_rows(st, e), z["edges"])
    assert np.array_equal(f.astype(np.int64), z["frag"])


def test_golden_streams_are_the_generators():
    # the committed streams are what the generators produce today
    z = np.load(os.path.join(GOLDEN, "cholesky_4096_nb256.npz"))
    st = S.cholesky(4096, 256)
    for f in ("task", "slot", "mode", "addr", "ld", "m", "n", "s"):
        assert np.array_equal(getattr(st, f), z[f]), f
    z = np.load(os.path.join(GOLDEN, "random2d_4000.npz"))
    st = S.random2d(4000)
    for f in ("task", "mode", "addr", "m", "n"):
        assert np.array_equal(getattr(st, f), z[f]), f


def test_partition_matches_reference_golden():
    z = np.load(os.path.join(GOLDEN, "partition_coherency.npz"))
    n = len(z["dev"])
    st = S.Stream(np.arange(n, dtype=np.uint32), np.zeros(n, np.uint16), z["mode"], z["dev"], z["addr"],
                  z["ld"], z["m"], z["n"], z["s"])
    r, nr = st.rects()
    off, parts = khp.partition(st.dev, r, nr)
    assert np.array_equal(off, z["part_off"])
    for i in range(n):
        a = parts[off[i]:off[i + 1]].astype(np.int64)
        b = z["parts"][off[i]:off[i + 1]]
        assert np.array_equal(a[np.lexsort(a.T[::-1])], b[np.lexsort(b.T[::-1])])


@pytest.mark.skipif(not refkhp.available(), reason="reference engine not built")
@pytest.mark.parametrize("mk", [lambda: S.lu(2048, 128), lambda: S.heat(2048, 128, 4, 3),
                                lambda: S.random2d(3000, rows=256, cols=256, maxsz=64, seed=2)])
def test_oracle_matches_reference_engine(mk):
    st = mk()
    r, nr = st.rects()
    e1, f1 = refkhp.resolve(st.task, st.mode, r, nr, want_frags=True)
    e2, f2 = khp.resolve(st.task, st.mode, r, nr, want_frags=True)
    assert np.array_equal(e1, e2)
    assert np.array_equal(f1, f2)


@pytest.mark.parametrize("seed", range(8))
def test_oracle_matches_pointwise_single_rect(seed):
    st = tiny_stream(250, 12, 16, 2, 8, 6, seed)
    r, nr = st.rects()
    assert (nr == 1).all()
    got = set(map(tuple, khp.resolve(st.task, st.mode, r, nr).tolist()))
    assert got == pointwise.resolve(st.task, st.mode, r, nr, 24, 24)


def test_multirect_overshoot_quirk():
    # W [0,8)x[0,3); R of a wrapped tile (addr 5, m 6, n 1: [5,8)x[0,1) + [0,3)x[1,2)); W [3,5)x[0,1).
    # The reference marks the whole split-off fragment [0,5)x[0,3) with the reader (it touches the
    # second rect), so the last writer depends on the reader, not on the first writer.
    st = S.Stream(np.array([0, 1, 2], np.uint32), np.zeros(3, np.uint16), np.array([2, 1, 2], np.uint8),
                  np.ones(3, np.uint8), np.array([0, 5, 3], np.uint64), np.full(3, 8, np.uint64),
                  np.array([8, 6, 2], np.uint64), np.array([3, 1, 1], np.uint64), np.ones(3, np.uint64))
    r, nr = st.rects()
    assert khp.resolve(st.task, st.mode, r, nr).tolist() == [[0, 1], [1, 2]]
    if refkhp.available():
        assert refkhp.resolve(st.task, st.mode, r, nr).tolist() == [[0, 1], [1, 2]]
    assert pointwise.resolve(st.task, st.mode, r, nr, 8, 4) == {(0, 1), (0, 2)}


def test_planted_mistake_is_caught():
    # An oracle that forgot the WAR rule (writers link only to the last writer) must fail the golden
    # (the heat stream has write-after-read on every halo strip).
    z = np.load(os.path.join(GOLDEN, "heat_1024_t128_s4.npz"))
    st = load_stream(z)
    r, nr = st.rects()
    bad_mode = np.where(st.mode == 1, 0, st.mode).astype(np.uint8)  # readers become invisible
    e = khp.resolve(st.task, bad_mode, r, nr)
    assert not np.array_equal(edge_rows(st, e), z["edges"])


@pytest.mark.parametrize("seed", range(6))
def test_coherency_matches_pointwise(seed):
    rng = np.random.default_rng(seed)
    n, ld, cols, s = 300, 12, 14, 2
    mode = rng.choice(np.array([1, 2, 3], np.uint8), n, p=[0.5, 0.25, 0.25])
    m = rng.integers(1, 11, n).astype(np.uint64)
    nn = rng.integers(1, 9, n).astype(np.uint64)
    row = rng.integers(0, ld, n).astype(np.uint64)
    col = rng.integers(0, cols, n).astype(np.uint64)
    m = np.minimum(m, ld - row)
    nn = np.minimum(nn, cols - col)
    dev = rng.integers(0, 2, n).astype(np.uint8)
    st = S.Stream(np.arange(n, dtype=np.uint32), np.zeros(n, np.uint16), mode, dev, (col * ld + row) * s,
                  np.full(n, ld, np.uint64), m, nn, np.full(n, s, np.uint64))
    r, nr = st.rects()
    cb = st.m * st.n * st.s
    plan = khp.coherency(st.dev, st.mode, r, nr, cb)
    out, fail = pointwise.coherency(st.dev, st.mode, r, nr, cb, 24, 16)
    assert plan.fail_at == fail
    last = fail if fail >= 0 else n
    for i in range(last):
        before, fetch = out[i]
        got_b, got_f = {}, {}
        for p in plan.parts_of(i):
            for x in range(p[0], p[1]):
                for y in range(p[2], p[3]):
                    assert (x, y) not in got_b
                    got_b[(x, y)] = int(p[4])
        for f in plan.fetches_of(i):
            for x in range(f[0], f[1]):
                for y in range(f[2], f[3]):
                    got_f[(x, y)] = tuple(int(v) for v in f[4:8])
        assert got_b == before, i
        assert got_f == fetch, i


def test_coherency_known_answers():
    # single GPU, R then W on the same rect, then a host R: H2D, no fetch, D2H
    st = S.Stream(np.arange(3, dtype=np.uint32), np.zeros(3, np.uint16), np.array([1, 2, 1], np.uint8),
                  np.array([1, 1, 0], np.uint8), np.zeros(3, np.uint64), np.full(3, 64, np.uint64),
                  np.full(3, 8, np.uint64), np.full(3, 4, np.uint64), np.full(3, 8, np.uint64))
    r, nr = st.rects()
    plan = khp.coherency(st.dev, st.mode, r, nr, st.m * st.n * st.s)
    assert plan.fetches_of(0).tolist() == [[0, 64, 0, 4, 0, -1, 1, 0]]
    assert plan.fetches_of(1).tolist() == []
    assert plan.fetches_of(2).tolist() == [[0, 64, 0, 4, 1, 0, 0, -1]]
    # host-only stream: nothing is ever planned
    st2 = st.take(slice(0, 3))
    st2.dev[:] = 0
    plan = khp.coherency(st2.dev, st2.mode, r, nr, st2.m * st2.n * st2.s)
    assert len(plan.fetches) == 0 and len(plan.parts) == 0


@pytest.mark.parametrize("seed", range(4))
def test_who_owns_matches_pointwise(seed):
    # memory-tree.hpp:1903-1933 on the final state: bytes per coherent device, the maxima win
    rng = np.random.default_rng(100 + seed)
    n, ld, cols, s = 200, 12, 14, 2
    mode = rng.choice(np.array([1, 2, 3], np.uint8), n, p=[0.5, 0.25, 0.25])
    m = rng.integers(1, 11, n).astype(np.uint64)
    nn = rng.integers(1, 9, n).astype(np.uint64)
    row = rng.integers(0, ld, n).astype(np.uint64)
    col = rng.integers(0, cols, n).astype(np.uint64)
    m = np.minimum(m, ld - row)
    nn = np.minimum(nn, cols - col)
    dev = rng.integers(0, 2, n).astype(np.uint8)
    st = S.Stream(np.arange(n, dtype=np.uint32), np.zeros(n, np.uint16), mode, dev, (col * ld + row) * s,
                  np.full(n, ld, np.uint64), m, nn, np.full(n, s, np.uint64))
    r, nr = st.rects()
    cb = st.m * st.n * st.s
    q = st.take(slice(0, 60))
    qr, qn = q.rects()
    plan = khp.coherency(st.dev, st.mode, r, nr, cb, queries=(qr, qn))
    _, fail, coh = pointwise.coherency(st.dev, st.mode, r, nr, cb, 24, 16, want_state=True)
    assert plan.fail_at == fail
    for k in range(len(qn)):
        by = [0, 0]
        for (x0, x1, y0, y1) in [tuple(int(v) for v in qr[k, 4 * j:4 * j + 4]) for j in range(qn[k])]:
            for x in range(x0, x1):
                for y in range(y0, y1):
                    for d in range(2):
                        by[d] += (int(coh[x, y]) >> d) & 1
        best = max(by)
        want = 0 if best == 0 else sum(1 << d for d in range(2) if by[d] == best)
        assert int(plan.owners[k]) == want, k

