"""GPU parity of the dependency resolver (libxkrt_hip.so through its C ABI)
against the golden fixtures of the reference engine and against the oracle."""
import os

import numpy as np
import pytest

from oracle import khp
from xkaapi_amd import resolver as R
from xkaapi_amd import streams as S

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rows(st, e):
    e = np.asarray(e, np.int64).reshape(-1, 2)
    r = np.stack([st.task[e[:, 0]].astype(np.int64), st.slot[e[:, 0]].astype(np.int64),
                  st.task[e[:, 1]].astype(np.int64), st.slot[e[:, 1]].astype(np.int64)], axis=1)
    return r[np.lexsort(r.T[::-1])] if len(r) else r.reshape(0, 4)


def oracle_edges(st):
    r, nr = st.rects()
    return khp.resolve(st.task, st.mode, r, nr)


def same(a, b):
    # one edge per (pred access, succ task) (include/xkrt/task/task.hpp:494-512): no row twice
    a = np.asarray(a, np.int64).reshape(-1, 2)
    ua = np.unique(a, axis=0)
    assert len(ua) == len(a), "duplicate edges"
    ub = np.unique(np.asarray(b, np.int64).reshape(-1, 2), axis=0)
    return ua.shape == ub.shape and np.array_equal(ua, ub)


def load_stream(z):
    n = len(z["task"])
    return S.Stream(z["task"], z["slot"], z["mode"], np.ones(n, np.uint8), z["addr"], z["ld"], z["m"], z["n"], z["s"])


def test_cholesky_t2_known_answer():
    st = S.cholesky(512, 256)
    e = R.resolve_stream(st, tasks_per_batch=1, bin_x_bytes=2048, bin_y_cols=256)
    assert rows(st, e).tolist() == [[0, 0, 1, 0], [1, 1, 2, 0], [2, 1, 3, 0]]


@pytest.mark.parametrize("name,bx,by,tpb", [
    ("cholesky_4096_nb256.npz", 2048, 256, 4096),
    ("cholesky_4096_nb256.npz", 512, 64, 37),
    ("lu_1024_nb128.npz", 1024, 128, 50),
    ("heat_1024_t128_s4.npz", 512, 128, 100),
    ("heat_1024_t128_s4.npz", 128, 32, 7),
    ("random2d_4000.npz", 1024, 256, 4096),
    ("random2d_4000.npz", 4096, 512, 333),
])
def test_golden_fixture(name, bx, by, tpb):
    z = np.load(os.path.join(GOLDEN, name))
    st = load_stream(z)
    e = R.resolve_stream(st, tasks_per_batch=tpb, bin_x_bytes=bx, bin_y_cols=by, kernel=1)
    assert np.array_equal(rows(st, e), z["edges"])


def tiny_stream(n, ld, cols, s, maxm, maxn, seed, modes=(1, 2, 3, 4, 5, 6, 8, 9), per_task=3):
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
    m = np.minimum(m, ld - row)
    col = rng.integers(0, cols, n).astype(np.uint64)
    return S.Stream(task, slot, mode, np.ones(n, np.uint8), (col * ld + row) * s, np.full(n, ld, np.uint64),
                    m, nn, np.full(n, s, np.uint64))


@pytest.mark.parametrize("seed", range(4))
def test_random_small_space_all_modes(seed):
    st = tiny_stream(3000, 64, 64, 4, 40, 40, seed)
    want = oracle_edges(st)
    for bx, by, tpb, kc in ((256, 64, 400, 2), (64, 16, 300, 2), (16, 4, 97, 1), (64, 16, 1, 1)):
        got = R.resolve_stream(st, tasks_per_batch=tpb, bin_x_bytes=bx, bin_y_cols=by, kernel=kc)
        assert same(got, want), (bx, by, tpb)


def test_cholesky_8192_nb128():
    st = S.cholesky(8192, 128)
    want = oracle_edges(st)
    got = R.resolve_stream(st, tasks_per_batch=4096, bin_x_bytes=1024, bin_y_cols=128, kernel=0)
    assert len(want) == 131040
    assert same(got, want)


def test_random2d_prefix_10k():
    st = S.random2d(10000)
    want = oracle_edges(st)
    got = R.resolve_stream(st, tasks_per_batch=4096, bin_x_bytes=4096, bin_y_cols=512, kernel=2)
    assert same(got, want)


def test_heat_2048():
    st = S.heat(2048, 128, 4, 6)
    want = oracle_edges(st)
    bs = S.geometric_batches(st.ntasks, mean=200.0, maxb=2000, seed=3)
    got = R.resolve_stream(st, tasks_per_batch=bs, bin_x_bytes=512, bin_y_cols=128, kernel=0)
    assert same(got, want)


def test_wrapped_tile_refused_when_reference_would_overshoot():
    # a full-column write, then a wrapped tile: the reference marks beyond the tile -> refused
    acc = R.to_access_array(np.array([0, 1], np.uint32), np.array([2, 1], np.uint8),
                            np.array([0, 5], np.uint64), np.array([8, 6], np.uint64), np.array([3, 1], np.uint64))
    with R.DependencyDomain(8, 1, bin_x_bytes=8, bin_y_cols=4) as d:
        with pytest.raises(R.ResolverError) as ei:
            d.submit(acc)
        assert ei.value.rc == -95  # -ENOTSUP


def test_wrapped_tile_exact_when_no_overshoot():
    # wrapped tiles with only narrow earlier rects resolve exactly like the reference
    rng = np.random.default_rng(5)
    n, ld, s = 400, 64, 1
    task = np.arange(n, dtype=np.uint32)
    mode = rng.choice(np.array([1, 2, 3], np.uint8), n)
    m = rng.integers(1, 5, n).astype(np.uint64)
    nn = rng.integers(1, 4, n).astype(np.uint64)
    col = rng.integers(0, 16, n).astype(np.uint64)
    row = rng.integers(0, ld, n).astype(np.uint64)
    st = S.Stream(task, np.zeros(n, np.uint16), mode, np.ones(n, np.uint8), (col * ld + row) * s,
                  np.full(n, ld, np.uint64), m, nn, np.full(n, s, np.uint64))
    r, nr = st.rects()
    assert (nr == 2).any()
    got = R.resolve_stream(st, tasks_per_batch=50, bin_x_bytes=16, bin_y_cols=4, kernel=1)
    assert same(got, oracle_edges(st))


def _stream(rows_cols_mn_mode, ld, s):
    """One access per task: (row, col, m, n, mode) in elements of an (ld, s) matrix."""
    a = np.asarray(rows_cols_mn_mode, np.int64)
    n = len(a)
    return S.Stream(np.arange(n, dtype=np.uint32), np.zeros(n, np.uint16), a[:, 4].astype(np.uint8),
                    np.ones(n, np.uint8), ((a[:, 1] * ld + a[:, 0]) * s).astype(np.uint64),
                    np.full(n, ld, np.uint64), a[:, 2].astype(np.uint64), a[:, 3].astype(np.uint64),
                    np.full(n, s, np.uint64))


def test_register_replay_overflow_drops_its_unique_edges():
    # One bin (256 x 128 elements of 4 B = 1024 B x 128 columns), one launch: a writer, readers
    # inside it (single-bin predecessor, one record per task: their edges are known unique and
    # staged as such), then 80 disjoint small writers that push the map past 64 fragments, so
    # the register replay gives the bin up and LDS tier 2 replays it; then more readers.  The
    # edges staged before the overflow must be dropped, or they would be found twice.
    acc = [(0, 0, 64, 32, 2)]
    acc += [(4 * k, k, 8, 8, 1) for k in range(10)]
    acc += [(16 * (k % 16), 40 + 8 * (k // 16), 4, 4, 2) for k in range(80)]
    acc += [(16 * (k % 16), 40 + 8 * (k // 16), 8, 6, 1) for k in range(0, 80, 3)]
    acc += [(0, 0, 128, 128, 3)]
    st = _stream(acc, 256, 4)
    want = oracle_edges(st)
    got = R.resolve_stream(st, tasks_per_batch=len(acc), bin_x_bytes=1024, bin_y_cols=128, kernel=1)
    assert same(got, want)


def test_unique_edges_mixed_with_shared_keys():
    # multi-record tasks (two accesses of one task in one bin) and rects spanning several bins
    # beside single-bin ones: edges known unique and table candidates in one launch
    # (launches of at most 400 tasks, capacity class 2: a bin's records fit its 1024-record bucket)
    rng = np.random.default_rng(21)
    n, ld, s = 2000, 512, 4
    task = np.repeat(np.arange(n // 2, dtype=np.uint32), 2)
    mode = rng.choice(np.array([1, 1, 2, 3], np.uint8), n)
    big = rng.random(n) < 0.1
    m = np.where(big, rng.integers(100, 300, n), rng.integers(1, 40, n)).astype(np.uint64)
    nn = np.where(big, rng.integers(50, 200, n), rng.integers(1, 40, n)).astype(np.uint64)
    row = rng.integers(0, ld - 300, n).astype(np.uint64)
    col = rng.integers(0, 400, n).astype(np.uint64)
    st = S.Stream(task, np.tile(np.array([0, 1], np.uint16), n // 2), mode, np.ones(n, np.uint8),
                  (col * ld + row) * s, np.full(n, ld, np.uint64), m, nn, np.full(n, s, np.uint64))
    want = oracle_edges(st)
    for tpb in (64, 400):
        got = R.resolve_stream(st, tasks_per_batch=tpb, bin_x_bytes=512, bin_y_cols=64, kernel=2)
        assert same(got, want), tpb
