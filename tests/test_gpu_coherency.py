"""GPU parity of the coherency fetch planner (libxkrt_hip.so through its C ABI)
against the oracle's restatement of the reference memory tree."""
import numpy as np
import pytest

from oracle import khp
from xkaapi_amd import coherency as K
from xkaapi_amd import streams as S
from xkaapi_amd.resolver import ResolverError

pytestmark = pytest.mark.gpu


def oracle_plan(st, queries=None):
    r, nr = st.rects()
    return khp.coherency(st.dev, st.mode, r, nr, st.m * st.n * st.s, queries=queries)


def assert_same_plan(got, want, n):
    assert got.fail_at == want.fail_at
    last = want.fail_at if want.fail_at >= 0 else n
    assert np.array_equal(got.part_off[:last + 1], want.part_off[:last + 1])
    assert np.array_equal(got.fetch_off[:last + 1], want.fetch_off[:last + 1])
    assert np.array_equal(got.parts[:want.part_off[last]], want.parts[:want.part_off[last]])
    assert np.array_equal(got.fetches[:want.fetch_off[last]], want.fetches[:want.fetch_off[last]])


def tiny(seed, n=300, ld=12, cols=14, s=2, maxm=10, maxn=8, phost=0.5):
    rng = np.random.default_rng(seed)
    mode = rng.choice(np.array([1, 2, 3], np.uint8), n, p=[0.5, 0.25, 0.25])
    m = rng.integers(1, maxm + 1, n).astype(np.uint64)
    nn = rng.integers(1, maxn + 1, n).astype(np.uint64)
    row = rng.integers(0, ld, n).astype(np.uint64)
    col = rng.integers(0, cols, n).astype(np.uint64)
    m = np.minimum(m, ld - row)
    nn = np.minimum(nn, cols - col)
    dev = (rng.random(n) >= phost).astype(np.uint8)
    return S.Stream(np.arange(n, dtype=np.uint32), np.zeros(n, np.uint16), mode, dev, (col * ld + row) * s,
                    np.full(n, ld, np.uint64), m, nn, np.full(n, s, np.uint64))


def test_known_answers():
    # one GPU: R then W on the same rect, then a host R -> H2D, nothing, D2H
    st = S.Stream(np.arange(3, dtype=np.uint32), np.zeros(3, np.uint16), np.array([1, 2, 1], np.uint8),
                  np.array([1, 1, 0], np.uint8), np.zeros(3, np.uint64), np.full(3, 64, np.uint64),
                  np.full(3, 8, np.uint64), np.full(3, 4, np.uint64), np.full(3, 8, np.uint64))
    got = K.plan_stream(st, (0, 4))
    assert got.fetches_of(0).tolist() == [[0, 64, 0, 4, 0, -1, 1, 0]]
    assert got.fetches_of(1).tolist() == []
    assert got.fetches_of(2).tolist() == [[0, 64, 0, 4, 1, 0, 0, -1]]
    # host-only: the tree stays empty, nothing is planned
    st.dev[:] = 0
    got = K.plan_stream(st, (0, 4))
    assert len(got.parts) == 0 and len(got.fetches) == 0


@pytest.mark.parametrize("seed", range(6))
def test_random_tiny_matches_oracle(seed):
    st = tiny(seed)
    assert_same_plan(K.plan_stream(st, (0, 14)), oracle_plan(st), len(st))


@pytest.mark.parametrize("seed", range(3))
def test_tall_accesses_span_row_blocks(seed):
    # columns of fragments taller than one work item (1024 rows): pieces, and
    # empty row blocks, on both sides of item boundaries
    st = tiny(70 + seed, n=80, ld=12, cols=3000, maxn=3000)
    assert int(st.n.max()) > 2048
    assert_same_plan(K.plan_stream(st, (0, 3000)), oracle_plan(st), len(st))


@pytest.mark.parametrize("seed", range(3))
def test_batches_match_one_call(seed):
    st = tiny(50 + seed, n=400, ld=40, cols=30, maxm=20, maxn=12)
    want = oracle_plan(st)
    for b in (1, 37, 400):
        assert_same_plan(K.plan_stream(st, (0, 30), batch=b), want, len(st))


def test_tile_edges_and_offset_frame():
    # 70 x 70 elements: several 32 x 32 state tiles, accesses crossing tile edges, a frame not at column 0
    st = tiny(7, n=600, ld=70, cols=70, s=4, maxm=50, maxn=50, phost=0.3)
    st.addr = st.addr + np.uint64(5 * 70 * 4)  # columns [5, 75)
    assert_same_plan(K.plan_stream(st, (5, 75)), oracle_plan(st), len(st))


def test_ninth_view_abort_matches_oracle():
    # nested GPU writes on growing rects stack allocation views until the reference aborts
    n = 12
    m = np.arange(1, n + 1, dtype=np.uint64)
    st = S.Stream(np.arange(n, dtype=np.uint32), np.zeros(n, np.uint16), np.full(n, 2, np.uint8),
                  np.ones(n, np.uint8), np.zeros(n, np.uint64), np.full(n, 16, np.uint64), m, m,
                  np.full(n, 8, np.uint64))
    want = oracle_plan(st)
    assert want.fail_at == 8
    got = K.plan_stream(st, (0, 16))
    assert_same_plan(got, want, n)


def test_who_owns_matches_oracle():
    st = tiny(11, n=100)
    q = tiny(12, n=80)
    qr, qn = q.rects()
    want = oracle_plan(st, queries=(qr, qn))
    assert want.fail_at == -1
    acc = K.to_mcc_array(st.dev, st.mode, st.addr, st.m, st.n)
    with K.MemoryTree(12, 2, (0, 14)) as t:
        t.plan(acc)
        got = t.who_owns(q.addr, q.m, q.n)
    assert np.array_equal(got, want.owners)


def test_who_owns_refused_after_abort():
    n = 12
    m = np.arange(1, n + 1, dtype=np.uint64)
    acc = K.to_mcc_array(np.ones(n, np.uint8), np.full(n, 2, np.uint8), np.zeros(n, np.uint64), m, m)
    with K.MemoryTree(16, 8, (0, 16)) as t:
        assert t.plan_counts(acc)[2] == 8
        with pytest.raises(ResolverError) as ei:
            t.who_owns(acc["addr"][:1], acc["m"][:1], acc["n"][:1])
        assert ei.value.rc == -125  # -ECANCELED


def test_two_view_stream_prefix():
    # the mixed-view coherency stream: one tree per (ld, sizeof_type), cut where the reference aborts
    st = S.coherency(6000)
    for ld, s in ((8192, 8), (16384, 4)):
        sel = np.nonzero((st.ld == ld) & (st.s == s))[0]
        sub = st.take(sel)
        want = oracle_plan(sub)
        got = K.plan_stream(sub, (0, 8192), batch=1024)
        assert_same_plan(got, want, len(sub))
