"""Pools: many dependency domains resolved together (one launch over the
batches of several clients), against each client resolved alone and against
the oracle.  Through the C ABI."""
import numpy as np
import pytest

from oracle import khp
from xkaapi_amd import resolver as R
from xkaapi_amd import streams as S

pytestmark = pytest.mark.gpu


def unique_rows(e):
    # sorted rows; a row twice would break one edge per (pred access, succ task)
    e = np.asarray(e, np.int64).reshape(-1, 2)
    u = np.unique(e, axis=0) if len(e) else e
    assert len(u) == len(e), "duplicate edges"
    return u


def oracle_edges(st):
    r, nr = st.rects()
    return unique_rows(khp.resolve(st.task, st.mode, r, nr))


def clients():
    return [("cholesky", S.cholesky(1024, 128)), ("lu", S.lu(1024, 128)), ("heat", S.heat(1024, 128, 4, 3)),
            ("random", S.random2d(3000, rows=4096, cols=4096, maxsz=512, seed=9)), ("cholesky2", S.cholesky(2048, 256))]


def col_range(st):
    ld, s = int(st.ld[0]), int(st.s[0])
    cols = (st.addr // (ld * s)).astype(np.int64)
    return int(cols.min()), int((cols + st.n.astype(np.int64)).max()) + 1


def run_pool(streams, tasks_per_batch, graph=False, kernel=1, **pool_kw):
    """Stage the streams' batches interleaved, one launch per round; edges per client."""
    out = {}
    with R.Pool(kernel=kernel, graph=graph, max_rects=1 << 16, **pool_kw) as pool:
        doms, batches, firsts = {}, {}, {}
        for name, st in streams:
            doms[name] = pool.domain(int(st.ld[0]), int(st.s[0]), col_range(st), bin_x_bytes=1024, bin_y_cols=128)
            batches[name] = R.batches_by_tasks(st.task, tasks_per_batch)
            firsts[name] = None
        acc = {name: R.to_access_array(st.task, st.mode, st.addr, st.m, st.n) for name, st in streams}
        idx = {name: doms[name].index for name, _ in streams}
        got = {name: [] for name, _ in streams}
        rnd = 0
        while any(rnd < len(b) for b in batches.values()):
            for name, _ in streams:
                if rnd < len(batches[name]):
                    a, b = batches[name][rnd]
                    f = pool.stage(doms[name], acc[name][a:b])
                    if firsts[name] is None:
                        firsts[name] = f
            pool.launch()
            rnd += 1
        e, d = pool.edges()
        for name, _ in streams:
            got[name] = e[d == idx[name]].astype(np.int64) - (firsts[name] or 0)
        stats = pool.stats()
    for name, _ in streams:
        out[name] = unique_rows(got[name])
    return out, stats


def test_pool_mix_matches_alone_and_oracle():
    streams = clients()
    mixed, stats = run_pool(streams, tasks_per_batch=40)
    assert stats["domains"] == 0 or stats["domains"] == len(streams)
    for name, st in streams:
        alone = unique_rows(R.resolve_stream(st, tasks_per_batch=40, bin_x_bytes=1024, bin_y_cols=128, kernel=1))
        want = oracle_edges(st)
        assert np.array_equal(alone, want), name
        assert np.array_equal(mixed[name], want), name


def test_pool_graph_replay_matches_oracle():
    streams = clients()[:3]
    mixed, stats = run_pool(streams, tasks_per_batch=25, graph=True)
    assert stats["graph"] == 1
    for name, st in streams:
        assert np.array_equal(mixed[name], oracle_edges(st)), name


@pytest.mark.parametrize("kernel,direct", [(0, True), (1, True), (1, False)])
def test_pool_small_launches_match_oracle(kernel, direct):
    # records read in place from pinned host memory, resolve from tier 1, both dedup forms;
    # the rounds of large batches exceed the small-launch limit and take the upload path
    streams = clients()[:4]
    for tpb in (12, 300):
        mixed, _ = run_pool(streams, tasks_per_batch=tpb, kernel=kernel, direct_dedup=direct,
                            small_launch_rects=2048)
        for name, st in streams:
            assert np.array_equal(mixed[name], oracle_edges(st)), (name, tpb)


def test_pool_direct_dedup_graph_matches_oracle():
    streams = clients()[:3]
    mixed, _ = run_pool(streams, tasks_per_batch=25, graph=True, direct_dedup=True)
    for name, st in streams:
        assert np.array_equal(mixed[name], oracle_edges(st)), name


def test_pool_admission_by_budget():
    st = S.cholesky(1024, 128)
    cr = col_range(st)
    probe = R.Pool(kernel=0, max_rects=1 << 12)
    d0 = probe.domain(1024, 8, cr)
    per_dom = probe.stats()["bins_used"] * probe.stats()["bytes_per_bin"]
    base = probe.stats()["device_bytes"] - probe.stats()["bins_capacity"] * probe.stats()["bytes_per_bin"]
    probe.leave(d0)
    probe.close()
    budget = base + 3 * per_dom  # room for exactly three such domains
    with R.Pool(kernel=0, max_rects=1 << 12, budget_bytes=budget) as pool:
        ds = [pool.domain(1024, 8, cr) for _ in range(3)]
        with pytest.raises(R.ResolverError) as ei:
            pool.domain(1024, 8, cr)
        assert ei.value.rc == -12  # -ENOMEM: not admitted
        pool.leave(ds[1])
        d3 = pool.domain(1024, 8, cr)  # admitted once a client has left
        acc = R.to_access_array(st.task, st.mode, st.addr, st.m, st.n)
        f = pool.stage(d3, acc)
        pool.launch()
        e, d = pool.edges()
        got = unique_rows(e[d == d3.index].astype(np.int64) - f)
        assert np.array_equal(got, oracle_edges(st))


def test_pool_domain_reuse_after_leave():
    # a domain index freed by one client and taken by the next must not see the old state
    st = S.cholesky(1024, 128)
    acc = R.to_access_array(st.task, st.mode, st.addr, st.m, st.n)
    with R.Pool(kernel=0, max_rects=1 << 12) as pool:
        for _ in range(3):
            d = pool.domain(1024, 8, col_range(st))
            f = pool.stage(d, acc)
            pool.launch()
            e, dd = pool.edges()
            assert np.array_equal(unique_rows(e[dd == d.index].astype(np.int64) - f), oracle_edges(st))
            pool.leave(d)


def test_pool_stage_room_counts_rects():
    # an open launch holds max_rects rects; a tile takes one, a column-wrapped tile two
    ld, s = 64, 8
    ntile = 6
    tiles = R.to_access_array(np.arange(ntile), np.full(ntile, 2), (2 + 4 * np.arange(ntile)) * ld * s,
                              np.full(ntile, 8), np.full(ntile, 2))
    # a tile of 8 rows starting at row 60 wraps into the next column: two rects
    wrapped = R.to_access_array(np.arange(4), np.full(4, 1), (40 + 4 * np.arange(4)) * ld * s + 60 * s,
                                np.full(4, 8), np.full(4, 1))
    with R.Pool(kernel=0, max_rects=8) as pool:
        d = pool.domain(ld, s, (0, 64))
        pool.stage(d, tiles)             # 6 rects
        with pytest.raises(R.ResolverError) as ei:
            pool.stage(d, wrapped[:2])   # 4 more: the open launch is full
        assert ei.value.rc == -28
        pool.launch()
        with pytest.raises(R.ResolverError) as ei:
            pool.stage(d, np.concatenate([wrapped, wrapped[:1]]))
        assert ei.value.rc == -22        # 10 rects never fit one launch
        pool.stage(d, wrapped)           # 8 rects: exactly full
        pool.launch()
        e, dd = pool.edges()
        assert len(e) == 0               # disjoint accesses
