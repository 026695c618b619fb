"""GPU parity of retire (id epochs) through the C ABI: after the accesses
below a horizon are retired, a domain's later edges equal the oracle's edges
minus those from a retired predecessor (the rule pinned in
tests/test_oracle_retire.py), with ids reported relative to the domain's id
base -- also across 2^32 ids over the domain's life and with a frame beyond
2^32 columns."""
import ctypes as C

import numpy as np
import pytest

from oracle import khp
from xkaapi_amd import resolver as R
from xkaapi_amd import streams as S

pytestmark = pytest.mark.gpu


def oracle_edges(st):
    r, nr = st.rects()
    return np.asarray(khp.resolve(st.task, st.mode, r, nr), np.int64).reshape(-1, 2)


def filtered(edges, done):
    return edges[edges[:, 0] >= done[edges[:, 1]]]


def uniq(e):
    # sorted rows; a row twice would break one edge per (pred access, succ task)
    e = np.asarray(e, np.int64).reshape(-1, 2)
    u = np.unique(e, axis=0)
    assert len(u) == len(e), "duplicate edges"
    return u


def schedule(st, seed, nb):
    rng = np.random.default_rng(seed)
    starts = st.task_starts()
    cut = np.unique(np.r_[0, np.sort(rng.choice(starts[1:-1], nb - 1, replace=False)), len(st)])
    hz = []
    h = 0
    for b in cut[1:]:
        h = int(rng.integers(h, b + 1))
        hz.append(h)
    return cut, hz


def run_domain(st, cut, hz, shift_cols=0, **kw):
    """Submit batch by batch, read the edges, retire; edges in stream positions."""
    ld, s = int(st.ld[0]), int(st.s[0])
    addr = st.addr + np.uint64(shift_cols * ld * s)
    acc = R.to_access_array(st.task, st.mode, addr, st.m, st.n)
    done = np.zeros(len(st), np.int64)
    out = []
    h_prev = 0
    with R.DependencyDomain(ld, s, **kw) as d:
        for (a, b), h in zip(zip(cut[:-1], cut[1:]), hz):
            base = d.id_base
            assert base == h_prev
            first = d.submit(acc[a:b])
            assert first + base == a
            out.append(d.edges().astype(np.int64) + base)
            done[a:b] = h_prev
            d.retire(h - base)
            h_prev = h
        assert d.id_base == hz[-1]
    return np.concatenate(out) if out else np.zeros((0, 2), np.int64), done


@pytest.mark.parametrize("name,seed", [("cholesky", 0), ("cholesky", 1), ("random", 2), ("random", 3)])
def test_retire_matches_filtered_oracle(name, seed):
    if name == "cholesky":
        st = S.cholesky(2048, 128)
        kw = dict(bin_x_bytes=2048, bin_y_cols=256)
    else:
        st = S.random2d(10000, seed=seed)
        kw = dict(bin_x_bytes=4096, bin_y_cols=512, kernel=2)
    cut, hz = schedule(st, seed, 7)
    got, done = run_domain(st, cut, hz, **kw)
    want = filtered(oracle_edges(st), done)
    assert len(want) < len(oracle_edges(st))  # retiring removed edges
    assert np.array_equal(uniq(got), uniq(want))


def test_retire_with_a_frame_beyond_2_32_columns():
    # the same geometry 2^33 columns up: the frame is rebased, the edges do not change
    st = S.cholesky(2048, 128)
    cut, hz = schedule(st, 5, 5)
    got, done = run_domain(st, cut, hz, shift_cols=1 << 33, bin_x_bytes=2048, bin_y_cols=256)
    assert np.array_equal(uniq(got), uniq(filtered(oracle_edges(st), done)))


def test_retire_refusals_and_no_op():
    st = S.cholesky(1024, 128)
    acc = R.to_access_array(st.task, st.mode, st.addr, st.m, st.n)
    with R.DependencyDomain(1024, 8, bin_x_bytes=2048, bin_y_cols=256) as d:
        d.submit(acc)
        d.sync()
        with pytest.raises(R.ResolverError) as ei:
            d.retire(1)  # edges not read yet
        assert ei.value.rc == -16  # -EBUSY
        d.edges()
        with pytest.raises(R.ResolverError) as ei:
            d.retire(len(st) + 1)  # past the ids handed out
        assert ei.value.rc == -22
        d.retire(0)
        assert d.id_base == 0
        d.retire(len(st))  # everything: a fresh domain with ids continuing
        assert d.id_base == len(st)
        assert d.submit(acc) == 0
        assert np.array_equal(uniq(d.edges()), uniq(oracle_edges(st)))


def test_caller_ids_across_2_32():
    # pool staging with caller ids near 2^32; retiring lets the ids of the domain's
    # life pass 2^32 while every id given stays 32-bit
    st = S.cholesky(2048, 128)
    n = len(st)
    off = (1 << 32) - 300  # the life id of stream position 0
    acc = R.to_access_array(st.task, st.mode, st.addr, st.m, st.n)
    starts = st.task_starts()
    cut = [0, int(starts[np.searchsorted(starts, 250)]), int(starts[np.searchsorted(starts, n // 2)]), n]
    hz = [cut[1] - 40, cut[2] - 10, n]  # stream positions
    L = R.hip_lib()
    got, done = [], np.zeros(n, np.int64)
    with R.Pool(kernel=1, max_rects=1 << 16) as pool:
        d = pool.domain(2048, 8, (0, 2048), bin_x_bytes=2048, bin_y_cols=256)
        h_prev = 0
        for (a, b), h in zip(zip(cut[:-1], cut[1:]), hz):
            base = d.id_base
            assert base == (0 if h_prev == 0 else off + h_prev)
            ids = (np.arange(a, b, dtype=np.int64) + off - base).astype(np.uint32)
            assert (np.arange(a, b, dtype=np.int64) + off - base).max() < (1 << 32)
            batch = np.ascontiguousarray(acc[a:b])
            dh = (C.c_void_p * 1)(d.h)
            ap = (C.c_void_p * 1)(batch.ctypes.data)
            ip = (C.c_void_p * 1)(ids.ctypes.data)
            cn = (C.c_uint64 * 1)(b - a)
            first = (C.c_uint64 * 1)()
            rc = L.xkrt_hip_pool_stage_group(pool.h, 1, dh, ap, ip, cn, first)
            assert rc == 0
            pool.launch()
            e, doms = pool.edges()
            got.append(e.astype(np.int64) + base - off)  # stream positions
            done[a:b] = h_prev
            d.retire(off + h - base)
            h_prev = h
        assert d.id_base == off + n
    want = filtered(oracle_edges(st), done)
    assert np.array_equal(uniq(np.concatenate(got)), uniq(want))
