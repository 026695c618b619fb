"""A domain that fails on the device fails alone: a bin that outgrows its
capacity class, or a record outside its domain's frame, marks only that
domain (-ENOSPC / -ERANGE); the pool's other domains, resolved in the same
launches, still equal the oracle."""
import numpy as np
import pytest

from oracle import khp
from xkaapi_amd import resolver as R
from xkaapi_amd import serving as V
from xkaapi_amd import streams as S

pytestmark = pytest.mark.gpu

ENOSPC, ERANGE, ECANCELED = -28, -34, -125


def unique_rows(e):
    # sorted rows; a row twice would break one edge per (pred access, succ task)
    e = np.asarray(e, np.int64).reshape(-1, 2)
    u = np.unique(e, axis=0) if len(e) else e
    assert len(u) == len(e), "duplicate edges"
    return u


def oracle_rows(st):
    r, nr = st.rects()
    return unique_rows(khp.resolve(st.task, st.mode, r, nr))


def dense_writers(n=600, seed=9):
    """Small random writes packed into one 1024 B x 64-column region (ld 128, fp64): the
    bin's writer map needs far more than class 0's 64 fragments."""
    rng = np.random.default_rng(seed)
    ld, s = 128, 8
    row = rng.integers(0, ld, n).astype(np.uint64)
    m = np.minimum(rng.integers(1, 9, n).astype(np.uint64), ld - row)
    col = rng.integers(0, 64, n).astype(np.uint64)
    return S.Stream(np.arange(n, dtype=np.uint32), np.zeros(n, np.uint16), np.full(n, 2, np.uint8),
                    np.ones(n, np.uint8), (col * ld + row) * s, np.full(n, ld, np.uint64), m,
                    rng.integers(1, 9, n).astype(np.uint64), np.full(n, s, np.uint64))


def acc_of(st):
    return R.to_access_array(st.task, st.mode, st.addr, st.m, st.n)


def test_capacity_overflow_fails_one_pool_domain():
    bad, good = dense_writers(), S.cholesky(1024, 128)
    with R.Pool(kernel=0, max_rects=1 << 14) as pool:
        da = pool.domain(128, 8, V.col_range(bad), 1024, 128)
        db = pool.domain(1024, 8, V.col_range(good), 1024, 128)
        # launches of 100 writes each (under the 256 records a class-0 bin takes per launch):
        # the dense domain's writer map outgrows 64 fragments in one of them
        bad_acc, first_b = acc_of(bad), None
        for k in range(6):
            pool.stage(da, bad_acc[100 * k:100 * (k + 1)])
            if k == 0:
                first_b = pool.stage(db, acc_of(good))
            pool.launch()
        pool.sync()  # the pool itself is healthy
        with pytest.raises(R.ResolverError) as ei:
            da.sync()
        assert ei.value.rc == ENOSPC and ei.value.err_bits & 4
        e, d = pool.edges()
        assert (d == db.index).all()  # the failed domain's edges are dropped
        assert np.array_equal(unique_rows(e.astype(np.int64) - first_b), oracle_rows(good))
        # the failed domain takes no batch until it is reset; the healthy one goes on
        with pytest.raises(R.ResolverError) as ei:
            pool.stage(da, acc_of(bad)[:10])
        assert ei.value.rc == ECANCELED
        assert pool.stats()["poisoned"] == 0


def test_record_outside_frame_fails_its_domain_in_a_shared_launch():
    # one launch holds the records of two domains; one of A's records leaves A's frame
    st_a, st_b = S.cholesky(1024, 128), S.lu(1024, 128)
    with R.Pool(kernel=1, max_rects=1 << 14) as pool:
        lo, hi = V.col_range(st_a)
        da = pool.domain(1024, 8, (lo, hi), 1024, 128)
        db = pool.domain(1024, 8, V.col_range(st_b), 1024, 128)
        ra, ylo, yhi = da.build_records(acc_of(st_a), 0)
        rb, _, _ = db.build_records(acc_of(st_b), 0)
        ra = ra.copy()
        ra["y1"][5] = hi - lo + 4096  # beyond A's grid
        ra["x1"][7] = 1024 * 8 + 64   # beyond the column (LDs)
        import torch
        recs = np.concatenate([ra, rb])
        dev = torch.from_numpy(recs.view(np.int32).reshape(-1)).cuda()
        torch.cuda.synchronize()  # the pool launches on a stream of its own
        da.submit_device(dev.data_ptr(), len(recs), len(st_a) + len(st_b), ylo, yhi)
        pool.sync()
        with pytest.raises(R.ResolverError) as ei:
            da.sync()
        assert ei.value.rc == ERANGE and ei.value.err_bits & 128
        db.sync()
        e, d = pool.edges()
        assert (d == db.index).all()
        assert np.array_equal(unique_rows(e), oracle_rows(st_b))


def test_record_naming_no_domain_poisons_only_its_pool():
    st = S.cholesky(1024, 128)
    with R.DependencyDomain(1024, 8, max_rects=1 << 12) as dom:
        recs, ylo, yhi = dom.build_records(acc_of(st), 0)
        recs = recs.copy()
        recs["dom"][3] = 999  # no such domain slot
        import torch
        dev = torch.from_numpy(recs.view(np.int32).reshape(-1)).cuda()
        torch.cuda.synchronize()  # the domain launches on a stream of its own
        dom.submit_device(dev.data_ptr(), len(recs), len(st), ylo, yhi)
        with pytest.raises(R.ResolverError) as ei:
            dom.sync()
        assert ei.value.rc == ERANGE
    # another domain of the process is untouched
    assert np.array_equal(unique_rows(R.resolve_stream(st, tasks_per_batch=64)), oracle_rows(st))


@pytest.mark.parametrize("prearm", [False, True])
def test_overflowing_client_fails_alone_in_serving(prearm):
    cs = V.make_clients(nclients=6, nmin=3, nmax=10, batch_tasks=32, seed=21)
    want = {c.cid: oracle_rows(c.stream) for c in cs}
    bad = dense_writers()
    bad_acc = acc_of(bad)
    pool = R.Pool(kernel=0, max_rects=1 << 14, small_launch_rects=4096 if prearm else 0)
    try:
        with R.Server(pool, depth=1, prearm=prearm) as srv:
            ids = {c.cid: srv.client(int(c.stream.ld[0]), int(c.stream.s[0]), V.col_range(c.stream)) for c in cs}
            ib = srv.client(128, 8, V.col_range(bad), 1024, 128)
            reqs, breq = {c.cid: [] for c in cs}, []
            # interleave: the bad client's batches arrive among the others'
            nb = max(len(c.batches) for c in cs)
            for k in range(nb):
                for c in cs:
                    if k < len(c.batches):
                        a, e = c.batches[k]
                        reqs[c.cid].append(srv.submit(ids[c.cid], c.acc[a:e]))
                if k < 4:
                    breq.append(srv.submit(ib, bad_acc[150 * k:150 * (k + 1)]))
            for c in cs:
                srv.close_client(ids[c.cid])
            srv.close_client(ib)
            srv.drain(timeout_s=60)
            _, _, st = srv.times()
            assert (st[breq] == ENOSPC).any() and set(st[breq].tolist()) <= {0, ENOSPC}
            for c in cs:
                assert (st[reqs[c.cid]] == 0).all(), c.cid
                assert np.array_equal(unique_rows(srv.edges(ids[c.cid])), want[c.cid]), c.cid
            s = srv.stats()
            assert s["error"] == 0 and s["left"] == len(cs) + 1
    finally:
        pool.close()
