"""Serving loop: every client's edges are the same alone and in the mix
(batched, plain, open-loop arrivals, admission-limited), and equal the oracle."""
import numpy as np
import pytest

from oracle import khp
from xkaapi_amd import resolver as R
from xkaapi_amd import serving as V

pytestmark = pytest.mark.gpu


def unique_rows(e):
    # sorted rows; a row twice would break one edge per (pred access, succ task)
    e = np.asarray(e, np.int64).reshape(-1, 2)
    u = np.unique(e, axis=0) if len(e) else e
    assert len(u) == len(e), "duplicate edges"
    return u


@pytest.fixture(scope="module")
def clients():
    cs = V.make_clients(nclients=6, nmin=3, nmax=10, batch_tasks=32, seed=21)
    for c in cs:
        r, nr = c.stream.rects()
        c.want = unique_rows(khp.resolve(c.stream.task, c.stream.mode, r, nr))
    return cs


def check(res, clients):
    for c in clients:
        assert np.array_equal(unique_rows(res.edges[c.cid]), c.want), (res.mode, c.cid, c.kind, c.ntiles)
    assert len(res.lat_us) == sum(len(c.batches) for c in clients)


def test_alone_equals_oracle(clients):
    for c in clients:
        got = R.resolve_stream(c.stream, tasks_per_batch=32, bin_x_bytes=1024, bin_y_cols=128, kernel=1)
        assert np.array_equal(unique_rows(got), c.want)


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_saturation_mix(clients, mode):
    res = V.saturation(clients, mode, max_rects=1 << 16)
    check(res, clients)
    if mode == "batched":
        assert res.launches < sum(len(c.batches) for c in clients)


def test_launch_cap_mix(clients):
    # a cap on rects per launch splits the closed loop's backlog over more launches; same edges.
    # A launch closes once it holds `cap` rects and takes at most one batch more (at most 3
    # rects per access), so the capped run needs at least total / (cap + 3 * largest batch)
    # launches.  How the uncapped run splits the flood depends on the engine's timing.
    cap = 128
    capped = V.saturation(clients, "batched", max_rects=1 << 16, launch_rects=cap)
    check(capped, clients)
    free = V.saturation(clients, "batched", max_rects=1 << 16, launch_rects=0)
    check(free, clients)
    total = sum(len(c.stream) for c in clients)
    largest = max(e - a for c in clients for a, e in c.batches)
    assert capped.launches >= -(-total // (cap + 3 * largest))
    assert capped.launches > 1


@pytest.mark.parametrize("driver", ["native", "python"])
def test_open_loop_mix(clients, driver):
    # arrivals submitted on schedule from C (the default) or from Python
    sat = V.saturation(clients, "batched", max_rects=1 << 16, driver=driver)
    rate = 0.5 * sum(len(c.batches) for c in clients) / sat.makespan_s
    res = V.serve(clients, V.poisson_schedule(clients, rate, seed=3), "batched", max_rects=1 << 16, driver=driver)
    check(res, clients)
    assert (res.lag_us >= 0).all()


def test_admission_limited_mix(clients):
    # a budget that holds the scratch and about two clients: the others wait and are admitted as clients leave
    probe = R.Pool(kernel=1, max_rects=1 << 14)
    base = probe.stats()["device_bytes"]
    bpb = probe.stats()["bytes_per_bin"]
    probe.close()
    big = max(clients, key=lambda c: c.ntiles)
    lo, hi = V.col_range(big.stream)
    bins = (big.ntiles * 128 * 8 // 1024) * ((hi - lo + 127) // 128 + 1)
    res = V.saturation(clients, "batched", max_rects=1 << 14, budget_bytes=base + int(2.5 * bins * bpb))
    check(res, clients)


def test_edge_buffer_reset_under_load(clients):
    # a small edge buffer: the engine copies edges out and resets the buffer several times
    res = V.saturation(clients, "plain", max_rects=1 << 12, max_edges=2048)
    check(res, clients)
    assert res.stats["consumes"] >= 1 and res.stats["error"] == 0


def test_refused_client_fails_alone(clients):
    # client 0's frame is one column wide: its first batch is refused (-ERANGE), and so are its
    # later ones; the other clients are served as alone
    pool = R.Pool(kernel=1, max_rects=1 << 14)
    try:
        with R.Server(pool) as srv:
            ids, reqs = {}, {}
            for c in clients:
                lo, hi = V.col_range(c.stream)
                rng = (lo, lo + 1) if c.cid == clients[0].cid else (lo, hi)
                ids[c.cid] = srv.client(int(c.stream.ld[0]), int(c.stream.s[0]), rng)
            for c in clients:
                reqs[c.cid] = [srv.submit(ids[c.cid], c.acc[a:e]) for a, e in c.batches]
                srv.close_client(ids[c.cid])
            empty = srv.client(64, 8, (0, 64))  # registered and closed without a batch
            srv.close_client(empty)
            srv.drain(timeout_s=60)
            _, _, st = srv.times()
            assert set(st[reqs[clients[0].cid]].tolist()) == {-34}
            for c in clients[1:]:
                assert (st[reqs[c.cid]] == 0).all()
                assert np.array_equal(unique_rows(srv.edges(ids[c.cid])), c.want)
            s = srv.stats()
            assert s["left"] == len(clients) + 1 and s["error"] == 0
    finally:
        pool.close()


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_prearmed_launches_mix(clients, mode):
    # the next launch waits on the device behind a gate the engine opens by writing pinned memory
    # (one launch in flight; two below)
    sat = V.saturation(clients, mode, depth=1, max_rects=1 << 16, prearm=True, small_launch_rects=4096)
    check(sat, clients)
    rate = 0.3 * sum(len(c.batches) for c in clients) / sat.makespan_s
    res = V.serve(clients, V.poisson_schedule(clients, rate, seed=5), mode, depth=1, max_rects=1 << 16,
                  prearm=True, small_launch_rects=4096)
    check(res, clients)
    assert res.stats["gated"] > 0 and res.stats["error"] == 0


@pytest.mark.parametrize("load", [0.3, 0.8])
def test_prearmed_launches_two_in_flight(clients, load):
    # two launches in flight: the gate of the launch after next is armed once the oldest launch
    # has freed its staging buffer, so a buffer can be armed again before the engine retires
    # the launch that used it
    sat = V.saturation(clients, "batched", depth=2, max_rects=1 << 16, prearm=True, small_launch_rects=4096)
    check(sat, clients)
    rate = load * sum(len(c.batches) for c in clients) / sat.makespan_s
    res = V.serve(clients, V.poisson_schedule(clients, rate, seed=9), "batched", depth=2, max_rects=1 << 16,
                  prearm=True, small_launch_rects=4096)
    check(res, clients)
    assert res.stats["gated"] > 0 and res.stats["error"] == 0
