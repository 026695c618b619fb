"""Serving loop: every client's edges are the same alone and in the mix
(batched, plain, open-loop arrivals, admission-limited), and equal the oracle."""
import numpy as np
import pytest

from oracle import khp
from xkaapi_amd import resolver as R
from xkaapi_amd import serving as V

pytestmark = pytest.mark.gpu


def unique_rows(e):
    e = np.asarray(e, np.int64).reshape(-1, 2)
    return np.unique(e, axis=0) if len(e) else e


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


def test_open_loop_mix(clients):
    sat = V.saturation(clients, "batched", max_rects=1 << 16)
    rate = 0.5 * sum(len(c.batches) for c in clients) / sat.makespan_s
    res = V.serve(clients, V.poisson_schedule(clients, rate, seed=3), "batched", max_rects=1 << 16)
    check(res, clients)


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
