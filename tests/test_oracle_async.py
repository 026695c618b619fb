"""The oracle's coherency controller with fetches in flight
(xko_coherency_events), pinned two ways:
  * completing every fetch right after its access is the synchronous model,
    which the reference engine's partition fixture and the per-point brute
    force already pin: the event log must equal khp.coherency's plan;
  * out-of-order completion schedules against a per-point brute force that
    keeps awaiting accesses and forwards per point and per view
    (oracle/pointwise.coherency_events) instead of on the tree's pieces;
    a brute force that forgets to clear a device's fetching bit must fail."""
import numpy as np
import pytest

from oracle import khp, pointwise
from xkaapi_amd import streams as S

XD, YD = 24, 16


def stream(seed, n=250, ld=12, cols=14, s=2, pdev=0.5, maxm=11, maxn=9):
    rng = np.random.default_rng(seed)
    mode = rng.choice(np.array([1, 2, 3], np.uint8), n, p=[0.6, 0.2, 0.2])
    m = rng.integers(1, maxm, n).astype(np.uint64)
    nn = rng.integers(1, maxn, n).astype(np.uint64)
    row = rng.integers(0, ld, n).astype(np.uint64)
    col = rng.integers(0, cols, n).astype(np.uint64)
    m = np.minimum(m, ld - row)
    nn = np.minimum(nn, cols - col)
    dev = (rng.random(n) < pdev).astype(np.uint8)
    return S.Stream(np.arange(n, dtype=np.uint32), np.zeros(n, np.uint16), mode, dev, (col * ld + row) * s,
                    np.full(n, ld, np.uint64), m, nn, np.full(n, s, np.uint64))


def schedule(n, seed, pcomp=0.6, burst=3):
    """Plans of the n accesses in order, with completions of random in-flight fetches
    (random ranks: out of order) between them, then completions until none is left."""
    rng = np.random.default_rng(seed)
    et, ea = [], []
    for _ in range(n):
        et.append(0); ea.append(0)
        while rng.random() < pcomp:
            for _ in range(int(rng.integers(1, burst + 1))):
                et.append(1); ea.append(int(rng.integers(0, 1 << 30)))
    for _ in range(40 * n):
        et.append(1); ea.append(int(rng.integers(0, 1 << 30)))
    return np.array(et, np.uint8), np.array(ea, np.uint64)


def expand(rows, key):
    out = {}
    for r in rows:
        for x in range(int(r[3]), int(r[4])):
            for y in range(int(r[5]), int(r[6])):
                assert (x, y) not in out
                out[(x, y)] = key(r)
    return out


def run(st, et, ea):
    r, nr = st.rects()
    cb = st.m * st.n * st.s
    log, fail = khp.coherency_events(st.dev, st.mode, r, nr, cb, et, ea)
    return r, nr, cb, log, fail


@pytest.mark.parametrize("seed", range(4))
def test_immediate_completion_is_the_synchronous_plan(seed):
    st = stream(seed)
    n = len(st)
    et, ea = [], []
    for _ in range(n):  # each access's fetches complete before the next access
        et += [0] + [1] * 64
        ea += [0] * 65
    r, nr, cb, log, fail = run(st, np.array(et, np.uint8), np.array(ea, np.uint64))
    plan = khp.coherency(st.dev, st.mode, r, nr, cb)
    assert fail == plan.fail_at
    assert not np.isin(log[:, 1], [2, 3, 4, 5]).any()  # nothing awaits or forwards
    last = fail if fail >= 0 else n
    for i in range(last):
        parts = log[(log[:, 1] == 0) & (log[:, 2] == i)]
        want = plan.parts_of(i)
        assert np.array_equal(parts[:, 3:9], want[:, :6]), i
        f = log[(log[:, 1] == 1) & (log[:, 2] == i)]
        assert np.array_equal(f[:, [3, 4, 5, 6, 7, 8, 9, 10]], plan.fetches_of(i)), i


def check(seed, mistake=False):
    """The oracle's log against the brute force on one seeded schedule; the kinds of rows seen."""
    st = stream(100 + seed)
    et, ea = schedule(len(st), seed)
    r, nr, cb, log, fail = run(st, et, ea)
    want = pointwise.coherency_events(st.dev, st.mode, r, nr, cb, et, ea, log, XD, YD,
                                      **({"forget_fetch_clear": True} if mistake else {}))
    kinds = {int(k) for k in log[:, 1]}
    for ev, w in enumerate(want):
        rows = log[log[:, 0] == ev]
        if w[0] == "fail":
            assert fail == w[1]
            break
        if w[0] == "done":
            # several registrations may fire on the same points (accesses awaiting one fetch)
            fired = {}
            for q in rows[np.isin(rows[:, 1], [4, 5])]:
                k = (int(q[1]), int(q[2]), int(q[10]) if q[1] == 5 else -1)
                for x in range(int(q[3]), int(q[4])):
                    for y in range(int(q[5]), int(q[6])):
                        fired.setdefault((x, y), set()).add(k)
            assert fired == w[1], ev
            continue
        _, a, before, fetch, awaits, fwd = w
        assert expand(rows[rows[:, 1] == 0], lambda q: (int(q[7]), int(q[8]))) == before, ev
        assert expand(rows[rows[:, 1] == 1], lambda q: tuple(int(v) for v in q[7:11])) == fetch, ev
        assert expand(rows[rows[:, 1] == 2], lambda q: int(q[7])) == awaits, ev
        assert expand(rows[rows[:, 1] == 3], lambda q: (int(q[7]), int(q[8]))) == fwd, ev
    return kinds


@pytest.mark.parametrize("seed", range(8))
def test_out_of_order_completion_matches_pointwise(seed):
    check(seed)


def test_schedules_reach_every_branch():
    # across the seeds the schedules await, forward, release and issue forwards
    kinds = set()
    for seed in range(8):
        kinds |= check(seed)
    assert {0, 1, 2, 3, 4, 5} <= kinds


def test_planted_mistake_fails():
    # a brute force that never clears a device's fetching bit disagrees with the oracle
    with pytest.raises(AssertionError):
        for seed in range(8):
            check(seed, mistake=True)
