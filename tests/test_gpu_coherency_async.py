"""GPU coherency planner with fetches in flight (async_completion) against the
oracle's event-driven restatement (xko_coherency_events): random schedules of
plan calls (several accesses each) and completion calls (several fetches
each, out of order), every output compared row for row -- the valid-region
split with fetching bits, the fetch rows and their ids, the accesses that
await and the forwards they register, and per completion the released
accesses and the forwards issued."""
import numpy as np
import pytest

from oracle import khp
from xkaapi_amd import coherency as K
from xkaapi_amd import streams as S

pytestmark = pytest.mark.gpu

LD, COLS, SZ = 12, 14, 2


def stream(seed, n=250):
    rng = np.random.default_rng(seed)
    mode = rng.choice(np.array([1, 2, 3], np.uint8), n, p=[0.6, 0.2, 0.2])
    m = np.minimum(rng.integers(1, 11, n).astype(np.uint64), LD - (row := rng.integers(0, LD, n).astype(np.uint64)))
    col = rng.integers(0, COLS, n).astype(np.uint64)
    nn = np.minimum(rng.integers(1, 9, n).astype(np.uint64), COLS - col)
    dev = (rng.random(n) < 0.5).astype(np.uint8)
    st = S.Stream(np.arange(n, dtype=np.uint32), np.zeros(n, np.uint16), mode, dev, (col * LD + row) * SZ,
                  np.full(n, LD, np.uint64), m, nn, np.full(n, SZ, np.uint64))
    # cut before the reference's ninth-view abort (allocation does not depend on completion order)
    r, nr = st.rects()
    fail = khp.coherency(st.dev, st.mode, r, nr, st.m * st.n * st.s).fail_at
    return st if fail < 0 else st.take(slice(0, fail))


def drive(st, seed, max_plan=7, burst=4):
    """A random schedule run on the GPU tree; returns its outputs per call and the
    equivalent oracle schedule (a completion = the rank of its fetch among those in flight)."""
    rng = np.random.default_rng(seed)
    acc = K.to_mcc_array(st.dev, st.mode, st.addr, st.m, st.n)
    n = len(st)
    et, ea, calls, inflight = [], [], [], set()
    with K.MemoryTree(LD, SZ, (0, COLS), async_completion=True) as t:
        a = 0
        while a < n or inflight:
            if a < n and (not inflight or rng.random() < 0.5):
                b = min(n, a + int(rng.integers(1, max_plan + 1)))
                plan = t.plan(acc[a:b])
                assert plan.fail_at == -1
                off, regs = t.awaits(b - a)
                base = t.stats()["last_fetch_base"]
                inflight |= set(range(base, base + len(plan.fetches)))
                calls.append(("plan", a, b, plan, off, regs, base))
                et += [0] * (b - a)
                ea += [0] * (b - a)
                a = b
            else:
                ids = [int(i) for i in rng.choice(sorted(inflight), min(len(inflight), int(rng.integers(1, burst + 1))),
                                                  replace=False)]
                cur = sorted(inflight)
                for i in ids:  # forwards issued meanwhile take larger ids: the ranks of these stand
                    et.append(1)
                    ea.append(cur.index(i))
                    cur.remove(i)
                ro, rel, fo, fwd = t.complete(ids)
                inflight -= set(ids)
                inflight |= {int(f[0]) for f in fwd}
                calls.append(("done", ids, ro, rel, fo, fwd))
        st_end = t.stats()
    assert st_end["in_flight"] == 0 and st_end["registrations"] == 0
    return calls, np.array(et, np.uint8), np.array(ea, np.uint64)


def check(seed):
    """One seeded schedule on the GPU against the oracle; the kinds of registration rows seen."""
    st = stream(200 + seed)
    calls, et, ea = drive(st, seed)
    r, nr = st.rects()
    log, fail = khp.coherency_events(st.dev, st.mode, r, nr, st.m * st.n * st.s, et, ea)
    assert fail == -1
    ev, seen = 0, set()
    for c in calls:
        if c[0] == "plan":
            _, a0, b, plan, off, regs, base = c
            for i, a in enumerate(range(a0, b)):
                rows = log[(log[:, 0] == ev + i) & (log[:, 2] == a)]
                assert np.array_equal(plan.parts_of(i), rows[rows[:, 1] == 0][:, 3:9]), (seed, a)
                f = plan.fetches_of(i)
                fid = base + plan.fetch_off[i] + np.arange(len(f))
                want = rows[rows[:, 1] == 1]
                assert np.array_equal(np.c_[f, fid], want[:, [3, 4, 5, 6, 7, 8, 9, 10, 11]]), (seed, a)
                g = regs[off[i]:off[i + 1]]
                want = rows[np.isin(rows[:, 1], [2, 3])]
                wk = want[:, 1]
                assert np.array_equal(g[:, 4], wk), (seed, a)
                assert np.array_equal(g[:, :4], want[:, 3:7]), (seed, a)
                assert np.array_equal(g[:, 5], want[:, 7]), (seed, a)
                assert np.array_equal(g[wk == 3][:, 6], want[wk == 3][:, 8]), (seed, a)
                assert np.array_equal(g[:, 7], want[:, 11]), (seed, a)
                seen |= set(wk.tolist())
            ev += b - a0
        else:
            _, ids, ro, rel, fo, fwd = c
            for k in range(len(ids)):
                rows = log[log[:, 0] == ev + k]
                w4, w5 = rows[rows[:, 1] == 4], rows[rows[:, 1] == 5]
                assert np.array_equal(rel[ro[k]:ro[k + 1]], w4[:, [2, 3, 4, 5, 6, 7, 11]]), (seed, ev + k)
                assert np.array_equal(fwd[fo[k]:fo[k + 1]], w5[:, [11, 3, 4, 5, 6, 7, 8, 9, 10]]), (seed, ev + k)
                seen |= {4} if len(w4) else set()
                seen |= {5} if len(w5) else set()
            ev += len(ids)
    assert ev == len(et)
    return seen


@pytest.mark.parametrize("seed", range(6))
def test_out_of_order_completion_matches_oracle(seed):
    check(seed)


def test_schedules_reach_awaits_and_forwards():
    seen = set()
    for seed in range(6):
        seen |= check(seed)
    assert {2, 3, 4, 5} <= seen


def test_completing_each_access_at_once_is_the_synchronous_plan():
    st = stream(7)
    acc = K.to_mcc_array(st.dev, st.mode, st.addr, st.m, st.n)
    with K.MemoryTree(LD, SZ, (0, COLS)) as ts, K.MemoryTree(LD, SZ, (0, COLS), async_completion=True) as ta:
        for i in range(len(st)):
            ps = ts.plan(acc[i:i + 1])
            pa = ta.plan(acc[i:i + 1])
            assert np.array_equal(ps.parts, pa.parts) and np.array_equal(ps.fetches, pa.fetches), i
            base = ta.stats()["last_fetch_base"]
            if len(pa.fetches):
                ro, rel, fo, fwd = ta.complete(list(range(base, base + len(pa.fetches))))
                assert len(rel) == 0 and len(fwd) == 0
        assert ta.stats()["in_flight"] == 0


def test_complete_refuses_ids_not_in_flight():
    st = stream(8, n=40)
    acc = K.to_mcc_array(st.dev, st.mode, st.addr, st.m, st.n)
    with K.MemoryTree(LD, SZ, (0, COLS), async_completion=True) as t:
        p = t.plan(acc)
        assert len(p.fetches) > 0
        with pytest.raises(Exception):
            t.complete([len(p.fetches) + 5])      # never issued
        t.complete([0])
        with pytest.raises(Exception):
            t.complete([0])                        # already completed
        with pytest.raises(Exception):
            t.complete([1, 1])                     # twice in one call
