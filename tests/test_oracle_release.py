"""The oracle's successor release (oracle/release.py: wait counters and
successor lists, as the reference's __task_precedes / __task_commit /
__task_complete) against a closed form over sets, on random schedules that
interleave launches and out-of-order completions; planted mistakes (an edge
from a completed task counted, a detached access fulfilled by completion)
must fail."""
import numpy as np
import pytest

from oracle import khp
from oracle import release as REL
from oracle import segments as G


def program(seed, n=1500, pd=0.0):
    modes = (1, 2, 3) if pd == 0 else (1, 2, 3, 9, 10)
    p = G.program_stream(n, seed, modes=modes)
    e = np.asarray(khp.resolve_program(*[p[k] for k in ("task", "mode", "kind", "addr", "m", "n", "ld", "s")]),
                   np.int64)
    return p, e


def task_batches(task, k):
    starts = np.flatnonzero(np.r_[True, np.diff(task.astype(np.int64)) != 0])
    bounds = list(starts[::k]) + [len(task)]
    return list(zip(bounds[:-1], bounds[1:]))


def run(p, e, seed, batch_tasks=40, pcomp=0.7, mistake=None, check=True):
    """Drive one schedule through the oracle; returns the release events (per event, the tasks)."""
    rng = np.random.default_rng(seed)
    task, mode = p["task"], p["mode"]
    ntask = int(task.max()) + 1
    rel = REL.Release(task, mode)
    launched_at = np.full(ntask, np.inf)
    done_at = np.full(ntask, np.inf)
    cum, pool, events = set(), set(), []
    ev = 0

    def note(out):
        nonlocal ev
        cum.update(out)
        pool.update(out)
        events.append(sorted(out))
        if check:
            want = REL.released_closed_form(e, task, mode, launched_at, done_at, ev)
            assert cum == want, ev
        ev += 1

    def completion():
        pick = sorted(rng.choice(sorted(pool), min(len(pool), int(rng.integers(1, 8))), replace=False).tolist())
        for t in pick:
            pool.discard(t)
            done_at[t] = ev
        note(rel.complete(pick, honor_detached=mistake != "detached"))

    for a, b in task_batches(task, batch_tasks):
        t0, t1 = int(task[a]), int(task[b - 1]) + 1
        launched_at[t0:t1] = ev
        sel = (e[:, 1] >= a) & (e[:, 1] < b)
        note(rel.launch(e[sel], t0, t1, skip_completed=mistake != "count_completed"))
        while pool and rng.random() < pcomp:
            completion()
    while pool:
        completion()
    return events


@pytest.mark.parametrize("seed", range(5))
def test_release_matches_closed_form(seed):
    p, e = program(seed)
    events = run(p, e, seed)
    released = sorted(t for ev in events for t in ev)
    assert released == list(range(int(p["task"].max()) + 1))  # every task, once


def test_detached_accesses_block_their_successors():
    p, e = program(3, pd=0.3)
    events = run(p, e, 3)
    released = {t for ev in events for t in ev}
    assert len(released) < int(p["task"].max()) + 1  # some tasks wait on a detached access for ever


@pytest.mark.parametrize("mistake,pd", [("count_completed", 0.0), ("detached", 0.3)])
def test_planted_mistakes_fail(mistake, pd):
    with pytest.raises(AssertionError):
        for seed in range(5):
            p, e = program(seed, pd=pd)
            run(p, e, seed, mistake=mistake)


@pytest.mark.parametrize("seed,pd", [(0, 0.0), (1, 0.0), (2, 0.3)])
def test_c_release_run_matches_the_model(seed, pd):
    # oracle/khp_oracle.c's FIFO run (the CPU baseline of the device release): every task whose
    # predecessors all ran runs, in a topological order; the same tasks as the Python model
    # completing everything it releases
    p, e = program(seed, pd=pd)
    task, mode = p["task"], p["mode"]
    ran, order = khp.release_run(task, mode, e)
    rel = REL.Release(task, mode)
    out = rel.launch(e, 0, int(task.max()) + 1)
    done, pending = [], list(out)
    while pending:
        t = pending.pop(0)
        done.append(t)
        pending += rel.complete([t])
    assert ran == len(done) and set(order.tolist()) == set(done)
    pos = {int(t): i for i, t in enumerate(order.tolist())}
    for a, b in e:
        ta, tb = int(task[a]), int(task[b])
        if tb in pos:
            assert ta in pos and pos[ta] < pos[tb]
    if pd == 0:
        assert ran == int(task.max()) + 1


def handle_program(seed, n=1500):
    """A task-domain program with handle accesses (concurrent writes too): its
    edges (program ids, empty writes as VIRT | k) and, per empty write, the access
    it was inserted for (its first program successor)."""
    p = G.program_stream(n, seed, modes=(1, 2, 3))
    rng = np.random.default_rng(seed)
    hsel = rng.random(len(p["task"])) < 0.4
    kind = np.where(hsel, 2, p["kind"]).astype(np.uint8)
    addr = np.where(hsel, rng.integers(0, 12, len(kind)).astype(np.uint64) * 8 + 8, p["addr"]).astype(np.uint64)
    conc = np.where(hsel & (p["mode"] & 2 > 0), rng.choice(np.array([1, 4], np.uint8), len(kind)), 1).astype(np.uint8)
    args = [p["task"], p["mode"], kind, addr, p["m"], p["n"], p["ld"], p["s"]]
    e = np.asarray(khp.resolve_program(*args, conc=conc), np.int64)
    trig = {}
    for a, b in e:
        if a >= REL.VIRT and b < REL.VIRT:
            trig[int(a - REL.VIRT)] = min(trig.get(int(a - REL.VIRT), b), b)
    return dict(p, kind=kind, addr=addr, conc=conc), e, trig


def test_release_with_empty_writes_matches_closed_form():
    # inserted empty writes are tasks of their own, completed as soon as they are released
    for seed in range(3):
        p, e, trig = handle_program(seed)
        task, mode = p["task"], p["mode"]
        nv = len(trig)
        assert nv > 10 and sorted(trig) == list(range(nv))
        rng = np.random.default_rng(seed)
        rel = REL.Release(task, mode, nvirt=nv)
        ntask = int(task.max()) + 1
        la, da = np.full(ntask, np.inf), np.full(ntask, np.inf)
        vla = np.full(nv, np.inf)
        cum, pool, ev = set(), set(), 0
        kk = np.array([trig[k] for k in range(nv)])
        for a, b in task_batches(task, 30):
            t0, t1 = int(task[a]), int(task[b - 1]) + 1
            la[t0:t1] = ev
            vs = np.flatnonzero((kk >= a) & (kk < b))
            v = (int(vs.min()), int(vs.max()) + 1) if len(vs) else (0, 0)
            vla[v[0]:v[1]] = ev
            sel = ((e[:, 1] >= a) & (e[:, 1] < b)) | ((e[:, 1] >= REL.VIRT) & np.isin(e[:, 1] - REL.VIRT, vs))
            out = rel.launch(e[sel], t0, t1, virt=v)
            cum |= set(out); pool |= set(out)
            assert cum == REL.released_closed_form_virtual(e, task, mode, la, da, ev, vla), ev
            ev += 1
            while pool and rng.random() < 0.7:
                pick = sorted(rng.choice(sorted(pool), min(len(pool), 5), replace=False).tolist())
                pool.difference_update(pick)
                da[pick] = ev
                out = rel.complete(pick)
                cum |= set(out); pool |= set(out)
                assert cum == REL.released_closed_form_virtual(e, task, mode, la, da, ev, vla), ev
                ev += 1
        while pool:
            pick = sorted(pool)
            pool.clear()
            da[pick] = ev
            out = rel.complete(pick)
            cum |= set(out); pool |= set(out)
            ev += 1
        assert cum == set(range(ntask))


def test_empty_writes_left_guarded_block_their_successors():
    # a model that never drops the empty writes' guards releases fewer tasks than the closed form
    p, e, trig = handle_program(1)
    task, mode = p["task"], p["mode"]
    rel = REL.Release(task, mode, nvirt=len(trig))
    out = rel.launch(e, 0, int(task.max()) + 1, virt=(0, 0))
    done = set(out)
    pending = list(out)
    while pending:
        t = pending.pop()
        more = rel.complete([t])
        done |= set(more)
        pending += more
    assert len(done) < int(task.max()) + 1
