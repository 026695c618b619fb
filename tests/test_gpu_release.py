"""Successor release on the device (a task domain with release on) against
the oracle's restatement of __task_precedes / __task_commit /
__task_complete (oracle/release.py, pinned by a closed form): schedules that
interleave launches with out-of-order completions, the set of tasks released
after every step compared; detached accesses; misuse refused."""
import errno

import numpy as np
import pytest

from oracle import khp
from oracle import release as REL
from oracle import segments as G
from xkaapi_amd import resolver as R

pytestmark = pytest.mark.gpu

TD = dict(kernel=2, seg_col_bytes=4096, seg_bin=(1024, 1), blas_bin=(128, 8), release=True, max_tasks=1 << 14,
          max_accesses=1 << 16, max_succ=1 << 18)


def program(seed, n=3000, pd=0.0):
    modes = (1, 2, 3) if pd == 0 else (1, 2, 3, 9, 10)
    p = G.program_stream(n, seed, modes=modes)
    e = np.asarray(khp.resolve_program(*[p[k] for k in ("task", "mode", "kind", "addr", "m", "n", "ld", "s")]),
                   np.int64)
    return p, e


def task_batches(task, k):
    starts = np.flatnonzero(np.r_[True, np.diff(task.astype(np.int64)) != 0])
    bounds = list(starts[::k]) + [len(task)]
    return list(zip(bounds[:-1], bounds[1:]))


def drive(p, e, seed, batch_tasks=40, pcomp=0.7, wait=False):
    rng = np.random.default_rng(seed)
    task, mode = p["task"], p["mode"]
    acc = R.to_task_access_array(task, mode, p["kind"], p["addr"], p["m"], p["n"], p["ld"], p["s"])
    rel = REL.Release(task, mode)
    pool, steps = set(), 0
    with R.TaskDomain(**TD) as t:
        def step(want, sync=True):
            nonlocal steps
            if sync:
                t.sync()
            got = t.ready()
            assert sorted(got.tolist()) == sorted(want), (seed, steps)
            pool.update(want)
            steps += 1

        def completion():
            pick = sorted(rng.choice(sorted(pool), min(len(pool), int(rng.integers(1, 8))), replace=False).tolist())
            pool.difference_update(pick)
            t.complete(pick, wait=wait)
            # a completion that waited has released its successors: no sync before reading them
            step(rel.complete(pick), sync=not wait)

        for a, b in task_batches(task, batch_tasks):
            assert t.submit(acc[a:b]) == a
            sel = (e[:, 1] >= a) & (e[:, 1] < b)
            step(rel.launch(e[sel], int(task[a]), int(task[b - 1]) + 1))
            while pool and rng.random() < pcomp:
                completion()
        while pool:
            completion()
        st = t.stats()
    return st, steps


@pytest.mark.parametrize("seed", range(4))
def test_release_matches_oracle(seed):
    p, e = program(seed)
    print(f"release schedule seed {seed}: {int(p['task'].max()) + 1} tasks, {len(e)} edges", flush=True)
    st, steps = drive(p, e, seed)
    ntask = int(p["task"].max()) + 1
    assert st["released"] == ntask and st["completed"] == ntask
    assert st["succ_nodes"] <= len(e) and steps > 100


@pytest.mark.parametrize("seed", range(2))
def test_waiting_completions_release_before_returning(seed):
    p, e = program(10 + seed)
    print(f"waiting completions seed {seed}: {int(p['task'].max()) + 1} tasks, {len(e)} edges", flush=True)
    st, steps = drive(p, e, seed, wait=True)
    ntask = int(p["task"].max()) + 1
    assert st["released"] == ntask and st["completed"] == ntask and steps > 100


def test_detached_accesses_hold_their_successors():
    p, e = program(3, pd=0.3)
    st, _ = drive(p, e, 3)
    assert st["released"] < int(p["task"].max()) + 1


def test_misuse_is_refused():
    p, e = program(5, n=300)
    acc = R.to_task_access_array(p["task"], p["mode"], p["kind"], p["addr"], p["m"], p["n"], p["ld"], p["s"])
    with R.TaskDomain(**TD) as t:
        with pytest.raises(R.ResolverError) as ei:
            t.submit(acc[acc["task"] > 0])  # tasks must start at 0 and be consecutive
        assert ei.value.rc == -errno.EINVAL
        t.submit(acc)
        t.sync()
        with pytest.raises(R.ResolverError) as ei:
            t.edges()  # applied on the device
        assert ei.value.rc == -errno.EPERM
        ready = t.ready()
        waiting = sorted(set(range(int(p["task"].max()) + 1)) - set(ready.tolist()))
        assert len(waiting) > 0
        t.complete([waiting[0]])  # not released: the domain fails at the next sync
        with pytest.raises(R.ResolverError) as ei:
            t.sync()
        assert ei.value.rc == -errno.EINVAL


def test_cholesky_dag_runs_to_completion():
    # the measured loop (submit everything, then complete what is released) on a smaller Cholesky:
    # every task runs once, each after all its predecessors, as the oracle's FIFO run
    from xkaapi_amd import streams as S
    st = S.cholesky(2048, 128)
    ntask = int(st.task[-1]) + 1
    r, nr = st.rects()
    e = np.asarray(khp.resolve(st.task, st.mode, r, nr), np.int64)
    acc = R.to_task_access_array(st.task, st.mode, np.zeros(len(st), np.uint8), st.addr, st.m, st.n, st.ld, st.s)
    order = []
    with R.TaskDomain(kernel=1, release=True, max_tasks=1 << 14, max_accesses=1 << 16, max_succ=1 << 18) as t:
        for a, b in R.batches_by_tasks(st.task, 256):
            t.submit(acc[a:b])
        t.sync()
        for k in range(100000):
            rdy = t.ready()
            if len(rdy):
                order += rdy.tolist()
                t.complete(rdy, wait=k % 2 == 1)  # both forms, interleaved
            else:
                t.sync()
                if len(order) == ntask and not len(t.ready()):
                    break
        assert t.stats()["completed"] == ntask
    assert sorted(order) == list(range(ntask))
    pos = {tk: i for i, tk in enumerate(order)}
    for a, b in e:
        assert pos[int(st.task[a])] < pos[int(st.task[b])]
    ran, _ = khp.release_run(st.task, st.mode, e, ntask)
    assert ran == ntask


def test_release_with_handles_matches_oracle():
    # handle accesses in a released task domain: the empty writes the map inserts run on the device
    # (they never reach the queue); the released program tasks after every step equal the oracle's
    import sys, os
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_oracle_release import handle_program
    for seed in range(2):
        p, e, trig = handle_program(10 + seed, n=2500)
        task, mode = p["task"], p["mode"]
        nv = len(trig)
        kk = np.array([trig[k] for k in range(nv)])
        acc = R.to_task_access_array(task, mode, p["kind"], p["addr"], p["m"], p["n"], p["ld"], p["s"],
                                     concurrency=p["conc"])
        rel = REL.Release(task, mode, nvirt=nv)
        rng = np.random.default_rng(seed)
        pool, steps = set(), 0
        with R.TaskDomain(**TD) as t:
            def step(want):
                nonlocal steps
                t.sync()
                got = t.ready()
                assert sorted(got.tolist()) == sorted(want), (seed, steps)
                pool.update(want)
                steps += 1

            for a, b in task_batches(task, 30):
                assert t.submit(acc[a:b]) == a
                vs = np.flatnonzero((kk >= a) & (kk < b))
                v = (int(vs.min()), int(vs.max()) + 1) if len(vs) else (0, 0)
                sel = ((e[:, 1] >= a) & (e[:, 1] < b)) | ((e[:, 1] >= REL.VIRT) & np.isin(e[:, 1] - REL.VIRT, vs))
                step(rel.launch(e[sel], int(task[a]), int(task[b - 1]) + 1, virt=v))
                while pool and rng.random() < 0.7:
                    pick = sorted(rng.choice(sorted(pool), min(len(pool), 5), replace=False).tolist())
                    pool.difference_update(pick)
                    t.complete(pick)
                    step(rel.complete(pick))
            while pool:
                pick = sorted(pool)
                pool.clear()
                t.complete(pick)
                step(rel.complete(pick))
            assert t.stats()["completed"] == int(task.max()) + 1
