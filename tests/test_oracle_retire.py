"""Pins the rule behind a domain's retire: once the accesses below a horizon
belong to completed tasks, the reference's edges (completed predecessors
link to nothing, include/xkrt/task/task.hpp:452-487) are its edges without
completion minus those whose predecessor is below the horizon at the
successor's LINK.  Checked against the per-point brute force with completion
on tiny streams; the GPU tests compare against the filtered oracle."""
import numpy as np
import pytest

from oracle import khp, pointwise
from xkaapi_amd import streams as S


def tiny(seed, n=220, ld=12, cols=16, s=2, maxm=8, maxn=6, per_task=3):
    rng = np.random.default_rng(seed)
    k = rng.integers(1, per_task + 1, n)
    task = np.repeat(np.arange(n, dtype=np.uint32), k)[:n]
    slot = np.zeros(n, np.uint16)
    for i in range(1, n):
        slot[i] = slot[i - 1] + 1 if task[i] == task[i - 1] else 0
    mode = rng.choice(np.array([1, 2, 3], np.uint8), n, p=[0.5, 0.25, 0.25])
    m = rng.integers(1, maxm + 1, n).astype(np.uint64)
    nn = rng.integers(1, maxn + 1, n).astype(np.uint64)
    row = rng.integers(0, ld, n).astype(np.uint64)
    m = np.minimum(m, ld - row)
    col = rng.integers(0, cols, n).astype(np.uint64)
    return S.Stream(task, slot, mode, np.ones(n, np.uint8), (col * ld + row) * s, np.full(n, ld, np.uint64), m,
                    nn, np.full(n, s, np.uint64))


def schedule(st, seed, nb=6):
    """Batches at task boundaries and, after each, a horizon at or below the
    ids handed out so far; done[a] = the horizon in force when a is LINKed."""
    rng = np.random.default_rng(1000 + seed)
    starts = st.task_starts()
    cut = np.unique(np.r_[0, np.sort(rng.choice(starts[1:-1], nb - 1, replace=False)), len(st)])
    done = np.zeros(len(st), np.int64)
    h = 0
    for a, b in zip(cut[:-1], cut[1:]):
        done[a:b] = h
        h = int(rng.integers(h, b + 1))
    return cut, done


def filtered(edges, done):
    e = np.asarray(edges, np.int64).reshape(-1, 2)
    return e[e[:, 0] >= done[e[:, 1]]]


@pytest.mark.parametrize("seed", range(8))
def test_completion_equals_filtered_oracle(seed):
    st = tiny(seed)
    r, nr = st.rects()
    _, done = schedule(st, seed)
    want = pointwise.resolve(st.task, st.mode, r, nr, 24, 24, done=done)
    got = set(map(tuple, filtered(khp.resolve(st.task, st.mode, r, nr), done).tolist()))
    assert got == want
    # the horizons matter on these streams: completion removes edges
    assert len(want) < len(pointwise.resolve(st.task, st.mode, r, nr, 24, 24))


def test_planted_mistake_is_caught():
    # a retire that also forgets the accesses AT the horizon (off by one) must be seen
    hit = False
    for seed in range(8):
        st = tiny(seed)
        r, nr = st.rects()
        _, done = schedule(st, seed)
        want = pointwise.resolve(st.task, st.mode, r, nr, 24, 24, done=done)
        got = set(map(tuple, filtered(khp.resolve(st.task, st.mode, r, nr), done + (done > 0)).tolist()))
        hit = hit or got != want
    assert hit
