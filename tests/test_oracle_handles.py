"""The oracle's handle dependency map (oracle/handles.py) pinned by
  * the reference's own handle test (tests/task-dependency-handle.cc: W, R, RW
    on one handle gives the chain T1 -> T2 -> T3) and the two diagrams of
    dependency-map.hpp (reads then concurrent writes, concurrent writes then
    reads: an empty write in between);
  * sequential accesses: the same edges as the interval tree over one byte per
    handle (the K = 1 oracle, itself pinned by the reference's KHPTree<1>);
  * random streams with concurrent writes: every conflicting pair is ordered
    by a path, no two concurrent writes of one group are, edges go forward;
  * planted mistakes that fail."""
import numpy as np
import pytest

from oracle import handles as H
from oracle import khp

R, W, RW = 1, 2, 3
SEQ, COM, CNC = 1, 2, 4


def edges_set(e):
    return set(map(tuple, np.asarray(e, np.int64).reshape(-1, 2).tolist()))


def test_reference_handle_test_chain():
    e, nv = H.resolve([0, 1, 2], [7, 7, 7], [W, R, RW], [SEQ] * 3)
    assert edges_set(e) == {(0, 1), (1, 2)} and nv == 0


def test_reads_then_concurrent_writes_get_an_empty_write():
    # seq-r O O O -> X -> conc-w O O  (dependency-map.hpp:205-214)
    e, nv = H.resolve([0, 1, 2, 3, 4], [5] * 5, [R, R, R, W, W], [SEQ, SEQ, SEQ, CNC, CNC])
    X = H.VIRT
    assert nv == 1
    assert edges_set(e) == {(0, X), (1, X), (2, X), (X, 3), (X, 4)}


def test_concurrent_writes_then_reads_get_an_empty_write():
    # conc-w O O O -> X -> seq-r O O  (dependency-map.hpp:238-247)
    e, nv = H.resolve([0, 1, 2, 3, 4, 5], [5] * 6, [W, W, W, W, R, R], [SEQ, CNC, CNC, CNC, SEQ, SEQ])
    X = H.VIRT
    assert nv == 1
    assert edges_set(e) == {(0, 1), (0, 2), (0, 3), (1, X), (2, X), (3, X), (X, 4), (X, 5)}


def test_sequential_writer_after_concurrent_writes_depends_on_all():
    e, nv = H.resolve([0, 1, 2], [1, 1, 1], [W, W, RW], [CNC, CNC, SEQ])
    assert edges_set(e) == {(0, 2), (1, 2)} and nv == 0


def stream(seed, n, nh, pcnc=0.0, ptask=0.3, distinct=False):
    """Random handle accesses; distinct: a task's accesses are on distinct handles."""
    rng = np.random.default_rng(seed)
    task = np.cumsum(rng.random(n) > ptask).astype(np.uint32)
    task -= task[0]
    hid = rng.integers(0, nh, n)
    if distinct:
        for i in range(1, n):
            if task[i] == task[i - 1]:
                used = {int(hid[k]) for k in range(i - 1, -1, -1) if task[k] == task[i]}
                free = [h for h in range(nh) if h not in used]
                hid[i] = free[int(rng.integers(0, len(free)))] if free else hid[i]
    handle = hid.astype(np.uint64) * 64 + 0x7F00_0000_0000
    mode = rng.choice(np.array([R, W, RW], np.uint8), n)
    conc = np.where(rng.random(n) < pcnc, CNC, rng.choice(np.array([SEQ, COM], np.uint8), n)).astype(np.uint8)
    conc = np.where(mode & W, conc, SEQ).astype(np.uint8)
    return task, handle, mode, conc


@pytest.mark.parametrize("seed", range(5))
def test_sequential_accesses_match_the_interval_tree(seed):
    task, handle, mode, conc = stream(seed, 3000, 40)
    e, nv = H.resolve(task, handle, mode, conc)
    seg = np.stack([handle, handle + np.uint64(1)], axis=1)
    assert nv == 0
    assert edges_set(e) == edges_set(khp.resolve_intervals(task, mode, seg))


def reach(edges, n_nodes_hint):
    succ = {}
    for p, s in edges:
        succ.setdefault(p, []).append(s)
    memo = {}

    def closure(a):
        if a in memo:
            return memo[a]
        out = set()
        for s in succ.get(a, []):
            out.add(s)
            out |= closure(s)
        memo[a] = out
        return out
    return closure


@pytest.mark.parametrize("seed", range(4))
def test_concurrent_streams_order_every_conflict(seed):
    import sys
    sys.setrecursionlimit(100000)
    # one access per task: orders between tasks are those of their handle (a task with a concurrent
    # write and a read of one handle drops the write from the list -- the reference's put of the read
    # clears the concurrent writes, dependency-map.hpp:320-324 -- and a task on two handles orders
    # through both)
    task, handle, mode, conc = stream(100 + seed, 600, 6, pcnc=0.4, ptask=0.0)
    e, nv = H.resolve(task, handle, mode, conc)
    assert nv > 0
    E = edges_set(e)
    for p, s in E:  # forward: a program access never precedes an earlier one
        if p < H.VIRT and s < H.VIRT:
            assert p < s
    # dependencies hold between tasks: an edge from an access to another task's access orders the tasks
    tk = lambda x: int(task[x]) if x < H.VIRT else x
    closure = reach({(tk(p), tk(q)) for p, q in E}, len(task))
    n = len(task)
    for b in range(n):
        for a in range(b):
            if handle[a] != handle[b] or task[a] == task[b]:
                continue
            ra, rb = not (mode[a] & W), not (mode[b] & W)
            ca, cb = conc[a] == CNC and not ra, conc[b] == CNC and not rb
            if ra and rb:
                continue
            between = [k for k in range(a + 1, b) if handle[k] == handle[a]]
            same_group = ca and cb and all(conc[k] == CNC and (mode[k] & W) for k in between)
            if same_group:
                assert tk(b) not in closure(tk(a)), (a, b)  # concurrent writes stay unordered
            else:
                assert tk(b) in closure(tk(a)), (a, b)


@pytest.mark.parametrize("mistake", ["no_empty_write", "no_dedup"])
def test_planted_mistakes_fail(mistake):
    with pytest.raises(AssertionError):
        e, nv = H.resolve([0, 1, 2, 3, 4], [5] * 5, [R, R, R, W, W], [SEQ, SEQ, SEQ, CNC, CNC], mistakes=[mistake])
        assert nv == 1 and edges_set(e) == {(0, H.VIRT), (1, H.VIRT), (2, H.VIRT), (H.VIRT, 3), (H.VIRT, 4)}
        # a task with two accesses on one handle reaches its predecessor once
        e, nv = H.resolve([0, 1, 1], [5] * 3, [W, R, R], [SEQ] * 3, mistakes=[mistake])
        assert len(e) == 1
