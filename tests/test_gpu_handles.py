"""Handle dependency domains on the GPU against the oracle's restatement of
the reference's DependencyMap (oracle/handles.py, pinned by the reference's
handle test, its diagrams and the interval tree): random streams with
sequential, commutative and concurrent writes, tasks with several accesses
(on one handle too), several batchings.  Inserted empty writes get their ids
in whatever order the device's handles ran: both sides name each one by the
program access it was inserted for (its first successor)."""
import numpy as np
import pytest

from oracle import handles as H
from xkaapi_amd import resolver as R

pytestmark = pytest.mark.gpu

RM, WM, RWM = 1, 2, 3


def stream(seed, n, nh, pcnc=0.3, ptask=0.3):
    rng = np.random.default_rng(seed)
    task = np.cumsum(rng.random(n) > ptask).astype(np.uint32)
    task -= task[0]
    handle = rng.integers(0, nh, n).astype(np.uint64) * 64 + np.uint64(0x7F00_0000_0000)
    mode = rng.choice(np.array([RM, WM, RWM], np.uint8), n)
    conc = np.where(rng.random(n) < pcnc, 4, rng.choice(np.array([1, 2], np.uint8), n)).astype(np.uint8)
    conc = np.where(mode & WM, conc, 1).astype(np.uint8)
    return task, handle, mode, conc


def canonical(e):
    """Edges with every empty write renamed ('X', its first program successor)."""
    e = [(int(p), int(q)) for p, q in np.asarray(e, np.int64).reshape(-1, 2)]
    first = {}
    for p, q in e:
        if p >= H.VIRT and q < H.VIRT:
            first[p] = min(first.get(p, q), q)
    name = lambda x: ("X", first[x]) if x >= H.VIRT else x
    return {(name(p), name(q)) for p, q in e}


@pytest.mark.parametrize("seed,n,nh,batch", [(0, 4000, 50, 100), (1, 6000, 8, 1000), (2, 20000, 2000, 5000),
                                             (3, 3000, 3, 7)])
def test_handles_match_oracle(seed, n, nh, batch):
    task, handle, mode, conc = stream(seed, n, nh)
    want, nv = H.resolve(task, handle, mode, conc)
    acc = R.to_handle_access_array(task, handle, mode, conc)
    print(f"handles seed {seed}: {n} accesses, {nh} handles, batches of {batch} tasks", flush=True)
    with R.HandleDomain(max_batch=1 << 16) as d:
        for a, b in R.batches_by_tasks(task, batch):
            assert d.submit(acc[a:b]) == a
        got = d.edges()
        _, gv = d.sync()
    assert gv == nv and nv > 0
    assert len(got) == len(want)
    assert canonical(got) == canonical(want)


def test_reference_handle_test_chain():
    acc = R.to_handle_access_array([0, 1, 2], [7, 7, 7], [WM, RM, RWM], [1, 1, 1])
    with R.HandleDomain() as d:
        d.submit(acc)
        got = d.edges()
    assert {tuple(x) for x in got.tolist()} == {(0, 1), (1, 2)}


def test_bad_accesses_are_refused():
    import errno
    with R.HandleDomain() as d:
        for bad in (R.to_handle_access_array([0], [1], [1], [3]),             # concurrency 3
                    R.to_handle_access_array([1, 0], [1, 1], [1, 1], [1, 1]),  # task order
                    R.to_handle_access_array([0], [(1 << 64) - 1], [1], [1])):  # the reserved handle
            with pytest.raises(R.ResolverError) as ei:
                d.submit(bad)
            assert ei.value.rc == -errno.EINVAL


def test_measured_handle_workload_matches_oracle():
    # the measured handle stream (a prefix), in its measured batching
    from xkaapi_amd import streams as S
    st = S.handle_stream(40000, 512)
    want, nv = H.resolve(st["task"], st["handle"], st["mode"], st["concurrency"])
    acc = R.to_handle_access_array(st["task"], st["handle"], st["mode"], st["concurrency"])
    with R.HandleDomain(max_batch=1 << 16) as d:
        for a, b in R.batches_by_tasks(st["task"], 16384):
            d.submit(acc[a:b])
        got = d.edges()
        _, gv = d.sync()
    assert gv == nv and canonical(got) == canonical(want)
