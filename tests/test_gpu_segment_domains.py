"""Segment dependency domains and task domains on the GPU, against the oracle.

* A segment domain (the reference's interval tree, interval/dependency-tree.hpp:
  187-363) gives the oracle's K = 1 edges (pinned by the reference's
  KHPTree<1>), for several column widths and bin shapes, in several batchings,
  at 64-bit addresses.
* A pool resolves a segment domain and a BLAS domain in one launch.
* A task domain over a mixed program (segments, tiles of three (ld, s) pairs,
  one pair's domain created late so that segments are replayed) gives the
  oracle's composition, pinned by the reference's own dispatch.
* Refusals leave nothing behind: a staged group is all or nothing, and a task
  domain batch refused with -ENOTSUP changes no later edge.
"""
import errno
import os

import numpy as np
import pytest

from oracle import khp
from oracle import segments as G
from xkaapi_amd import resolver as R

pytestmark = pytest.mark.gpu

BASE = 0x7F12_3456_0000
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODES = np.array([1, 2, 3, 4, 5, 6, 8, 9], np.uint8)


def pairs(x):
    # sorted rows; a row twice would break one edge per (pred access, succ task)
    x = np.asarray(x, np.int64).reshape(-1, 2)
    u = np.unique(x, axis=0)
    assert len(u) == len(x), "duplicate edges"
    return u


def same(x, y):
    return np.array_equal(pairs(x), pairs(y))


def seg_stream(seed, n, space, maxsz, per_task=3, base=BASE):
    rng = np.random.default_rng(seed)
    k = rng.integers(1, per_task + 1, n)
    task = np.repeat(np.arange(n, dtype=np.uint32), k)[:n]
    mode = rng.choice(MODES, n)
    a = rng.integers(0, space, n).astype(np.uint64)
    size = np.exp2(rng.random(n) * np.log2(maxsz)).astype(np.uint64) + 1
    b = np.minimum(a + size, space)
    return task, mode, np.stack([a, b], axis=1) + np.uint64(base)


def seg_acc(task, mode, seg):
    return R.to_access_array(task, mode, seg[:, 0], seg[:, 1] - seg[:, 0], 0,
                             kind=np.full(len(task), R.KIND_SEGMENT, np.uint8))


@pytest.mark.parametrize("seed,col,bx,by,tasks,kernel", [
    (0, 4096, 4096, 4, 500, 2),
    (1, 4096, 1024, 8, 2000, 2),
    (2, 1 << 16, 4096, 1, 100000, 1),
    (3, 256, 256, 64, 300, 2),
])
def test_segment_domain_matches_oracle(seed, col, bx, by, tasks, kernel):
    task, mode, seg = seg_stream(seed, 12000, 1 << 20, 1 << 14)
    want = khp.resolve_intervals(task, mode, seg)
    acc = seg_acc(task, mode, seg)
    print(f"segment domain seed {seed} col {col} bins {bx}x{by} tasks/batch {tasks}", flush=True)
    with R.SegmentDomain(col, bx, by, kernel=kernel) as d:
        for a, b in R.batches_by_tasks(task, tasks):
            d.submit(acc[a:b])
        d.sync()
        got = d.edges()
        st = d.stats()
    assert st["err"] == 0
    assert len(want) > 5000
    assert same(got, want)


def test_segment_domain_matches_golden_fixture():
    # the reference's KHPTree<1> edges (intervals_k1.npz, written by tests/golden/make_golden.py)
    z = np.load(os.path.join(GOLDEN, "intervals_k1.npz"))
    acc = seg_acc(z["task"], z["mode"], z["seg"])
    for col, bx, by, tasks in ((4096, 1024, 1, 100), (1 << 16, 4096, 1, 5000)):
        with R.SegmentDomain(col, bx, by, kernel=2) as d:
            for a, b in R.batches_by_tasks(z["task"], tasks):
                d.submit(acc[a:b])
            got = d.edges()
        assert same(got, z["edges"]), (col, bx, by, tasks)


def test_segment_domain_refuses_tiles():
    with R.SegmentDomain(4096, 4096, 4) as d:
        tile = R.to_access_array([0], [2], [BASE], [4], [4])
        with pytest.raises(R.ResolverError) as e:
            d.submit(tile)
        assert e.value.rc == -errno.EINVAL


def test_pool_resolves_segment_and_blas_domains_together():
    task, mode, seg = seg_stream(9, 4000, 1 << 18, 1 << 12, base=1 << 30)
    want_seg = khp.resolve_intervals(task, mode, seg)
    from xkaapi_amd import streams as S
    ch = S.cholesky(1024, 128)
    r, nr = ch.rects()
    want_ch = khp.resolve(ch.task, ch.mode, r, nr)
    print("pool: segment + cholesky domains", flush=True)
    with R.Pool(kernel=1, max_rects=1 << 18) as p:
        ds = p.segment_domain(4096, (0, 0), 1024, 1)
        dc = p.domain(1024, 8, (0, 1024))
        p.stage_group([ds, dc], [seg_acc(task, mode, seg), R.to_access_array(ch.task, ch.mode, ch.addr, ch.m, ch.n)])
        p.launch()
        p.sync()
        e, dom = p.edges()
        ids = ds.index, dc.index
    assert same(e[dom == ids[0]], want_seg)
    assert same(e[dom == ids[1]], want_ch)


# 1 KiB bins for every domain of the 64 KiB programs
TD = dict(kernel=2, seg_col_bytes=4096, seg_bin=(1024, 1), blas_bin=(128, 8))


def program_args(p):
    return [p[k] for k in ("task", "mode", "kind", "addr", "m", "n", "ld", "s")]


def program_acc(p):
    return R.to_task_access_array(p["task"], p["mode"], p["kind"], p["addr"], p["m"], p["n"], p["ld"], p["s"])


@pytest.mark.parametrize("seed,tasks,base", [(0, 50, 1 << 20), (1, 400, 1 << 20), (2, 3000, BASE), (3, 7, BASE)])
def test_task_domain_matches_oracle(seed, tasks, base):
    p = G.program_stream(6000, seed, base=base, span=1 << 16)
    want = khp.resolve_program(*program_args(p))
    acc = program_acc(p)
    print(f"task domain seed {seed} tasks/batch {tasks} base {base:#x}", flush=True)
    with R.TaskDomain(**TD) as t:
        for a, b in R.batches_by_tasks(p["task"], tasks):
            first = t.submit(acc[a:b])
            assert first == a
        got = t.edges()
        st = t.stats()
    assert st["blas_domains"] == 3 and st["segments"] == int((p["kind"] == 1).sum())
    assert len(want) > 3000
    assert same(got, want)


def test_task_domain_matches_golden_fixture():
    # the reference's dispatch over its own trees (program_mixed.npz)
    z = np.load(os.path.join(GOLDEN, "program_mixed.npz"))
    p = {k: z[k] for k in ("task", "mode", "kind", "addr", "m", "n", "ld", "s")}
    acc = program_acc(p)
    for tasks in (30, 2000):
        with R.TaskDomain(**TD) as t:
            for a, b in R.batches_by_tasks(p["task"], tasks):
                t.submit(acc[a:b])
            got = t.edges()
        assert same(got, z["edges"]), tasks


def test_task_domain_edges_in_pieces():
    # edges copied out after every few batches: their union is the program's edges
    p = G.program_stream(3000, 5)
    want = khp.resolve_program(*program_args(p))
    acc = program_acc(p)
    got = []
    with R.TaskDomain(**TD) as t:
        for k, (a, b) in enumerate(R.batches_by_tasks(p["task"], 200)):
            t.submit(acc[a:b])
            if k % 3 == 2:
                got.append(t.edges())
        got.append(t.edges())
        assert t.stats()["edges"] == sum(len(g) for g in got)
    assert same(np.concatenate(got), want)


def test_replay_refusals_and_wrapped_tile_guard():
    # a BLAS domain that gets a long segment replayed refuses a later wrapped tile; a segment
    # in the first ld * s bytes is refused and records nothing
    with R.DependencyDomain(64, 8) as d:  # LDs = 512 B
        seg = R.to_access_array([0], [2], [1 << 20], [3000], [0], kind=[R.KIND_SEGMENT])
        low = R.to_access_array([0], [2], [100], [10], [0], kind=[R.KIND_SEGMENT])
        with pytest.raises(R.ResolverError) as e:
            d.replay_segments(low)
        assert e.value.rc == -errno.ENOTSUP
        # before the replay a wrapped tile at (1 MiB + 496 B) with 4 rows is accepted elsewhere;
        # after it, the whole-column rect of the replayed segment spans its wrap gap
        wrapped = R.to_access_array([1], [2], [(1 << 20) + 8 * 62], [4], [2])
        d.replay_segments(seg)
        with pytest.raises(R.ResolverError) as e:
            d.submit(wrapped)
        assert e.value.rc == -errno.ENOTSUP
    with R.DependencyDomain(64, 8) as d:
        d.submit(R.to_access_array([1], [2], [(1 << 20) + 8 * 62], [4], [2]))
        d.sync()


def test_stage_group_is_all_or_nothing():
    # the second domain refuses its batch (a segment in its first column): the first domain's
    # batch is not staged either, and its next batch gets the same first id
    task, mode, seg = seg_stream(4, 500, 1 << 16, 1 << 10, base=1 << 24)
    acc = seg_acc(task, mode, seg)
    with R.Pool(kernel=1, max_rects=1 << 16) as p:
        ds = p.segment_domain(4096, (0, 0), 1024, 1)
        db = p.domain(64, 8, (0, 0))
        bad = R.to_access_array([0], [2], [16], [8], [0], kind=[R.KIND_SEGMENT])
        with pytest.raises(R.ResolverError) as e:
            p.stage_group([ds, db], [acc[:200], bad])
        assert e.value.rc == -errno.ENOTSUP
        assert p.open()[0] == 0
        first = p.stage_group([ds], [acc])
        assert first == [0]
        p.launch()
        p.sync()
        st = ds.stats()
        e, dom = p.edges()
        idx = ds.index
    with pytest.raises(R.ResolverError):
        ds.index  # the pool released its domains
    assert st["err"] == 0
    assert same(e[dom == idx], khp.resolve_intervals(task, mode, seg))


def test_task_domain_refused_batch_changes_nothing():
    p = G.program_stream(2000, 6)
    want = khp.resolve_program(*program_args(p))
    acc = program_acc(p)
    batches = R.batches_by_tasks(p["task"], 100)
    with R.TaskDomain(**TD) as t:
        for k, (a, b) in enumerate(batches):
            if k == len(batches) // 2:
                # a task with a tile and a segment in the first bytes of memory: refused whole
                bad = R.to_task_access_array([p["task"][a]] * 2, [2, 1], [0, 1], [1 << 20, 8], [1, 64], [1, 0],
                                             [32, 0], [4, 0])
                with pytest.raises(R.ResolverError) as e:
                    t.submit(bad)
                assert e.value.rc == -errno.ENOTSUP
            t.submit(acc[a:b])
        got = t.edges()
    assert same(got, want)


def test_stencil_workload_matches_oracle():
    # the measured 1-D stencil (fewer chunks and sweeps), in its measured domain configuration
    from xkaapi_amd import streams as S
    st = S.stencil1d(512, 1 << 16, 6, 256)
    n = len(st["task"])
    seg = np.stack([st["addr"], st["addr"] + st["size"]], axis=1)
    want = khp.resolve_intervals(st["task"], st["mode"], seg)
    acc = R.to_access_array(st["task"], st["mode"], st["addr"], st["size"], 0, kind=np.full(n, R.KIND_SEGMENT, np.uint8))
    with R.SegmentDomain(1 << 16, 1 << 16, 4, kernel=1, max_rects=1 << 16) as d:
        for a, b in R.batches_by_tasks(st["task"], 512):
            d.submit(acc[a:b])
        got = d.edges()
    assert len(want) > 5000 and same(got, want)


def canonical_virtual(e, virt):
    """Edges with each inserted empty write (id >= virt) renamed by its first program successor."""
    e = [(int(p), int(q)) for p, q in np.asarray(e).reshape(-1, 2).tolist()]
    first = {}
    for p, q in e:
        if p >= virt and q < virt:
            first[p] = min(first.get(p, q), q)
    name = lambda x: ("X", first[x]) if x >= virt else x
    return {(name(p), name(q)) for p, q in e}


@pytest.mark.parametrize("seed,tasks", [(0, 40), (1, 500)])
def test_task_domain_with_handles_matches_oracle(seed, tasks):
    # segments, tiles of three pairs and handle accesses (sequential, commutative and concurrent
    # writes) in one program: the handle map's edges join the others'
    p = G.program_stream(5000, 20 + seed)
    rng = np.random.default_rng(seed)
    hsel = rng.random(len(p["task"])) < 0.3
    kind = np.where(hsel, R.KIND_HANDLE, p["kind"]).astype(np.uint8)
    addr = np.where(hsel, rng.integers(0, 40, len(kind)).astype(np.uint64) * 8 + 8, p["addr"]).astype(np.uint64)
    conc = np.where(hsel & (p["mode"] & 2 > 0), rng.choice(np.array([1, 2, 4], np.uint8), len(kind)), 1)
    conc = conc.astype(np.uint8)
    args = [p["task"], p["mode"], kind, addr, p["m"], p["n"], p["ld"], p["s"]]
    want = khp.resolve_program(*args, conc=conc)
    acc = R.to_task_access_array(*args, concurrency=conc)
    with R.TaskDomain(**{k: v for k, v in TD.items()}) as t:
        for a, b in R.batches_by_tasks(p["task"], tasks):
            assert t.submit(acc[a:b]) == a
        got = t.edges()
    g = canonical_virtual(got.astype(object), R.TASKDOM_VIRTUAL)
    w = canonical_virtual(np.asarray(want, np.int64), 1 << 31)
    assert len(got) == len(want) and g == w


def test_task_domain_handle_limits():
    # more handle accesses in one batch than a handle launch takes are refused whole
    n = (1 << 18) + 1
    acc = R.to_task_access_array(np.arange(n, dtype=np.uint32), np.full(n, 2, np.uint8),
                                 np.full(n, R.KIND_HANDLE, np.uint8), np.arange(n, dtype=np.uint64) * 8 + 8,
                                 np.ones(n, np.uint64), np.zeros(n, np.uint64), np.zeros(n, np.uint64),
                                 np.zeros(n, np.uint64))
    with R.TaskDomain(**TD) as t:
        with pytest.raises(R.ResolverError) as ei:
            t.submit(acc)
        assert ei.value.rc == -errno.EINVAL
        assert t.submit(acc[:1000]) == 0
        assert len(t.edges()) == 0
