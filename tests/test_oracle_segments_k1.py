"""Segment (K = 1) dependency domains in the oracle, pinned by the reference's
own KHPTree<1> with the IntervalDependencyTree hooks (oracle/_ref) and by the
per-point brute force on a tiny address space; and the segment -> BLAS replay
(src/task/task.cc:177-198): segments put into a new BLAS tree without a LINK
change no edge of the later tiles (their marks meet only nodes starting at
column 0), checked on the reference engine."""
import os

import numpy as np
import pytest

from oracle import khp, pointwise, refkhp
from oracle import segments as G

MODES = np.array([1, 2, 3, 4, 5, 6, 8, 9], np.uint8)


def seg_stream(seed, n, space, maxsz, per_task=3):
    rng = np.random.default_rng(seed)
    k = rng.integers(1, per_task + 1, n)
    task = np.repeat(np.arange(n, dtype=np.uint32), k)[:n]
    mode = rng.choice(MODES, n)
    a = rng.integers(0, space, n).astype(np.uint64)
    size = np.exp2(rng.random(n) * np.log2(maxsz)).astype(np.uint64) + 1
    b = np.minimum(a + size, space)
    return task, mode, np.stack([a, b], axis=1)


def pairs(x):
    if isinstance(x, (set, frozenset)):
        x = sorted(x)
    return np.unique(np.asarray(x, np.int64).reshape(-1, 2), axis=0)


def same(x, y):
    return np.array_equal(pairs(x), pairs(y))


@pytest.mark.parametrize("seed", range(6))
def test_k1_oracle_matches_reference_engine(seed):
    if not refkhp.available():
        pytest.skip("oracle/_ref not built (needs /root/reference)")
    task, mode, seg = seg_stream(seed, 3000, 1 << 20, 1 << 14)
    assert same(khp.resolve_intervals(task, mode, seg), refkhp.resolve_intervals(task, mode, seg))


@pytest.mark.parametrize("seed", range(4))
def test_k1_oracle_matches_pointwise(seed):
    task, mode, seg = seg_stream(100 + seed, 600, 96, 40)
    rects = np.zeros((len(seg), 8), np.uint64)
    rects[:, 0], rects[:, 1], rects[:, 3] = seg[:, 0], seg[:, 1], 1
    want = pointwise.resolve(task, mode, rects, np.ones(len(seg), np.uint8), 96, 1)
    assert same(khp.resolve_intervals(task, mode, seg), want)


def test_k1_planted_mistake_fails():
    # segments shifted by one byte resolve differently: the comparison sees it
    task, mode, seg = seg_stream(7, 600, 96, 40)
    rects = np.zeros((len(seg), 8), np.uint64)
    rects[:, 0], rects[:, 1], rects[:, 3] = seg[:, 0], seg[:, 1], 1
    want = pointwise.resolve(task, mode, rects, np.ones(len(seg), np.uint8), 96, 1)
    bad = seg.copy()
    bad[:, 1] = np.maximum(bad[:, 1] - 1, bad[:, 0] + 1)
    assert not same(khp.resolve_intervals(task, mode, bad), want)


def tiles(rng, ntile, ld, cols, s):
    """Random tiles over the whole matrix, column 0 included: (addr, m, n, mode)."""
    row = rng.integers(0, ld, ntile).astype(np.uint64)
    col = rng.integers(0, cols, ntile).astype(np.uint64)
    m = np.minimum(rng.integers(1, 12, ntile).astype(np.uint64), ld - row)
    n = np.minimum(rng.integers(1, 4, ntile).astype(np.uint64), cols - col)
    return (col * ld + row) * s, m, n, rng.choice(np.array([1, 2, 3], np.uint8), ntile)


def replay_vs_alone(seg_a, seg_size, seg_mode, taddr, tm, tn, tmode, ld, s):
    """Edges of the tiles after the segments were put into the BLAS tree (no LINK), and of
    the tiles alone (their task ids shifted past the segments')."""
    nseg, ntile = len(seg_a), len(taddr)
    kinds = np.r_[np.ones(nseg, np.uint8), np.zeros(ntile, np.uint8)]
    rects, nrect, marks, nmark = G.stream_rects(kinds, np.r_[seg_a, taddr], np.r_[seg_size, tm],
                                                np.r_[np.zeros(nseg, np.uint64), tn], ld, s)
    mode = np.r_[seg_mode | refkhp.PUT_ONLY, tmode]
    ttask = np.arange(ntile, dtype=np.uint32)
    task = np.r_[np.arange(nseg, dtype=np.uint32), ttask + nseg]
    got = refkhp.resolve_marks(task, mode, rects, nrect, marks, nmark)
    r2, n2, m2, nm2 = G.stream_rects(np.zeros(ntile, np.uint8), taddr, tm, tn, ld, s)
    alone = refkhp.resolve_marks(ttask, tmode, r2, n2, m2, nm2).astype(np.int64) + nseg
    return got, alone


@pytest.mark.parametrize("seed", range(4))
def test_replayed_segments_change_no_blas_edge(seed):
    # a BLAS domain created after segment accesses: the reference puts them into the new tree
    # (no LINK); unless one starts in the first ld * sizeof_type bytes, the later tiles' edges
    # are those of the tiles alone
    if not refkhp.available():
        pytest.skip("oracle/_ref not built (needs /root/reference)")
    ld, s, cols = 32, 4, 8
    rng = np.random.default_rng(seed)
    nseg = 60
    a = rng.integers(ld * s, ld * s * cols, nseg).astype(np.uint64)  # none starts in column 0
    size = rng.integers(1, 3 * ld * s, nseg).astype(np.uint64)
    smode = rng.choice(np.array([1, 2, 3], np.uint8), nseg)
    got, alone = replay_vs_alone(a, size, smode, *tiles(rng, 300, ld, cols, s), ld, s)
    assert same(got, alone)


def test_replayed_segments_in_column_zero_can_change_blas_edges():
    # the exception: a replayed segment that starts in the first column marks the nodes there
    # that it meets, and a later tile on column 0 gains an edge from that segment's task (the
    # GPU refuses this case)
    if not refkhp.available():
        pytest.skip("oracle/_ref not built (needs /root/reference)")
    ld, s, cols = 32, 4, 8
    changed = 0
    for seed in range(6):
        rng = np.random.default_rng(50 + seed)
        nseg = 40
        a = rng.integers(0, ld * s, nseg).astype(np.uint64)
        size = rng.integers(1, 3 * ld * s, nseg).astype(np.uint64)
        got, alone = replay_vs_alone(a, size, np.full(nseg, 2, np.uint8), *tiles(rng, 200, ld, cols, s), ld, s)
        changed += not same(got, alone)
    assert changed > 0


@pytest.mark.parametrize("seed", range(6))
def test_program_composition_matches_reference_dispatch(seed):
    # a task domain over segments and tiles of three (ld, s) pairs, one pair's BLAS tree created
    # late (the replay): the composition of one segment domain and one BLAS domain per pair
    # equals the reference's dispatch over its own trees
    if not refkhp.available():
        pytest.skip("oracle/_ref not built (needs /root/reference)")
    p = G.program_stream(4000, seed)
    args = [p[k] for k in ("task", "mode", "kind", "addr", "m", "n", "ld", "s")]
    want = refkhp.resolve_program(*args)
    got = khp.resolve_program(*args)
    assert len(want) > 1000
    assert same(got, want)
    # the segments reach predecessors through the BLAS domains, and the tiles through the segment domain never
    kind = p["kind"]
    e = pairs(want)
    assert ((kind[e[:, 0]] == 0) & (kind[e[:, 1]] == 1)).any()
    assert ((kind[e[:, 0]] == 1) & (kind[e[:, 1]] == 1)).any()
    assert not ((kind[e[:, 0]] == 1) & (kind[e[:, 1]] == 0)).any()
    # every predecessor's edges come from one domain (a segment's from the segment domain, a
    # tile's from its pair's BLAS domain): the domains never give the same (pred, succ task)
    k = np.stack([e[:, 0], p["task"][e[:, 1]].astype(np.int64)], axis=1)
    assert len(np.unique(k, axis=0)) == len(k)


@pytest.mark.parametrize("mistake", ["segments_skip_blas", "no_segment_domain"])
def test_program_composition_planted_mistakes_fail(mistake):
    # segments that do not LINK into the BLAS domains, or a lost segment domain, break the
    # equality with the reference's dispatch
    if not refkhp.available():
        pytest.skip("oracle/_ref not built (needs /root/reference)")
    p = G.program_stream(3000, 11)
    args = [p[k] for k in ("task", "mode", "kind", "addr", "m", "n", "ld", "s")]
    want = refkhp.resolve_program(*args)
    assert same(khp.resolve_program(*args), want)
    assert not same(khp.resolve_program(*args, _mistake=mistake), want)


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_k1_oracle_matches_golden_fixture():
    # intervals_k1.npz: the reference's KHPTree<1> edges, written by tests/golden/make_golden.py
    z = np.load(f"{GOLDEN}/intervals_k1.npz")
    assert same(khp.resolve_intervals(z["task"], z["mode"], z["seg"]), z["edges"])


def test_program_oracle_matches_golden_fixture():
    # program_mixed.npz: the reference's task-domain dispatch over its own trees
    z = np.load(f"{GOLDEN}/program_mixed.npz")
    args = [z[k] for k in ("task", "mode", "kind", "addr", "m", "n", "ld", "s")]
    assert same(khp.resolve_program(*args), z["edges"])


def test_program_with_handles_is_the_union_of_its_domains():
    # handle accesses join the program: the segment / BLAS edges are unchanged and the handle map's
    # edges (between handle accesses only) are added
    p = G.program_stream(2000, 12)
    rng = np.random.default_rng(12)
    hsel = rng.random(len(p["task"])) < 0.3
    kind = np.where(hsel, 2, p["kind"]).astype(np.uint8)
    addr = np.where(hsel, rng.integers(0, 20, len(kind)).astype(np.uint64) * 8 + 8, p["addr"]).astype(np.uint64)
    conc = np.where(hsel & (p["mode"] & 2 > 0) & (rng.random(len(kind)) < 0.3), 4, 1).astype(np.uint8)
    args = [p["task"], p["mode"], kind, addr, p["m"], p["n"], p["ld"], p["s"]]
    got = pairs(khp.resolve_program(*args, conc=conc))
    keep = ~hsel
    sub = [np.asarray(a)[keep] for a in args]
    ids = np.flatnonzero(keep)
    base = ids[pairs(khp.resolve_program(*sub))]
    from oracle import handles as HM
    h = np.flatnonzero(hsel)
    he = pairs(HM.HandleMap().resolve(h, p["task"][h], addr[h], p["mode"][h], conc[h]))
    assert np.array_equal(got, np.unique(np.concatenate([base, he]), axis=0))
    assert len(he) > 100
