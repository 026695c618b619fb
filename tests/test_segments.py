"""Segment (1-D) accesses seen by a BLAS dependency domain (no GPU): the
interval geometry, and the restatement against the reference engine, a
brute force and the golden fixture."""
import os

import numpy as np
import pytest

from oracle import khp, pointwise, refkhp
from oracle import segments as G

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("ld,s", [(8, 1), (12, 2), (64, 8), (5, 4)])
def test_interval_rects_cover_exactly_the_segment(ld, s):
    rng = np.random.default_rng(ld * 7 + s)
    LDs = ld * s
    for _ in range(300):
        a = int(rng.integers(0, 20 * LDs))
        size = int(rng.integers(1, 6 * LDs))
        covered = []
        for (x0, x1, y0, y1) in G.interval_to_rects(a, size, ld, s):
            for y in range(y0, y1):
                covered.extend(y * LDs + x for x in range(x0, x1))
        assert sorted(covered) == list(range(a, a + size))  # every byte once, nothing else


def test_marking_rects_are_one_row():
    # built with ld = SIZE_MAX, s = 1: [a, a + size) x [0, 1) and two empty rects
    r = G.interval_to_rects(1000, 77, G.SIZE_MAX, 1)
    assert r[0] == (1000, 1077, 0, 1)
    assert r[1][2] == r[1][3] and r[2][0] == r[2][1]


def build(seed, n=600, ld=16, cols=24, s=2, **kw):
    st = G.mixed_stream(n, ld, cols, s, seed, **kw)
    rects, nrect, marks, nmark = G.stream_rects(st["kind"], st["addr"], st["m"], st["n"], ld, s)
    return st, rects, nrect, marks, nmark


@pytest.mark.skipif(not refkhp.available(), reason="reference engine not built")
@pytest.mark.parametrize("seed", range(4))
def test_restatement_matches_reference_engine(seed):
    st, rects, nrect, marks, nmark = build(seed)
    want = refkhp.resolve_marks(st["task"], st["mode"], rects, nrect, marks, nmark)
    got = khp.resolve_marks(st["task"], st["mode"], rects, nrect, marks, nmark)
    assert np.array_equal(np.unique(want, axis=0), np.unique(got, axis=0))


@pytest.mark.parametrize("seed", range(4))
def test_restatement_matches_pointwise(seed):
    ld, cols, s = 8, 10, 1
    st, rects, nrect, marks, nmark = build(100 + seed, n=250, ld=ld, cols=cols, s=s, maxseg=20)
    got = khp.resolve_marks(st["task"], st["mode"], rects, nrect, marks, nmark)
    # segments at or above the first column: their PUT changes nothing in the domain
    put = st["kind"] == 0
    want = pointwise.resolve(st["task"], st["mode"], rects, nrect, ld * s, cols + 1, put=put)
    assert set(map(tuple, got.astype(np.int64).tolist())) == want


def test_segment_marks_nodes_spanning_column_zero():
    # a tile over columns 0-1, a reader segment far away, then a writer on column 1: the segment's
    # empty marking rect [0, SIZE_MAX) x [1, 1) meets pieces of the node [0, 8) x [0, 2) (split at
    # the segment's x breakpoints), so the segment becomes a reader there and the writer depends on
    # it (the GPU refuses segments once column 0 has been touched)
    ld, s = 8, 1
    kind = np.array([0, 1, 0], np.uint8)
    addr = np.array([0, 5 * 8 + 2, 8], np.uint64)  # tile [0,8)x[0,2); segment in column 5; tile [0,8)x[1,2)
    m = np.array([8, 3, 8], np.uint64)
    n = np.array([2, 1, 1], np.uint64)
    rects, nrect, marks, nmark = G.stream_rects(kind, addr, m, n, ld, s)
    task = np.arange(3, dtype=np.uint32)
    mode = np.array([2, 1, 2], np.uint8)
    e = set(map(tuple, khp.resolve_marks(task, mode, rects, nrect, marks, nmark).astype(np.int64).tolist()))
    assert (1, 2) in e  # no byte of the segment is in the writer's rect
    if refkhp.available():
        e2 = set(map(tuple, refkhp.resolve_marks(task, mode, rects, nrect, marks, nmark).astype(np.int64).tolist()))
        assert e == e2


def test_golden_fixture():
    z = np.load(os.path.join(GOLDEN, "segments_mixed.npz"))
    rects, nrect, marks, nmark = G.stream_rects(z["kind"], z["addr"], z["m"], z["n"], int(z["ld"]), int(z["s"]))
    got = khp.resolve_marks(z["task"], z["mode"], rects, nrect, marks, nmark)
    assert np.array_equal(np.unique(got.astype(np.int64), axis=0), z["edges"])
