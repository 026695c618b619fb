"""ORACLE (test infrastructure): geometry of segment (1-D) accesses seen by a
BLAS dependency domain, restated from the reference:

  interval_to_rects   include/xkrt/memory/access/access.hpp:77-139
  the domain's rects  dependency-tree.hpp:358-364 (the domain's ld, sizeof_type)
  the marking rects   access.hpp:495-513: the segment's own rects, built with
                      ld = SIZE_MAX and sizeof_type = 1; on_insert tests tree
                      nodes against these (dependency-tree.hpp:241-256)

Interval overlap is strict (interval.hpp:110-119), so an empty interval
[c, c) still meets an interval that holds c strictly inside: the marking rect
[0, SIZE_MAX) x [1, 1) meets every node spanning columns 0 and 1.

Pinned by tests: the three rects cover exactly the bytes [a, a + size) of the
column-major frame, disjointly; and streams built with them resolve the same
on the reference engine (oracle/_ref) and on the restatement.
"""
from __future__ import annotations

import numpy as np

SIZE_MAX = (1 << 64) - 1


def interval_to_rects(a: int, size: int, ld: int, s: int):
    """Three rects [x0, x1, y0, y1] (some may be empty) of the segment [a, a + size)."""
    assert size > 0
    LDs = ld * s
    x0 = a % LDs
    x1 = min(x0 + size, LDs)
    dx10 = x1 - x0
    x2 = 0
    x3 = (size - dx10) % LDs
    dx32 = x3 - x2
    assert (size - dx10 - dx32) % LDs == 0
    y0 = a // LDs
    y1 = y0 + 1
    y3 = y0 + (size - dx10 - dx32) // LDs + 2
    y2 = y3 - 1
    return [(x0, x1, y0, y1), (0, LDs, y1, y2), (x2, x3, y2, y3)]


def stream_rects(kind, addr, m, n, ld, s):
    """Per access: inserted rects (12 u64), their count, marking rects (12 u64), their count.
    kind 0 = tile (addr, m rows, n columns; access.hpp:205-287), 1 = segment [addr, addr + m)."""
    cnt = len(addr)
    rects = np.zeros((cnt, 12), np.uint64)
    marks = np.zeros((cnt, 12), np.uint64)
    nrect = np.zeros(cnt, np.uint8)
    nmark = np.zeros(cnt, np.uint8)
    LDs = ld * s
    for i in range(cnt):
        a, mm, nn = int(addr[i]), int(m[i]), int(n[i])
        if kind[i] == 1:
            rs = interval_to_rects(a, mm, ld, s)
            ms = interval_to_rects(a, mm, SIZE_MAX, 1)
            for j, r in enumerate(rs):
                rects[i, 4 * j:4 * j + 4] = r
            for j, r in enumerate(ms):
                marks[i, 4 * j:4 * j + 4] = [min(v, SIZE_MAX) for v in r]
            nrect[i] = 3
            nmark[i] = 3
        else:
            X0, Y0, ms_ = a % LDs, a // LDs, mm * s
            if X0 + ms_ <= LDs:
                rr = [(X0, X0 + ms_, Y0, Y0 + nn)]
            else:
                rr = [(X0, LDs, Y0, Y0 + nn), (0, ms_ - (LDs - X0), Y0 + 1, Y0 + nn + 1)]
            for j, r in enumerate(rr):
                rects[i, 4 * j:4 * j + 4] = r
                marks[i, 4 * j:4 * j + 4] = r
            nrect[i] = len(rr)
            nmark[i] = len(rr)
    return rects, nrect, marks, nmark


def mixed_stream(n, ld, cols, s, seed, pseg=0.3, maxm=12, maxn=6, maxseg=None, modes=(1, 2, 3, 4, 5, 6, 8, 9)):
    """Tasks of 1..3 accesses, tiles and segments mixed, everything at or
    above the second column (address >= ld * s: column 0, whose nodes the
    segments' marking rects meet, stays untouched).  Returns dict of arrays."""
    rng = np.random.default_rng(seed)
    LDs = ld * s
    per = rng.integers(1, 4, n)
    task = np.repeat(np.arange(n, dtype=np.uint32), per)[:n]
    slot = np.zeros(n, np.uint16)
    for i in range(1, n):
        slot[i] = slot[i - 1] + 1 if task[i] == task[i - 1] else 0
    mode = rng.choice(np.array(modes, np.uint8), n)
    kind = (rng.random(n) < pseg).astype(np.uint8)
    m = rng.integers(1, maxm + 1, n).astype(np.uint64)
    nn = rng.integers(1, maxn + 1, n).astype(np.uint64)
    row = rng.integers(0, ld, n).astype(np.uint64)
    col = rng.integers(1, cols, n).astype(np.uint64)
    m = np.minimum(m, ld - row)                # tiles do not wrap
    nn = np.minimum(nn, cols - col)
    addr = (col * ld + row) * s
    maxseg = maxseg or 3 * LDs
    seg_a = rng.integers(LDs, (cols - 1) * LDs, n).astype(np.uint64)
    seg_len = rng.integers(1, maxseg + 1, n).astype(np.uint64)
    seg_len = np.minimum(seg_len, np.uint64(cols * LDs) - seg_a)
    addr = np.where(kind == 1, seg_a, addr)
    m = np.where(kind == 1, seg_len, m)
    nn = np.where(kind == 1, 1, nn).astype(np.uint64)
    return dict(task=task, slot=slot, mode=mode, kind=kind, addr=addr.astype(np.uint64), m=m.astype(np.uint64),
                n=nn, ld=ld, s=s)


def program_stream(n, seed, pairs=((32, 4), (16, 8), (24, 4)), base=1 << 20, span=1 << 15, pseg=0.35,
                   maxm=24, maxn=6, maxseg=1 << 11, modes=(1, 2, 3, 4, 5, 6, 8, 9), late=0.3):
    """A mixed program of one task domain: tasks of 1..3 accesses, segments and
    tiles of several (ld, sizeof_type) pairs over the bytes [base, base + span)
    (base >= every ld * s: nothing in the first column).  A pair's first tile
    comes after about late * n accesses, so its BLAS domain is created while
    segments are in flight.  Tiles do not wrap.  Returns a dict of arrays."""
    rng = np.random.default_rng(seed)
    per = rng.integers(1, 4, n)
    task = np.repeat(np.arange(n, dtype=np.uint32), per)[:n]
    mode = rng.choice(np.array(modes, np.uint8), n)
    kind = (rng.random(n) < pseg).astype(np.uint8)
    pair = rng.integers(0, len(pairs), n)
    # the last pair appears only in the later part of the program
    if len(pairs) > 1:
        pair[: int(late * n)] = rng.integers(0, len(pairs) - 1, int(late * n))
    ld = np.array([pairs[k][0] for k in pair], np.uint64)
    s = np.array([pairs[k][1] for k in pair], np.uint64)
    LDs = ld * s
    c0 = (np.uint64(base) + LDs - np.uint64(1)) // LDs
    ncol = np.uint64(span) // LDs
    col = c0 + (rng.random(n) * (ncol - np.uint64(maxn))).astype(np.uint64)
    row = (rng.random(n) * ld).astype(np.uint64)
    m = np.minimum(rng.integers(1, maxm + 1, n).astype(np.uint64), ld - row)
    nn = rng.integers(1, maxn + 1, n).astype(np.uint64)
    addr = (col * ld + row) * s
    sa = rng.integers(base, base + span - 1, n).astype(np.uint64)
    sl = np.minimum(np.exp2(rng.random(n) * np.log2(maxseg)).astype(np.uint64) + np.uint64(1),
                    np.uint64(base + span) - sa)
    addr = np.where(kind == 1, sa, addr).astype(np.uint64)
    m = np.where(kind == 1, sl, m).astype(np.uint64)
    nn = np.where(kind == 1, 0, nn).astype(np.uint64)
    return dict(task=task, mode=mode, kind=kind, addr=addr, m=m, n=nn, ld=ld, s=s)
