"""ORACLE (test infrastructure): ctypes binding of oracle/liboracle.so."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(_HERE, "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            raise RuntimeError(f"{SO} not built (make -C oracle)")
        _lib = C.CDLL(SO)
        p, u = C.c_void_p, C.c_uint64
        _lib.xko_resolve.argtypes = [u, p, p, p, p, u, p, p]
        _lib.xko_resolve.restype = C.c_int64
        _lib.xko_coherency.argtypes = [u, p, p, p, p, p, u, p, p, u, p, p, p, u, p, p, p]
        _lib.xko_coherency.restype = C.c_int64
        _lib.xko_resolve_marks.argtypes = [u, p, p, p, p, p, p, u, p]
        _lib.xko_resolve_marks.restype = C.c_int64
        _lib.xko_partition.argtypes = [u, p, p, p, u, p, p]
        _lib.xko_partition.restype = C.c_int64
        _lib.xko_coherency_events.argtypes = [u, p, p, p, p, p, u, p, p, u, p, p]
        _lib.xko_coherency_events.restype = C.c_int64
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dt)


def resolve(task, mode, rects, nrect, want_frags=False):
    """(pred, succ) access-index edges, grouped by predecessor (reference order)."""
    L = lib()
    n = len(task)
    task = _c(task, np.uint32); mode = _c(mode, np.uint8)
    rects = _c(rects, np.uint64); nrect = _c(nrect, np.uint8)
    frag = np.zeros(max(n, 1), np.uint64) if want_frags else None
    cap = max(1024, 4 * n)
    while True:
        edges = np.empty((cap, 2), np.uint32)
        ne = L.xko_resolve(n, task.ctypes.data, mode.ctypes.data, nrect.ctypes.data, rects.ctypes.data,
                           cap, edges.ctypes.data, frag.ctypes.data if frag is not None else None)
        if ne <= cap:
            break
        cap = int(ne)
    if want_frags:
        return edges[:ne].copy(), frag[:n]
    return edges[:ne].copy()


def resolve_marks(task, mode, rects, nrect, marks, nmark):
    """Edges with up to 3 inserted rects per access (12 u64) and separate marking rects (segments)."""
    L = lib()
    n = len(task)
    task = _c(task, np.uint32); mode = _c(mode, np.uint8)
    rects = _c(rects, np.uint64); nrect = _c(nrect, np.uint8)
    marks = _c(marks, np.uint64); nmark = _c(nmark, np.uint8)
    cap = max(1024, 4 * n)
    while True:
        edges = np.empty((cap, 2), np.uint32)
        ne = L.xko_resolve_marks(n, task.ctypes.data, mode.ctypes.data, nrect.ctypes.data, rects.ctypes.data,
                                 nmark.ctypes.data, marks.ctypes.data, cap, edges.ctypes.data)
        if ne <= cap:
            break
        cap = int(ne)
    return edges[:ne].copy()


def partition(dev, rects, nrect):
    L = lib()
    n = len(nrect)
    rects = _c(rects, np.uint64); nrect = _c(nrect, np.uint8)
    dev = _c(dev, np.uint8) if dev is not None else None
    off = np.zeros(n + 1, np.uint64)
    cap = max(1024, 8 * n)
    while True:
        parts = np.empty((cap, 4), np.uint64)
        np_ = L.xko_partition(n, dev.ctypes.data if dev is not None else None, nrect.ctypes.data,
                              rects.ctypes.data, cap, off.ctypes.data, parts.ctypes.data)
        if np_ <= cap:
            break
        cap = int(np_)
    return off.astype(np.int64), parts[:np_].copy()


class CoherencyPlan:
    """Per-access valid-region split and fetch list of one coherency tree."""

    owners = None  # who_owns answers for the queries, on the final state

    def __init__(self, part_off, parts, fetch_off, fetches, fail_at):
        self.part_off = part_off
        self.parts = parts          # (P, 6) x0 x1 y0 y1 coh_before fetch_before
        self.fetch_off = fetch_off
        self.fetches = fetches      # (F, 8) x0 x1 y0 y1 src_dev src_chunk dst_dev dst_chunk
        self.fail_at = fail_at      # index of the access that would abort, or -1

    def parts_of(self, i):
        return self.parts[self.part_off[i]:self.part_off[i + 1]]

    def fetches_of(self, i):
        return self.fetches[self.fetch_off[i]:self.fetch_off[i + 1]]


def coherency(dev, mode, rects, nrect, chunk_bytes, queries=None):
    """queries: optional (rects, nrect) answered with who_owns after the stream."""
    L = lib()
    if queries is not None:
        qr, qn = _c(queries[0], np.uint64), _c(queries[1], np.uint8)
    else:
        qr, qn = np.zeros((1, 8), np.uint64), np.zeros(1, np.uint8)
    nq = len(qn) if queries is not None else 0
    owners = np.zeros(max(nq, 1), np.uint32)
    n = len(dev)
    dev = _c(dev, np.uint8); mode = _c(mode, np.uint8)
    rects = _c(rects, np.uint64); nrect = _c(nrect, np.uint8)
    cb = _c(chunk_bytes, np.uint64)
    po = np.zeros(n + 1, np.int64); fo = np.zeros(n + 1, np.int64)
    fail = np.zeros(1, np.int64)
    pc, fc = max(1024, 8 * n), max(1024, 8 * n)
    while True:
        parts = np.empty((pc, 6), np.int64); fetches = np.empty((fc, 8), np.int64)
        rc = L.xko_coherency(n, dev.ctypes.data, mode.ctypes.data, nrect.ctypes.data, rects.ctypes.data,
                             cb.ctypes.data, pc, po.ctypes.data, parts.ctypes.data, fc, fo.ctypes.data,
                             fetches.ctypes.data, fail.ctypes.data, nq, qn.ctypes.data, qr.ctypes.data,
                             owners.ctypes.data)
        last = int(fail[0]) if rc < 0 else n
        tp, tf = int(po[last]), int(fo[last])
        if tp <= pc and tf <= fc:
            break
        pc, fc = max(pc, tp), max(fc, tf)
    last = int(fail[0]) if rc < 0 else n
    if rc < 0:
        po[last + 1:] = po[last]
        fo[last + 1:] = fo[last]
    plan = CoherencyPlan(po, parts[:po[last]].copy(), fo, fetches[:fo[last]].copy(), int(fail[0]) if rc < 0 else -1)
    plan.owners = owners[:nq]
    return plan


def coherency_events(dev, mode, rects, nrect, chunk_bytes, ev_type, ev_arg):
    """Coherency with fetches in flight, driven by an event schedule (ev_type 0:
    plan the next access; 1: complete the in-flight fetch of rank ev_arg, in id
    order).  Returns (log rows (N, 12) int64, fail_at): see xko_coherency_events."""
    L = lib()
    n = len(dev)
    dev = _c(dev, np.uint8); mode = _c(mode, np.uint8)
    rects = _c(rects, np.uint64); nrect = _c(nrect, np.uint8)
    cb = _c(chunk_bytes, np.uint64)
    et = _c(ev_type, np.uint8); ea = _c(ev_arg, np.uint64)
    fail = np.zeros(1, np.int64)
    cap = max(1024, 16 * n)
    while True:
        rows = np.empty((cap, 12), np.int64)
        rc = L.xko_coherency_events(n, dev.ctypes.data, mode.ctypes.data, nrect.ctypes.data, rects.ctypes.data,
                                    cb.ctypes.data, len(et), et.ctypes.data, ea.ctypes.data, cap, rows.ctypes.data,
                                    fail.ctypes.data)
        if rc <= cap:
            return rows[:rc].copy(), int(fail[0])
        cap = int(rc)


def resolve_intervals(task, mode, seg):
    """Segment (K = 1) domain: each access a byte segment [a, b) (seg (n, 2)), the
    same LINK / PUT rule over one dimension -- the 2D restatement with every
    segment as [a, b) x [0, 1)."""
    seg = np.asarray(seg, np.uint64).reshape(-1, 2)
    rects = np.zeros((len(seg), 8), np.uint64)
    rects[:, 0] = seg[:, 0]
    rects[:, 1] = seg[:, 1]
    rects[:, 3] = 1
    return resolve(task, mode, rects, np.ones(len(seg), np.uint8))


def release_run(acc_task, acc_mode, edges, ntask=None):
    """Runs the task graph on one host core (empty tasks, FIFO): the reference's
    edge application, commit and completion.  Returns (tasks run, completion order)."""
    L = lib()
    if not getattr(L, "_rel_typed", False):
        L.xko_release_run.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                      C.c_void_p]
        L.xko_release_run.restype = C.c_int64
        L._rel_typed = True
    acc_task = _c(acc_task, np.uint32)
    acc_mode = _c(acc_mode, np.uint8)
    edges = _c(np.asarray(edges).reshape(-1, 2), np.uint32)
    ntask = int(acc_task.max()) + 1 if ntask is None else ntask
    order = np.empty(max(ntask, 1), np.uint32)
    ran = L.xko_release_run(ntask, len(acc_task), acc_task.ctypes.data, acc_mode.ctypes.data, len(edges),
                            edges.ctypes.data, order.ctypes.data)
    return int(ran), order[:max(int(ran), 0)]


def resolve_program(task, mode, kind, addr, m, n, ld, s, _mistake=None, conc=None):
    """A task domain over a mixed program (src/task/task.cc:153-299): segments
    resolved by the segment domain (resolve_intervals); each (ld, s) pair's
    BLAS domain, created at its first tile, resolves its tiles and the segments
    that LINK into it (those after its creation), with their marking rects
    (resolve_marks).  The segments put into a BLAS domain with no LINK (the
    replay at creation, and those of the creating task before the tile) mark
    nothing, since no access of the program touches the first ld * s bytes;
    ValueError otherwise.  Edges of all domains in program indices, deduplicated
    per (predecessor, successor task) to the lowest successor (__access_precedes,
    task.hpp:494-512; one successor list per access across its domains).
    _mistake plants an error for the tests that show the comparison catches it:
    "segments_skip_blas" gives no segment to the BLAS domains, "no_segment_domain"
    drops the segment domain's edges."""
    from . import segments as G
    task = np.asarray(task, np.uint32)
    kind = np.asarray(kind, np.uint8)
    addr = np.asarray(addr, np.uint64)
    m = np.asarray(m, np.uint64)
    n = np.asarray(n, np.uint64)
    ld = np.asarray(ld, np.uint64)
    s = np.asarray(s, np.uint64)
    mode = np.asarray(mode, np.uint8)
    out = []
    hnd = np.flatnonzero(kind == 2)
    if len(hnd):
        # handle accesses: the dependency map (oracle/handles.py) over program ids; an inserted
        # empty write keeps the map's id (handles.VIRT | k)
        from . import handles as HM
        cc = np.ones(len(kind), np.uint8) if conc is None else np.asarray(conc, np.uint8)
        cc = np.where(cc == 0, 1, cc)
        hm = HM.HandleMap()
        out.append(np.asarray(hm.resolve(hnd, task[hnd], addr[hnd], mode[hnd], cc[hnd]), np.int64).reshape(-1, 2))
    seg = np.flatnonzero(kind == 1)
    if len(seg) and _mistake != "no_segment_domain":
        e = resolve_intervals(task[seg], mode[seg], np.stack([addr[seg], addr[seg] + m[seg]], axis=1))
        out.append(seg[np.asarray(e, np.int64)])
    tiles = np.flatnonzero(kind == 0)
    pairs = []
    for i in tiles:
        key = (int(ld[i]), int(s[i]))
        if key not in [p[0] for p in pairs]:
            pairs.append((key, int(i)))
    for (pld, ps), first in pairs:
        if ((addr < np.uint64(pld * ps)) & (kind != 2)).any():
            raise ValueError("an access in the first ld * s bytes: the replayed segments would mark nodes")
        late_seg = (kind == 1) & (np.arange(len(kind)) > first) & (_mistake != "segments_skip_blas")
        mine = np.flatnonzero(((kind == 0) & (ld == pld) & (s == ps)) | late_seg)
        if not len(mine):
            continue
        r, nr, mk, nmk = G.stream_rects(kind[mine], addr[mine], m[mine], n[mine], pld, ps)
        e = resolve_marks(task[mine], mode[mine], r, nr, mk, nmk)
        out.append(mine[np.asarray(e, np.int64)])
    if not out:
        return np.zeros((0, 2), np.int64)
    e = np.concatenate(out).reshape(-1, 2)
    tk = lambda x: np.where(x >= (1 << 31), x, task[np.minimum(x, len(task) - 1)].astype(np.int64))
    key = np.stack([e[:, 0], tk(e[:, 1]), e[:, 1]], axis=1)
    key = key[np.lexsort((key[:, 2], key[:, 1], key[:, 0]))]
    first = np.r_[True, (np.diff(key[:, 0]) != 0) | (np.diff(key[:, 1]) != 0)]
    return key[first][:, [0, 2]]
