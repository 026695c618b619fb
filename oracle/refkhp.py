"""Test infrastructure: ctypes binding of oracle/_ref/libkhpref.so, the
reference KHP engine driven by oracle/ref_harness.cc.  Only tests, smoke() and
bench.py's CPU-baseline leg may use this module."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(_HERE, "_ref", "libkhpref.so")
_lib = None


def available() -> bool:
    return os.path.exists(REF_SO)


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(REF_SO)
        p, u = C.c_void_p, C.c_uint64
        _lib.khpref_resolve.argtypes = [u, p, p, p, p, u, p, p, p]
        _lib.khpref_resolve.restype = C.c_int64
        _lib.khpref_time.argtypes = [u, p, p, p, p, p]
        _lib.khpref_time.restype = u
        _lib.khpref_partition.argtypes = [u, p, p, p, u, p, p]
        _lib.khpref_resolve_marks.argtypes = [u, p, p, p, p, p, p, u, p]
        _lib.khpref_resolve_marks.restype = C.c_int64
        _lib.khpref_partition.restype = C.c_int64
        _lib.khpref_resolve_intervals.argtypes = [u, p, p, p, u, p]
        _lib.khpref_resolve_intervals.restype = C.c_int64
        _lib.khpref_resolve_program.argtypes = [u, p, p, p, p, p, p, p, p, u, p]
        _lib.khpref_resolve_program.restype = C.c_int64
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dt)


def resolve(task, mode, rects, nrect, want_frags=False, want_times=False):
    """Edges (E, 2) uint32 (pred, succ) access indices; optional per-access
    fragment counts and per-task LINK+PUT nanoseconds."""
    L = lib()
    n = len(task)
    task = _c(task, np.uint32); mode = _c(mode, np.uint8)
    rects = _c(rects, np.uint64); nrect = _c(nrect, np.uint8)
    frag = np.zeros(n, np.uint64) if want_frags else None
    ntask = int(np.count_nonzero(np.diff(task.astype(np.int64))) + 1) if n else 0
    tns = np.zeros(ntask, np.uint64) if want_times else None
    cap = max(1024, 4 * n)
    while True:
        edges = np.empty((cap, 2), np.uint32)
        ne = L.khpref_resolve(n, task.ctypes.data, mode.ctypes.data, nrect.ctypes.data, rects.ctypes.data,
                              cap, edges.ctypes.data, frag.ctypes.data if frag is not None else None,
                              tns.ctypes.data if tns is not None else None)
        if ne <= cap:
            break
        cap = int(ne)
    out = [edges[:ne].copy()]
    if want_frags:
        out.append(frag)
    if want_times:
        out.append(tns)
    return out[0] if len(out) == 1 else tuple(out)


def resolve_marks(task, mode, rects, nrect, marks, nmark):
    """The reference engine with up to 3 inserted rects per access and separate marking rects."""
    L = lib()
    n = len(task)
    task = _c(task, np.uint32); mode = _c(mode, np.uint8)
    rects = _c(rects, np.uint64); nrect = _c(nrect, np.uint8)
    marks = _c(marks, np.uint64); nmark = _c(nmark, np.uint8)
    cap = max(1024, 4 * n)
    while True:
        edges = np.empty((cap, 2), np.uint32)
        ne = L.khpref_resolve_marks(n, task.ctypes.data, mode.ctypes.data, nrect.ctypes.data, rects.ctypes.data,
                                    nmark.ctypes.data, marks.ctypes.data, cap, edges.ctypes.data)
        if ne <= cap:
            break
        cap = int(ne)
    return edges[:ne].copy()


# mode flag of an access replayed without a LINK (resolve_marks): the segment
# accesses put into a new BLAS tree (src/task/task.cc:177-198)
PUT_ONLY = 0x40


def resolve_intervals(task, mode, seg):
    """The reference's segment (K = 1) tree over [a, b) byte segments (seg: (n, 2) u64)."""
    L = lib()
    n = len(task)
    task = _c(task, np.uint32); mode = _c(mode, np.uint8); seg = _c(seg, np.uint64)
    cap = max(1024, 4 * n)
    while True:
        edges = np.empty((cap, 2), np.uint32)
        ne = L.khpref_resolve_intervals(n, task.ctypes.data, mode.ctypes.data, seg.ctypes.data, cap,
                                        edges.ctypes.data)
        if ne <= cap:
            break
        cap = int(ne)
    return edges[:ne].copy()


def resolve_program(task, mode, kind, addr, m, n, ld, s):
    """The reference's task-domain dispatch over a mixed program: one interval
    tree, one BLAS tree per (ld, s) created on first use (with the replay), one
    successor list per access.  Edges in program access indices."""
    L = lib()
    cnt = len(task)
    a = [_c(task, np.uint32), _c(mode, np.uint8), _c(kind, np.uint8), _c(addr, np.uint64), _c(m, np.uint64),
         _c(n, np.uint64), _c(ld, np.uint64), _c(s, np.uint64)]
    cap = max(1024, 4 * cnt)
    while True:
        edges = np.empty((cap, 2), np.uint32)
        ne = L.khpref_resolve_program(cnt, *[x.ctypes.data for x in a], cap, edges.ctypes.data)
        if ne <= cap:
            break
        cap = int(ne)
    return edges[:ne].copy()


def time_resolve(task, mode, rects, nrect):
    L = lib()
    n = len(task)
    task = _c(task, np.uint32); mode = _c(mode, np.uint8)
    rects = _c(rects, np.uint64); nrect = _c(nrect, np.uint8)
    ne = np.zeros(1, np.uint64)
    ns = L.khpref_time(n, task.ctypes.data, mode.ctypes.data, nrect.ctypes.data, rects.ctypes.data, ne.ctypes.data)
    return int(ns), int(ne[0])


def partition(dev, rects, nrect):
    L = lib()
    n = len(dev)
    dev = _c(dev, np.uint8); rects = _c(rects, np.uint64); nrect = _c(nrect, np.uint8)
    off = np.zeros(n + 1, np.uint64)
    cap = max(1024, 8 * n)
    while True:
        parts = np.empty((cap, 4), np.uint64)
        np_ = L.khpref_partition(n, dev.ctypes.data, nrect.ctypes.data, rects.ctypes.data, cap,
                                 off.ctypes.data, parts.ctypes.data)
        if np_ <= cap:
            break
        cap = int(np_)
    return off.astype(np.int64), parts[:np_].copy()
