"""Host-side binding of the MI355X coherency fetch planner (libxkrt_hip.so).

Mirrors the reference's BLAS memory coherency controller
(include/xkrt/memory/access/blas/memory-tree.hpp): one tree per
``(ld, sizeof_type)`` (src/task/task.cc:50-150), a host and one GPU, and for
each access the valid-region split and the fetch list that ``fetch``
(memory-tree.hpp:1806-1874) computes, with every fetch completing before the
next access.  ``who_owns`` follows memory-tree.hpp:1903-1933.

There is no CPU fallback: the library must be built, and device errors raise
``ResolverError``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .resolver import ResolverError, hip_lib

_u64, _u32, _p = C.c_uint64, C.c_uint32, C.c_void_p


class MccConfigT(C.Structure):
    _fields_ = [("y_lo", _u64), ("y_hi", _u64), ("budget_bytes", _u64), ("async_completion", _u32), ("pad", _u32)]


class MccAccessT(C.Structure):
    _fields_ = [("addr", _u64), ("m", _u64), ("n", _u64), ("dev", C.c_uint8), ("mode", C.c_uint8),
                ("pad", C.c_uint8 * 6)]


MCC_ACCESS_DTYPE = np.dtype([("addr", "<u8"), ("m", "<u8"), ("n", "<u8"), ("dev", "u1"), ("mode", "u1"),
                             ("pad", "u1", (6,))])
assert MCC_ACCESS_DTYPE.itemsize == C.sizeof(MccAccessT) == 32


class MccStatsT(C.Structure):
    _fields_ = [(n, _u64) for n in ("accesses", "gpu_accesses", "chunks", "last_parts", "last_fetches",
                                    "device_bytes", "poisoned")] + [("last_ms", C.c_float), ("pad", _u32)] + \
               [(n, _u64) for n in ("last_fetch_base", "in_flight", "registrations")]


_BOUND = None


def mcc_lib():
    global _BOUND
    if _BOUND is not None:
        return _BOUND
    L = hip_lib()
    L.xkrt_hip_mcc_create.argtypes = [_u64, _u64, C.POINTER(MccConfigT), _p, C.POINTER(_p)]
    L.xkrt_hip_mcc_destroy.argtypes = [_p]
    L.xkrt_hip_mcc_plan.argtypes = [_p, _p, _u64, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(C.c_int64)]
    L.xkrt_hip_mcc_parts_copy.argtypes = [_p, _p, _p, _u64]
    L.xkrt_hip_mcc_fetches_copy.argtypes = [_p, _p, _p, _u64]
    L.xkrt_hip_mcc_who_owns.argtypes = [_p, _p, _u64, _p]
    L.xkrt_hip_mcc_who_owns_static.argtypes = [_p, _p, _u64, _p]
    L.xkrt_hip_mcc_stats.argtypes = [_p, C.POINTER(MccStatsT)]
    L.xkrt_hip_mcc_awaits_copy.argtypes = [_p, _p, _p, _u64]
    L.xkrt_hip_mcc_complete.argtypes = [_p, _p, _u64, C.POINTER(_u64), C.POINTER(_u64)]
    L.xkrt_hip_mcc_completed_copy.argtypes = [_p, _p, _p, _u64, _p, _p, _u64]
    L.xkrt_hip_mcc_exec_bind.argtypes = [_p, _p, _p, _u64]
    L.xkrt_hip_mcc_exec.argtypes = [_p, _p, _u64, _p]
    L.xkrt_hip_mcc_access_chunks.argtypes = [_p, _p]
    L.xkrt_hip_mcc_chunk_ptr.argtypes = [_p, _u32, C.POINTER(_p), C.POINTER(_u64), _p]
    _BOUND = L
    return L


def _chk(rc: int, what: str):
    if rc != 0:
        raise ResolverError(rc, what)


class CoherencyPlan:
    """Per-access valid-region split and fetch list (rows as the C ABI gives them)."""

    def __init__(self, part_off, parts, fetch_off, fetches, fail_at):
        self.part_off = part_off    # (n + 1,) row offsets per access
        self.parts = parts          # (P, 6) x0 x1 y0 y1 coh_before fetch_before
        self.fetch_off = fetch_off  # (n + 1,)
        self.fetches = fetches      # (F, 8) x0 x1 y0 y1 src_dev src_chunk dst_dev dst_chunk
        self.fail_at = fail_at      # access index at which the reference aborts, or -1

    def parts_of(self, i):
        return self.parts[self.part_off[i]:self.part_off[i + 1]]

    def fetches_of(self, i):
        return self.fetches[self.fetch_off[i]:self.fetch_off[i + 1]]


def to_mcc_array(dev, mode, addr, m, n) -> np.ndarray:
    a = np.zeros(len(addr), MCC_ACCESS_DTYPE)
    a["dev"] = dev
    a["mode"] = mode
    a["addr"] = addr
    a["m"] = m
    a["n"] = n
    return a


class MemoryTree:
    """A GPU-resident coherency tree for one (ld, sizeof_type) over columns [y_lo, y_hi)."""

    def __init__(self, ld: int, sizeof_type: int, y_range: tuple[int, int], budget_bytes: int = 0,
                 stream: int | None = None, device: int = 0, async_completion: bool = False):
        self.L = mcc_lib()
        _chk(self.L.xkrt_hip_init(device), "xkrt_hip_init")
        cfg = MccConfigT(int(y_range[0]), int(y_range[1]), budget_bytes, 1 if async_completion else 0, 0)
        h = _p()
        _chk(self.L.xkrt_hip_mcc_create(ld, sizeof_type, C.byref(cfg), stream, C.byref(h)), "xkrt_hip_mcc_create")
        self.h = h
        self.ld, self.s = ld, sizeof_type

    def close(self):
        if getattr(self, "h", None):
            self.L.xkrt_hip_mcc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def plan_counts(self, acc: np.ndarray):
        """Plan a batch; results stay on the device.  Returns (parts, fetches, fail_at)."""
        acc = np.ascontiguousarray(acc, MCC_ACCESS_DTYPE)
        npart, nfetch, fail = _u64(), _u64(), C.c_int64()
        _chk(self.L.xkrt_hip_mcc_plan(self.h, acc.ctypes.data, len(acc), C.byref(npart), C.byref(nfetch),
                                      C.byref(fail)), "xkrt_hip_mcc_plan")
        return int(npart.value), int(nfetch.value), int(fail.value)

    def plan(self, acc: np.ndarray) -> CoherencyPlan:
        """Plan a batch and copy its rows back."""
        n = len(acc)
        npart, nfetch, fail = self.plan_counts(acc)
        po = np.zeros(n + 1, np.uint64)
        fo = np.zeros(n + 1, np.uint64)
        parts = np.zeros((max(npart, 1), 6), np.int64)
        fetches = np.zeros((max(nfetch, 1), 8), np.int64)
        _chk(self.L.xkrt_hip_mcc_parts_copy(self.h, po.ctypes.data, parts.ctypes.data, len(parts)),
             "xkrt_hip_mcc_parts_copy")
        _chk(self.L.xkrt_hip_mcc_fetches_copy(self.h, fo.ctypes.data, fetches.ctypes.data, len(fetches)),
             "xkrt_hip_mcc_fetches_copy")
        return CoherencyPlan(po.astype(np.int64), parts[:npart], fo.astype(np.int64), fetches[:nfetch], fail)

    def awaits(self, n: int):
        """Registrations of the last plan call of n accesses (asynchronous trees):
        (offsets (n + 1), rows (R, 8): x0, x1, y0, y1, kind 2 await / 3 forward, chunk, dst chunk, id)."""
        off = np.zeros(n + 1, np.uint64)
        _chk(self.L.xkrt_hip_mcc_awaits_copy(self.h, off.ctypes.data, None, 0), "xkrt_hip_mcc_awaits_copy")
        rows = np.zeros((max(int(off[n]), 1), 8), np.int64)
        _chk(self.L.xkrt_hip_mcc_awaits_copy(self.h, off.ctypes.data, rows.ctypes.data, len(rows)),
             "xkrt_hip_mcc_awaits_copy")
        return off.astype(np.int64), rows[:int(off[n])]

    def complete(self, ids):
        """Complete fetches in flight, in order.  Returns per completion k
        (released offsets, released (R, 7), forward offsets, forwards (F, 9)):
        released rows = access, x0, x1, y0, y1, chunk, registration id; forward rows =
        fetch id, x0, x1, y0, y1, src dev, src chunk, dst dev, dst chunk."""
        ids = np.ascontiguousarray(ids, np.uint64)
        nr, nf = _u64(), _u64()
        _chk(self.L.xkrt_hip_mcc_complete(self.h, ids.ctypes.data, len(ids), C.byref(nr), C.byref(nf)),
             "xkrt_hip_mcc_complete")
        ro = np.zeros(len(ids) + 1, np.uint64)
        fo = np.zeros(len(ids) + 1, np.uint64)
        rel = np.zeros((max(nr.value, 1), 7), np.int64)
        fwd = np.zeros((max(nf.value, 1), 9), np.int64)
        _chk(self.L.xkrt_hip_mcc_completed_copy(self.h, ro.ctypes.data, rel.ctypes.data, len(rel), fo.ctypes.data,
                                                fwd.ctypes.data, len(fwd)), "xkrt_hip_mcc_completed_copy")
        return ro.astype(np.int64), rel[:nr.value], fo.astype(np.int64), fwd[:nf.value]

    def exec_bind(self, host_ptr: int, arena_ptr: int, arena_bytes: int):
        """Memory the fetches move: the host matrix (pinned, from column y_lo) and the chunk arena."""
        _chk(self.L.xkrt_hip_mcc_exec_bind(self.h, host_ptr, arena_ptr, arena_bytes), "xkrt_hip_mcc_exec_bind")

    def exec(self, ids, stream: int):
        """Enqueue the copies of these fetches in flight on `stream` (one kernel)."""
        ids = np.ascontiguousarray(ids, np.uint64)
        _chk(self.L.xkrt_hip_mcc_exec(self.h, ids.ctypes.data, len(ids), stream or None), "xkrt_hip_mcc_exec")

    def access_chunks(self, n: int) -> np.ndarray:
        """The GPU chunk of each access of the last plan call (-1: host)."""
        out = np.zeros(max(n, 1), np.int64)
        _chk(self.L.xkrt_hip_mcc_access_chunks(self.h, out.ctypes.data), "xkrt_hip_mcc_access_chunks")
        return out[:n]

    def chunk_ptr(self, chunk: int):
        """(device address, pitch bytes, rect (x0, x1 bytes, y0, y1 columns)) of a chunk."""
        ptr, pitch = _p(), _u64()
        rect = np.zeros(4, np.int64)
        _chk(self.L.xkrt_hip_mcc_chunk_ptr(self.h, chunk, C.byref(ptr), C.byref(pitch), rect.ctypes.data),
             "xkrt_hip_mcc_chunk_ptr")
        return int(ptr.value or 0), int(pitch.value), rect

    def who_owns(self, addr, m, n) -> np.ndarray:
        q = to_mcc_array(np.zeros(len(addr), np.uint8), np.zeros(len(addr), np.uint8), addr, m, n)
        out = np.zeros(max(len(q), 1), np.uint32)
        _chk(self.L.xkrt_hip_mcc_who_owns(self.h, q.ctypes.data, len(q), out.ctypes.data), "xkrt_hip_mcc_who_owns")
        return out[:len(q)]

    # queries per who_owns_static call (include/xkrt_hip.h)
    WHO_OWNS_STATIC_MAX = 32768

    def who_owns_static(self, addr, m, n) -> np.ndarray:
        """who_owns with no device allocation: staging in the library's static device array."""
        q = to_mcc_array(np.zeros(len(addr), np.uint8), np.zeros(len(addr), np.uint8), addr, m, n)
        out = np.zeros(max(len(q), 1), np.uint32)
        _chk(self.L.xkrt_hip_mcc_who_owns_static(self.h, q.ctypes.data, len(q), out.ctypes.data),
             "xkrt_hip_mcc_who_owns_static")
        return out[:len(q)]

    def stats(self) -> dict:
        st = MccStatsT()
        _chk(self.L.xkrt_hip_mcc_stats(self.h, C.byref(st)), "xkrt_hip_mcc_stats")
        return {f: (float(getattr(st, f)) if f == "last_ms" else int(getattr(st, f))) for f, _ in MccStatsT._fields_}


def concat_plans(plans) -> CoherencyPlan:
    """Join the plans of consecutive batches of one tree."""
    po, fo, parts, fetches = [np.zeros(1, np.int64)], [np.zeros(1, np.int64)], [], []
    pbase = fbase = 0
    fail = -1
    nacc = 0
    for p in plans:
        po.append(p.part_off[1:] + pbase)
        fo.append(p.fetch_off[1:] + fbase)
        parts.append(p.parts)
        fetches.append(p.fetches)
        pbase += int(p.part_off[-1])
        fbase += int(p.fetch_off[-1])
        if p.fail_at >= 0 and fail < 0:
            fail = nacc + p.fail_at
        nacc += len(p.part_off) - 1
    return CoherencyPlan(np.concatenate(po), np.concatenate(parts) if parts else np.zeros((0, 6), np.int64),
                         np.concatenate(fo), np.concatenate(fetches) if fetches else np.zeros((0, 8), np.int64), fail)


def plan_stream(stream, y_range, batch: int = 0, **tree_kw) -> CoherencyPlan:
    """Plan a single-tree stream on the GPU (in batches of `batch` accesses; 0 = one call)."""
    ld, s = int(stream.ld[0]), int(stream.s[0])
    acc = to_mcc_array(stream.dev, stream.mode, stream.addr, stream.m, stream.n)
    plans = []
    with MemoryTree(ld, s, y_range, **tree_kw) as t:
        step = batch or len(acc)
        for a in range(0, len(acc), step):
            p = t.plan(acc[a:a + step])
            plans.append(p)
            if p.fail_at >= 0:
                break
    out = concat_plans(plans)
    n = len(acc)
    if len(out.part_off) < n + 1:  # stopped early: the later accesses have no rows
        out.part_off = np.r_[out.part_off, np.full(n + 1 - len(out.part_off), out.part_off[-1])]
        out.fetch_off = np.r_[out.fetch_off, np.full(n + 1 - len(out.fetch_off), out.fetch_off[-1])]
    return out
