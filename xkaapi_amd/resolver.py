"""Host-side binding of the MI355X dependency resolver (libxkrt_hip.so).

Mirrors the reference's BLAS dependency domain
(include/xkrt/memory/access/blas/dependency-tree.hpp:188-428): one domain per
``(ld, sizeof_type)`` (src/task/task.cc:153-201), accesses LINKed then PUT per
task (src/task/task.cc:285-299), edges ``pred -> succ`` with one entry per
(pred access, succ task) at the lowest succ slot (include/xkrt/task/task.hpp:494-512).

There is no CPU fallback: if the HIP library is missing the constructor
raises, and every device error is raised as ``ResolverError``.
"""
from __future__ import annotations

import ctypes as C
import errno
import os
import weakref

import numpy as np

from ._loader import LibraryMissing, lib_path

_u64, _u32, _p = C.c_uint64, C.c_uint32, C.c_void_p


class AccessT(C.Structure):
    _fields_ = [("addr", C.c_uint64), ("m", C.c_uint64), ("n", C.c_uint64), ("task", C.c_uint32),
                ("mode", C.c_uint8), ("kind", C.c_uint8), ("pad", C.c_uint8 * 2)]


KIND_TILE, KIND_SEGMENT = 0, 1
ACCESS_DTYPE = np.dtype([("addr", "<u8"), ("m", "<u8"), ("n", "<u8"), ("task", "<u4"), ("mode", "u1"),
                         ("kind", "u1"), ("pad", "u1", (2,))])
assert ACCESS_DTYPE.itemsize == C.sizeof(AccessT) == 32


class ConfigT(C.Structure):
    _fields_ = [("bin_x_bytes", C.c_uint32), ("bin_y_cols", C.c_uint32), ("kernel", C.c_uint32),
                ("pad", C.c_uint32), ("max_rects", C.c_uint64), ("max_edges", C.c_uint64),
                ("budget_bytes", C.c_uint64), ("y_hint_lo", C.c_uint64), ("y_hint_hi", C.c_uint64)]


class StatsT(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("accesses", "rects", "edges", "candidates", "launches", "bins",
                                          "max_frags", "max_reads", "frag_visits", "device_bytes", "err")]


RECT_DTYPE = np.dtype([(n, "<u4") for n in ("x0", "x1", "y0", "y1", "acc", "tm", "dom", "pad")])
assert RECT_DTYPE.itemsize == 32


class PoolConfigT(C.Structure):
    _fields_ = [("kernel", C.c_uint32), ("graph", C.c_uint32), ("max_rects", C.c_uint64),
                ("max_edges", C.c_uint64), ("budget_bytes", C.c_uint64), ("direct_dedup", C.c_uint32),
                ("small_launch_rects", C.c_uint32), ("host_edges", C.c_uint32), ("pad", C.c_uint32)]


class PoolStatsT(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("domains", "bins_capacity", "bins_used", "bytes_per_bin", "device_bytes",
                                          "launches", "graph", "poisoned")]


class ServerConfigT(C.Structure):
    _fields_ = [("depth", C.c_uint32), ("max_batches", C.c_uint32), ("max_rects", C.c_uint64),
                ("reserve_bytes", C.c_uint64), ("prearm", C.c_uint32), ("pad", C.c_uint32)]


class ServerStatsT(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("launches", "batches", "admitted", "left", "waited", "harvests",
                                          "consumes", "gated")] + [("error", C.c_int64)]


class TaskAccessT(C.Structure):
    _fields_ = [("addr", C.c_uint64), ("m", C.c_uint64), ("n", C.c_uint64), ("ld", C.c_uint64),
                ("sizeof_type", C.c_uint32), ("task", C.c_uint32), ("mode", C.c_uint8), ("kind", C.c_uint8),
                ("concurrency", C.c_uint8), ("pad", C.c_uint8 * 5)]


KIND_HANDLE = 2
TASKDOM_VIRTUAL = 1 << 63
TASK_ACCESS_DTYPE = np.dtype([("addr", "<u8"), ("m", "<u8"), ("n", "<u8"), ("ld", "<u8"), ("sizeof_type", "<u4"),
                              ("task", "<u4"), ("mode", "u1"), ("kind", "u1"), ("concurrency", "u1"),
                              ("pad", "u1", (5,))])
assert TASK_ACCESS_DTYPE.itemsize == C.sizeof(TaskAccessT) == 48


class TaskDomConfigT(C.Structure):
    _fields_ = [("kernel", C.c_uint32), ("release", C.c_uint32), ("max_rects", C.c_uint64),
                ("max_edges", C.c_uint64), ("budget_bytes", C.c_uint64), ("seg_col_bytes", C.c_uint64),
                ("seg_bin_x_bytes", C.c_uint32), ("seg_bin_y_cols", C.c_uint32), ("blas_bin_x_bytes", C.c_uint32),
                ("blas_bin_y_cols", C.c_uint32), ("max_tasks", C.c_uint64), ("max_accesses", C.c_uint64),
                ("max_succ", C.c_uint64), ("max_virtual", C.c_uint64)]


class TaskDomStatsT(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("accesses", "segments", "blas_domains", "launches", "edges", "succ_nodes",
                                          "released", "completed")]


class HandleAccessT(C.Structure):
    _fields_ = [("handle", C.c_uint64), ("task", C.c_uint32), ("mode", C.c_uint8), ("concurrency", C.c_uint8),
                ("pad", C.c_uint8 * 2)]


HANDLE_ACCESS_DTYPE = np.dtype([("handle", "<u8"), ("task", "<u4"), ("mode", "u1"), ("concurrency", "u1"),
                                ("pad", "u1", (2,))])
assert HANDLE_ACCESS_DTYPE.itemsize == C.sizeof(HandleAccessT) == 16
CONC_SEQUENTIAL, CONC_COMMUTATIVE, CONC_CONCURRENT = 1, 2, 4
VIRTUAL_ID = 1 << 31


class HandleConfigT(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("max_handles", "max_accesses", "max_virtual", "max_batch", "max_edges")]


# device error bits: 2, 4, 8 and 128 belong to one domain (the pool's other domains go on);
# 16, 32 and 64 to the pool (shared launch scratch overflowed)
ERR_BITS = {2: "bin record list", 4: "bin fragments", 8: "bin reader log", 128: "record outside its frame",
            16: "edge candidates", 32: "dedup table", 64: "edge buffer"}


class ResolverError(RuntimeError):
    def __init__(self, rc: int, what: str, err_bits: int = 0):
        name = errno.errorcode.get(-rc, str(rc))
        extra = ""
        if err_bits:
            extra = " [overflow: " + ", ".join(v for k, v in ERR_BITS.items() if err_bits & k) + "]"
        super().__init__(f"{what}: {name} ({rc}){extra}")
        self.rc = rc
        self.err_bits = err_bits


_LIB = None


def hip_lib():
    """Load libxkrt_hip.so (after torch, so both share torch's HIP runtime)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    try:
        import torch  # noqa: F401  -- one HIP runtime per process
    except Exception:
        pass
    path = lib_path("libxkrt_hip.so")
    if not os.path.exists(path):
        raise LibraryMissing(f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(path)
    L.xkrt_hip_init.argtypes = [C.c_int]
    L.xkrt_hip_dep_domain_create.argtypes = [_u64, _u64, C.POINTER(ConfigT), _p, C.POINTER(_p)]
    L.xkrt_hip_dep_domain_destroy.argtypes = [_p]
    L.xkrt_hip_dep_submit.argtypes = [_p, _p, _u64, C.POINTER(_u64)]
    L.xkrt_hip_dep_submit_device.argtypes = [_p, _p, _u64, _u64, _u32, _u32]
    L.xkrt_hip_dep_build_records.argtypes = [_p, _p, _u64, _u32, _u64, _p, C.POINTER(_u64), C.POINTER(_u32),
                                             C.POINTER(_u32)]
    L.xkrt_hip_pool_create.argtypes = [C.POINTER(PoolConfigT), _p, C.POINTER(_p)]
    L.xkrt_hip_pool_destroy.argtypes = [_p]
    L.xkrt_hip_pool_domain_create.argtypes = [_p, _u64, _u64, _u32, _u32, _u64, _u64, C.POINTER(_p)]
    L.xkrt_hip_seg_domain_create.argtypes = [_u64, C.POINTER(ConfigT), _p, C.POINTER(_p)]
    L.xkrt_hip_pool_seg_domain_create.argtypes = [_p, _u64, _u32, _u32, _u64, _u64, C.POINTER(_p)]
    L.xkrt_hip_dep_replay_segments.argtypes = [_p, _p, _u64]
    L.xkrt_hip_pool_stage.argtypes = [_p, _p, _p, _u64, C.POINTER(_u64)]
    L.xkrt_hip_pool_launch.argtypes = [_p, C.POINTER(_u64)]
    L.xkrt_hip_pool_poll.argtypes = [_p, C.POINTER(_u64)]
    L.xkrt_hip_pool_wait.argtypes = [_p, _u64]
    L.xkrt_hip_pool_open.argtypes = [_p, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64)]
    L.xkrt_hip_pool_sync.argtypes = [_p, C.POINTER(_u64)]
    L.xkrt_hip_pool_edges_copy.argtypes = [_p, _p, _p, _u64, C.POINTER(_u64)]
    L.xkrt_hip_pool_stats.argtypes = [_p, C.POINTER(PoolStatsT)]
    L.xkrt_hip_pool_timing.argtypes = [_p, C.c_int]
    L.xkrt_hip_pool_launch_times.argtypes = [_p, _p, _u64, C.POINTER(_u64), C.c_int]
    L.xkrt_hip_pool_stream.argtypes = [_p]
    L.xkrt_hip_pool_stream.restype = _p
    L.xkrt_hip_dep_index.argtypes = [_p, C.POINTER(_u32)]
    L.xkrt_hip_dep_frame.argtypes = [_p, C.POINTER(_u64)]
    L.xkrt_hip_dep_sync.argtypes = [_p, C.POINTER(_u64)]
    L.xkrt_hip_dep_edges_device.argtypes = [_p, C.POINTER(_p), C.POINTER(_u64)]
    L.xkrt_hip_dep_edges_copy.argtypes = [_p, _p, _u64, C.POINTER(_u64)]
    L.xkrt_hip_dep_edges_consume.argtypes = [_p]
    L.xkrt_hip_dep_stats.argtypes = [_p, C.POINTER(StatsT)]
    L.xkrt_hip_dep_stream.argtypes = [_p]
    L.xkrt_hip_dep_stream.restype = _p
    L.xkrt_hip_dep_reset.argtypes = [_p]
    L.xkrt_hip_dep_retire.argtypes = [_p, _u64]
    L.xkrt_hip_dep_id_base.argtypes = [_p, C.POINTER(_u64)]
    L.xkrt_hip_dep_timing.argtypes = [_p, C.c_int]
    L.xkrt_hip_dep_kernel_times.argtypes = [_p, C.POINTER(C.c_double), C.POINTER(_u64), C.c_int]
    L.xkrt_hip_dep_launch_times.argtypes = [_p, _p, _u64, C.POINTER(_u64), C.c_int]
    L.xkrt_hip_copy.argtypes = [_p, _p, _u64, _p]
    L.xkrt_hip_server_create.argtypes = [_p, C.POINTER(ServerConfigT), C.POINTER(_p)]
    L.xkrt_hip_server_destroy.argtypes = [_p]
    L.xkrt_hip_server_client.argtypes = [_p, _u64, _u64, _u32, _u32, _u64, _u64, C.POINTER(_u32)]
    L.xkrt_hip_server_submit.argtypes = [_p, _u32, _p, _u64, C.POINTER(_u64)]
    L.xkrt_hip_server_close.argtypes = [_p, _u32]
    L.xkrt_hip_server_drain.argtypes = [_p, _u64]
    L.xkrt_hip_server_times.argtypes = [_p, _u64, _u64, _p, _p, _p]
    L.xkrt_hip_server_edges.argtypes = [_p, _u32, _p, _u64, C.POINTER(_u64)]
    L.xkrt_hip_server_stats.argtypes = [_p, C.POINTER(ServerStatsT)]
    L.xkrt_hip_pool_stage_group.argtypes = [_p, _u32, _p, _p, _p, _p, _p]
    L.xkrt_hip_pool_edges_device.argtypes = [_p, C.POINTER(_p), C.POINTER(_p), C.POINTER(_p)]
    L.xkrt_hip_pool_edges_consume.argtypes = [_p]
    L.xkrt_hip_taskdom_ready.argtypes = [_p, _p, _u64, C.POINTER(_u64)]
    L.xkrt_hip_handle_domain_create.argtypes = [C.POINTER(HandleConfigT), _p, C.POINTER(_p)]
    L.xkrt_hip_handle_domain_destroy.argtypes = [_p]
    L.xkrt_hip_handle_submit.argtypes = [_p, _p, _u64, C.POINTER(_u64)]
    L.xkrt_hip_handle_sync.argtypes = [_p, C.POINTER(_u64), C.POINTER(_u64)]
    L.xkrt_hip_handle_edges_copy.argtypes = [_p, _p, _u64, C.POINTER(_u64)]
    L.xkrt_hip_handle_submit_ids.argtypes = [_p, _p, _p, _u64]
    L.xkrt_hip_handle_edges_device.argtypes = [_p, C.POINTER(_p), C.POINTER(_p), C.POINTER(_u32), C.POINTER(_u64),
                                               C.POINTER(_p)]
    L.xkrt_hip_handle_edges_consume.argtypes = [_p]
    L.xkrt_hip_taskdom_complete.argtypes = [_p, _p, _u64]
    L.xkrt_hip_taskdom_complete_wait.argtypes = [_p, _p, _u64]
    L.xkrt_hip_taskdom_create.argtypes = [C.POINTER(TaskDomConfigT), _p, C.POINTER(_p)]
    L.xkrt_hip_taskdom_destroy.argtypes = [_p]
    L.xkrt_hip_taskdom_submit.argtypes = [_p, _p, _u64, C.POINTER(_u64)]
    L.xkrt_hip_taskdom_sync.argtypes = [_p, C.POINTER(_u64)]
    L.xkrt_hip_taskdom_edges_copy.argtypes = [_p, _p, _u64, C.POINTER(_u64)]
    L.xkrt_hip_taskdom_stats.argtypes = [_p, C.POINTER(TaskDomStatsT)]
    _LIB = L
    return L


def _chk(rc: int, what: str):
    if rc != 0:
        raise ResolverError(rc, what)


def device_copy(dst_ptr: int, src_ptr: int, nbytes: int, stream: int = 0):
    """Streaming device copy (the calibration kernel for HBM byte counters)."""
    _chk(hip_lib().xkrt_hip_copy(dst_ptr, src_ptr, nbytes, stream or None), "xkrt_hip_copy")


def to_access_array(task, mode, addr, m, n, kind=None) -> np.ndarray:
    """Accesses for submit: tiles (addr, m rows, n columns) or, where kind is
    KIND_SEGMENT, byte segments [addr, addr + m)."""
    a = np.zeros(len(task), ACCESS_DTYPE)
    a["task"] = task
    a["mode"] = mode
    a["addr"] = addr
    a["m"] = m
    a["n"] = n
    if kind is not None:
        a["kind"] = kind
    return a


class DependencyDomain:
    """A GPU-resident BLAS dependency domain for one (ld, sizeof_type)."""

    def __init__(self, ld: int, sizeof_type: int, bin_x_bytes: int = 1024, bin_y_cols: int = 128,
                 kernel: int = 0, max_rects: int = 1 << 20, max_edges: int = 0, budget_bytes: int = 0,
                 y_hint: tuple[int, int] | None = None, stream: int | None = None, device: int = 0,
                 _segments: bool = False):
        self.L = hip_lib()
        _chk(self.L.xkrt_hip_init(device), "xkrt_hip_init")
        cfg = ConfigT(bin_x_bytes, bin_y_cols, kernel, 0, max_rects, max_edges, budget_bytes,
                      y_hint[0] if y_hint else 0, y_hint[1] if y_hint else 0)
        h = _p()
        if _segments:
            _chk(self.L.xkrt_hip_seg_domain_create(ld, C.byref(cfg), stream, C.byref(h)),
                 "xkrt_hip_seg_domain_create")
        else:
            _chk(self.L.xkrt_hip_dep_domain_create(ld, sizeof_type, C.byref(cfg), stream, C.byref(h)),
                 "xkrt_hip_dep_domain_create")
        self.h = h
        self.ld, self.s = ld, sizeof_type
        self.max_rects = max_rects

    def replay_segments(self, acc: np.ndarray):
        """Put uncompleted segment accesses into this (new) BLAS domain, with no
        LINK (src/task/task.cc:177-198): no edge changes; the inserted rects'
        extent is recorded for the wrapped-tile check."""
        acc = np.ascontiguousarray(acc, ACCESS_DTYPE)
        _chk(self.L.xkrt_hip_dep_replay_segments(self.h, acc.ctypes.data, len(acc)), "xkrt_hip_dep_replay_segments")

    def close(self):
        if getattr(self, "h", None):
            self.L.xkrt_hip_dep_domain_destroy(self.h)
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

    @property
    def stream(self) -> int:
        return int(self.L.xkrt_hip_dep_stream(self.h) or 0)

    def submit(self, acc: np.ndarray) -> int:
        """Enqueue a batch of accesses (ACCESS_DTYPE); returns the first access id."""
        acc = np.ascontiguousarray(acc, ACCESS_DTYPE)
        first = _u64()
        _chk(self.L.xkrt_hip_dep_submit(self.h, acc.ctypes.data, len(acc), C.byref(first)), "xkrt_hip_dep_submit")
        return int(first.value)

    def build_records(self, acc: np.ndarray, first_id: int):
        """Host conversion to device-frame rect records (RECT_DTYPE), ylo, yhi."""
        acc = np.ascontiguousarray(acc, ACCESS_DTYPE)
        cap = 3 * len(acc)  # a segment gives up to 3 rects, a tile up to 2
        out = np.zeros(max(cap, 1), RECT_DTYPE)
        nrect, ylo, yhi = _u64(), _u32(), _u32()
        _chk(self.L.xkrt_hip_dep_build_records(self.h, acc.ctypes.data, len(acc), first_id, cap, out.ctypes.data,
                                               C.byref(nrect), C.byref(ylo), C.byref(yhi)),
             "xkrt_hip_dep_build_records")
        return out[:int(nrect.value)], int(ylo.value), int(yhi.value)

    def submit_device(self, ptr: int, nrect: int, naccess: int, ylo: int, yhi: int):
        """Enqueue device-resident rect records (a device pointer to RECT_DTYPE records)."""
        _chk(self.L.xkrt_hip_dep_submit_device(self.h, ptr, nrect, naccess, ylo, yhi), "xkrt_hip_dep_submit_device")

    @property
    def index(self) -> int:
        i = _u32()
        _chk(self.L.xkrt_hip_dep_index(self.h, C.byref(i)), "xkrt_hip_dep_index")
        return int(i.value)

    def sync(self) -> int:
        n = _u64()
        rc = self.L.xkrt_hip_dep_sync(self.h, C.byref(n))
        if rc != 0:
            st = StatsT()
            self.L.xkrt_hip_dep_stats(self.h, C.byref(st))
            raise ResolverError(rc, "xkrt_hip_dep_sync", int(st.err))
        return int(n.value)

    def edges(self, out: np.ndarray | None = None) -> np.ndarray:
        """Copy and consume every unconsumed edge: (E, 2) uint32 (pred id, succ id).
        out: a C-contiguous (cap, 2) uint32 buffer to copy into (pinned host memory
        takes the copy at DMA speed); the result is its first E rows."""
        n = self.sync()
        if out is None:
            out = np.empty((max(n, 1), 2), np.uint32)
        elif out.dtype != np.uint32 or out.ndim != 2 or out.shape[1] != 2 or not out.flags.c_contiguous \
                or out.shape[0] < n:
            raise ValueError(f"edges: out must be a C-contiguous (>= {n}, 2) uint32 array")
        got = _u64()
        _chk(self.L.xkrt_hip_dep_edges_copy(self.h, out.ctypes.data, n, C.byref(got)), "xkrt_hip_dep_edges_copy")
        return out[:n]

    def edges_device(self):
        ptr, n = _p(), _u64()
        _chk(self.L.xkrt_hip_dep_edges_device(self.h, C.byref(ptr), C.byref(n)), "xkrt_hip_dep_edges_device")
        return int(ptr.value or 0), int(n.value)

    def consume(self):
        _chk(self.L.xkrt_hip_dep_edges_consume(self.h), "xkrt_hip_dep_edges_consume")

    def reset(self):
        """Forget every access (asynchronous); the grid and frame stay."""
        _chk(self.L.xkrt_hip_dep_reset(self.h), "xkrt_hip_dep_reset")

    def retire(self, horizon: int):
        """The accesses with id < horizon completed: forget them and rebase the ids
        (ids reported later + id_base = ids over the domain's life).  The edges
        must have been read first."""
        _chk(self.L.xkrt_hip_dep_retire(self.h, horizon), "xkrt_hip_dep_retire")

    @property
    def id_base(self) -> int:
        b = _u64()
        _chk(self.L.xkrt_hip_dep_id_base(self.h, C.byref(b)), "xkrt_hip_dep_id_base")
        return int(b.value)

    def set_timing(self, on: bool = True):
        _chk(self.L.xkrt_hip_dep_timing(self.h, 1 if on else 0), "xkrt_hip_dep_timing")

    # the library's six timing slots; slots 1 and 2 read 0 (kept for the layout)
    KERNELS = ("k_count", "-", "--", "k_resolve", "k_dedup", "k_final")

    def kernel_times(self, clear: bool = True):
        """({kernel: total ms}, timed launches) from HIP events on the domain's stream."""
        ms = (C.c_double * 6)()
        n = _u64()
        _chk(self.L.xkrt_hip_dep_kernel_times(self.h, ms, C.byref(n), 1 if clear else 0), "xkrt_hip_dep_kernel_times")
        return {k: float(ms[i]) for i, k in enumerate(self.KERNELS)}, int(n.value)

    def launch_times(self, clear: bool = True) -> np.ndarray:
        """Per timed launch: ms from its first kernel start to its last kernel end."""
        n = _u64()
        _chk(self.L.xkrt_hip_dep_launch_times(self.h, None, 0, C.byref(n), 0), "xkrt_hip_dep_launch_times")
        out = np.zeros(max(int(n.value), 1), np.float32)
        _chk(self.L.xkrt_hip_dep_launch_times(self.h, out.ctypes.data, len(out), C.byref(n), 1 if clear else 0),
             "xkrt_hip_dep_launch_times")
        return out[:int(n.value)]

    def stats(self) -> dict:
        st = StatsT()
        _chk(self.L.xkrt_hip_dep_stats(self.h, C.byref(st)), "xkrt_hip_dep_stats")
        return {f: int(getattr(st, f)) for f, _ in StatsT._fields_}


class SegmentDomain(DependencyDomain):
    """A GPU-resident segment dependency domain (the interval tree of
    interval/dependency-tree.hpp:187-363): byte segments [addr, addr + m) only,
    cut into a frame of col_bytes-byte columns."""

    def __init__(self, col_bytes: int = 1 << 16, bin_x_bytes: int = 4096, bin_y_cols: int = 1, **kw):
        super().__init__(col_bytes, 1, bin_x_bytes, bin_y_cols, _segments=True, **kw)

    def replay_segments(self, acc):
        raise ResolverError(-22, "replay_segments: a segment domain takes no replay")


class PoolDomain(DependencyDomain):
    """A domain admitted into a Pool (its bins live in the pool's arrays)."""

    def __init__(self, pool: "Pool", ld: int, sizeof_type: int, y_range: tuple[int, int], bin_x_bytes: int = 1024,
                 bin_y_cols: int = 128, segments: bool = False):
        self.L = hip_lib()
        h = _p()
        if segments:
            _chk(self.L.xkrt_hip_pool_seg_domain_create(pool.h, ld, bin_x_bytes, bin_y_cols, int(y_range[0]),
                                                        int(y_range[1]), C.byref(h)), "xkrt_hip_pool_seg_domain_create")
        else:
            _chk(self.L.xkrt_hip_pool_domain_create(pool.h, ld, sizeof_type, bin_x_bytes, bin_y_cols,
                                                    int(y_range[0]), int(y_range[1]), C.byref(h)),
                 "xkrt_hip_pool_domain_create")
        self.h = h
        self.pool = pool
        self.ld, self.s = ld, sizeof_type
        pool._domains.add(self)

    def close(self):
        # a destroyed pool has already released its domains
        if getattr(self, "h", None) and getattr(self.pool, "h", None):
            self.L.xkrt_hip_dep_domain_destroy(self.h)
        self.h = None


class Pool:
    """Many dependency domains resident on the device, resolved together: batches
    of any domains are staged and resolved by one launch (optionally a replayed
    hipGraph); the device-memory budget decides which domains are admitted."""

    def __init__(self, kernel: int = 1, graph: bool = False, max_rects: int = 1 << 20, max_edges: int = 0,
                 budget_bytes: int = 0, stream: int | None = None, device: int = 0, direct_dedup: bool = False,
                 small_launch_rects: int = 0, host_edges: bool = False):
        self.L = hip_lib()
        _chk(self.L.xkrt_hip_init(device), "xkrt_hip_init")
        cfg = PoolConfigT(kernel, 1 if graph else 0, max_rects, max_edges, budget_bytes, 1 if direct_dedup else 0,
                          small_launch_rects, 1 if host_edges else 0, 0)
        h = _p()
        _chk(self.L.xkrt_hip_pool_create(C.byref(cfg), stream, C.byref(h)), "xkrt_hip_pool_create")
        self.h = h
        self.max_rects = max_rects
        self._domains = weakref.WeakSet()

    def close(self):
        if getattr(self, "h", None):
            # the pool releases its domains: their handles die with it (later calls raise)
            for d in list(self._domains):
                d.h = None
            self.L.xkrt_hip_pool_destroy(self.h)
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

    def domain(self, ld: int, sizeof_type: int, y_range, bin_x_bytes: int = 1024, bin_y_cols: int = 128):
        """Admit a domain; raises ResolverError(-ENOMEM) when the budget cannot hold it now."""
        return PoolDomain(self, ld, sizeof_type, y_range, bin_x_bytes, bin_y_cols)

    def segment_domain(self, col_bytes: int, y_range, bin_x_bytes: int = 4096, bin_y_cols: int = 1):
        """Admit a segment domain over addresses [y_range[0], y_range[1]) * col_bytes."""
        return PoolDomain(self, col_bytes, 1, y_range, bin_x_bytes, bin_y_cols, segments=True)

    def leave(self, dom: PoolDomain):
        """The domain's client is done: its bins return to the pool."""
        dom.close()

    def stage_group(self, doms, accs):
        """Stage one batch per domain as one unit (all or none); returns the first ids."""
        nd = len(doms)
        accs = [np.ascontiguousarray(a, ACCESS_DTYPE) for a in accs]
        dh = (_p * max(nd, 1))(*[d.h for d in doms])
        ap = (_p * max(nd, 1))(*[a.ctypes.data for a in accs])
        cn = (_u64 * max(nd, 1))(*[len(a) for a in accs])
        first = (_u64 * max(nd, 1))()
        _chk(self.L.xkrt_hip_pool_stage_group(self.h, nd, dh, ap, None, cn, first), "xkrt_hip_pool_stage_group")
        return [int(first[k]) for k in range(nd)]

    def stage(self, dom: PoolDomain, acc: np.ndarray) -> int:
        acc = np.ascontiguousarray(acc, ACCESS_DTYPE)
        first = _u64()
        _chk(self.L.xkrt_hip_pool_stage(self.h, dom.h, acc.ctypes.data, len(acc), C.byref(first)),
             "xkrt_hip_pool_stage")
        return int(first.value)

    def launch(self) -> int:
        """Resolve every staged batch in one launch; returns its completion ticket."""
        t = _u64()
        _chk(self.L.xkrt_hip_pool_launch(self.h, C.byref(t)), "xkrt_hip_pool_launch")
        return int(t.value)

    def poll(self) -> int:
        """Highest ticket whose launch (and every earlier one) has completed."""
        d = _u64()
        _chk(self.L.xkrt_hip_pool_poll(self.h, C.byref(d)), "xkrt_hip_pool_poll")
        return int(d.value)

    def wait(self, ticket: int):
        _chk(self.L.xkrt_hip_pool_wait(self.h, ticket), "xkrt_hip_pool_wait")

    def open(self):
        r, a, b = _u64(), _u64(), _u64()
        _chk(self.L.xkrt_hip_pool_open(self.h, C.byref(r), C.byref(a), C.byref(b)), "xkrt_hip_pool_open")
        return int(r.value), int(a.value), int(b.value)

    def sync(self) -> int:
        n = _u64()
        rc = self.L.xkrt_hip_pool_sync(self.h, C.byref(n))
        if rc != 0:
            raise ResolverError(rc, "xkrt_hip_pool_sync")
        return int(n.value)

    def edges(self):
        """Copy and consume every unconsumed edge: ((E, 2) pred/succ ids, (E,) domain index)."""
        n = self.sync()
        e = np.empty((max(n, 1), 2), np.uint32)
        d = np.empty(max(n, 1), np.uint32)
        got = _u64()
        _chk(self.L.xkrt_hip_pool_edges_copy(self.h, e.ctypes.data, d.ctypes.data, n, C.byref(got)),
             "xkrt_hip_pool_edges_copy")
        return e[:n], d[:n]

    def stats(self) -> dict:
        st = PoolStatsT()
        _chk(self.L.xkrt_hip_pool_stats(self.h, C.byref(st)), "xkrt_hip_pool_stats")
        return {f: int(getattr(st, f)) for f, _ in PoolStatsT._fields_}

    def set_timing(self, on: bool = True):
        _chk(self.L.xkrt_hip_pool_timing(self.h, 1 if on else 0), "xkrt_hip_pool_timing")

    def launch_times(self, clear: bool = True) -> np.ndarray:
        n = _u64()
        _chk(self.L.xkrt_hip_pool_launch_times(self.h, None, 0, C.byref(n), 0), "xkrt_hip_pool_launch_times")
        out = np.zeros(max(int(n.value), 1), np.float32)
        _chk(self.L.xkrt_hip_pool_launch_times(self.h, out.ctypes.data, len(out), C.byref(n), 1 if clear else 0),
             "xkrt_hip_pool_launch_times")
        return out[:int(n.value)]

    @property
    def stream(self) -> int:
        return int(self.L.xkrt_hip_pool_stream(self.h) or 0)


class Server:
    """Serving engine over a Pool: a native host thread admits clients against
    the pool's budget, batches every arrived batch into the next launch while
    fewer than `depth` launches are in flight (max_batches=1: one batch per
    launch), and stamps submit and ready times.  Clients are small ints."""

    def __init__(self, pool: Pool, depth: int = 2, max_batches: int = 0, max_rects: int = 0,
                 reserve_bytes: int = 0, prearm: bool = False):
        self.L = hip_lib()
        self.pool = pool
        cfg = ServerConfigT(depth, max_batches, max_rects, reserve_bytes, 1 if prearm else 0, 0)
        h = _p()
        _chk(self.L.xkrt_hip_server_create(pool.h, C.byref(cfg), C.byref(h)), "xkrt_hip_server_create")
        self.h = h
        self.nreq = 0

    def close(self):
        if getattr(self, "h", None):
            rc = self.L.xkrt_hip_server_destroy(self.h)
            self.h = None
            if rc:
                raise ResolverError(rc, "xkrt_hip_server_destroy")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def client(self, ld: int, sizeof_type: int, y_range, bin_x_bytes: int = 1024, bin_y_cols: int = 128) -> int:
        c = _u32()
        _chk(self.L.xkrt_hip_server_client(self.h, ld, sizeof_type, bin_x_bytes, bin_y_cols, int(y_range[0]),
                                           int(y_range[1]), C.byref(c)), "xkrt_hip_server_client")
        return int(c.value)

    def submit(self, client: int, acc: np.ndarray) -> int:
        acc = np.ascontiguousarray(acc, ACCESS_DTYPE)
        r = _u64()
        _chk(self.L.xkrt_hip_server_submit(self.h, client, acc.ctypes.data, len(acc), C.byref(r)),
             "xkrt_hip_server_submit")
        self.nreq = int(r.value) + 1
        return int(r.value)

    def close_client(self, client: int):
        _chk(self.L.xkrt_hip_server_close(self.h, client), "xkrt_hip_server_close")

    def open_loop(self, t_s, clients, batches, last) -> np.ndarray:
        """Submits batch i (an ACCESS_DTYPE array) of client clients[i] at t_s[i] seconds from
        the call, and closes the client after it when last[i], from native code (no Python
        between submissions); returns how late each submit started, in seconds."""
        from .streams import _lib as streams_lib
        n = len(t_s)
        t = np.ascontiguousarray(t_s, np.float64)
        cl = np.ascontiguousarray(clients, np.uint32)
        bs = [np.ascontiguousarray(b, ACCESS_DTYPE) for b in batches]
        ptrs = (C.c_void_p * max(n, 1))(*[b.ctypes.data for b in bs])
        cnt = np.array([len(b) for b in bs], np.uint64)
        lst = np.ascontiguousarray(last, np.uint8)
        lag = np.zeros(max(n, 1), np.float64)
        req = np.zeros(max(n, 1), np.uint64)
        lib = streams_lib()
        rc = lib.xks_open_loop(self.h, C.cast(self.L.xkrt_hip_server_submit, C.c_void_p),
                               C.cast(self.L.xkrt_hip_server_close, C.c_void_p), n, t.ctypes.data,
                               cl.ctypes.data, ptrs, cnt.ctypes.data, lst.ctypes.data, lag.ctypes.data,
                               req.ctypes.data)
        _chk(rc, "xks_open_loop")
        if n:
            self.nreq = int(req[n - 1]) + 1
        return lag[:n]

    def drain(self, timeout_s: float = 0.0):
        _chk(self.L.xkrt_hip_server_drain(self.h, int(timeout_s * 1e6)), "xkrt_hip_server_drain")

    def times(self, first: int = 0, n: int | None = None):
        """(submit_us, ready_us, status) per request."""
        n = self.nreq - first if n is None else n
        t0 = np.zeros(max(n, 1), np.float64)
        t1 = np.zeros(max(n, 1), np.float64)
        st = np.zeros(max(n, 1), np.int32)
        _chk(self.L.xkrt_hip_server_times(self.h, first, n, t0.ctypes.data, t1.ctypes.data, st.ctypes.data),
             "xkrt_hip_server_times")
        return t0[:n], t1[:n], st[:n]

    def edges(self, client: int) -> np.ndarray:
        """The client's edges: (E, 2) uint32 (pred, succ) in its own access ids."""
        n = _u64()
        _chk(self.L.xkrt_hip_server_edges(self.h, client, None, 0, C.byref(n)), "xkrt_hip_server_edges")
        out = np.empty((max(int(n.value), 1), 2), np.uint32)
        _chk(self.L.xkrt_hip_server_edges(self.h, client, out.ctypes.data, len(out), C.byref(n)),
             "xkrt_hip_server_edges")
        return out[:int(n.value)]

    def stats(self) -> dict:
        st = ServerStatsT()
        _chk(self.L.xkrt_hip_server_stats(self.h, C.byref(st)), "xkrt_hip_server_stats")
        return {f: int(getattr(st, f)) for f, _ in ServerStatsT._fields_}


def to_task_access_array(task, mode, kind, addr, m, n, ld, sizeof_type, concurrency=None) -> np.ndarray:
    """Task-domain accesses: tiles (kind 0: addr, m rows, n columns, ld, sizeof_type),
    segments (kind 1: [addr, addr + m)) and handles (kind 2: addr is the handle,
    with a concurrency)."""
    a = np.zeros(len(task), TASK_ACCESS_DTYPE)
    a["task"], a["mode"], a["kind"] = task, mode, kind
    a["addr"], a["m"], a["n"] = addr, m, n
    a["ld"], a["sizeof_type"] = ld, sizeof_type
    if concurrency is not None:
        a["concurrency"] = concurrency
    return a


class TaskDomain:
    """One task domain's dependency resolution over segments and tiles of any
    (ld, sizeof_type) -- the dispatch of task_dependency_resolve
    (src/task/task.cc:204-299) over a segment domain and one BLAS domain per
    pair, all in one pool.  Access ids are program-wide."""

    def __init__(self, kernel: int = 1, max_rects: int = 1 << 20, max_edges: int = 1 << 22, budget_bytes: int = 0,
                 seg_col_bytes: int = 1 << 16, seg_bin: tuple[int, int] = (0, 0), blas_bin: tuple[int, int] = (1024, 128),
                 release: bool = False, max_tasks: int = 0, max_accesses: int = 0, max_succ: int = 0,
                 max_virtual: int = 0, stream: int | None = None, device: int = 0):
        self.L = hip_lib()
        _chk(self.L.xkrt_hip_init(device), "xkrt_hip_init")
        cfg = TaskDomConfigT(kernel, 1 if release else 0, max_rects, max_edges, budget_bytes, seg_col_bytes,
                             seg_bin[0], seg_bin[1], blas_bin[0], blas_bin[1], max_tasks, max_accesses, max_succ,
                             max_virtual)
        h = _p()
        _chk(self.L.xkrt_hip_taskdom_create(C.byref(cfg), stream, C.byref(h)), "xkrt_hip_taskdom_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.xkrt_hip_taskdom_destroy(self.h)
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

    def submit(self, acc: np.ndarray) -> int:
        acc = np.ascontiguousarray(acc, TASK_ACCESS_DTYPE)
        first = _u64()
        _chk(self.L.xkrt_hip_taskdom_submit(self.h, acc.ctypes.data, len(acc), C.byref(first)),
             "xkrt_hip_taskdom_submit")
        return int(first.value)

    def sync(self) -> int:
        n = _u64()
        _chk(self.L.xkrt_hip_taskdom_sync(self.h, C.byref(n)), "xkrt_hip_taskdom_sync")
        return int(n.value)

    def edges(self) -> np.ndarray:
        """Sync, then copy out and consume the waiting edges: (E, 2) uint64 program ids."""
        n = self.sync()
        out = np.empty((max(n, 1), 2), np.uint64)
        got = _u64()
        _chk(self.L.xkrt_hip_taskdom_edges_copy(self.h, out.ctypes.data, n, C.byref(got)),
             "xkrt_hip_taskdom_edges_copy")
        return out[:int(got.value)]

    def ready(self, cap: int = 1 << 16) -> np.ndarray:
        """Released tasks not read yet (non-blocking), in release order."""
        buf = getattr(self, "_rbuf", None)
        if buf is None or len(buf) < cap:
            buf = self._rbuf = np.empty(max(cap, 1), np.uint32)
            self._rn = _u64()
        _chk(self.L.xkrt_hip_taskdom_ready(self.h, buf.ctypes.data, cap, C.byref(self._rn)), "xkrt_hip_taskdom_ready")
        return buf[:int(self._rn.value)].copy()

    def complete(self, tasks, wait: bool = False) -> None:
        """Completes released tasks.  wait: return once their successors are
        released (as __task_complete, src/task.cc:160-247), so the next ready()
        sees them; otherwise the completion is only enqueued."""
        t = np.ascontiguousarray(tasks, np.uint32)
        if wait:
            _chk(self.L.xkrt_hip_taskdom_complete_wait(self.h, t.ctypes.data, len(t)),
                 "xkrt_hip_taskdom_complete_wait")
        else:
            _chk(self.L.xkrt_hip_taskdom_complete(self.h, t.ctypes.data, len(t)), "xkrt_hip_taskdom_complete")

    def stats(self) -> dict:
        st = TaskDomStatsT()
        _chk(self.L.xkrt_hip_taskdom_stats(self.h, C.byref(st)), "xkrt_hip_taskdom_stats")
        return {f: int(getattr(st, f)) for f, _ in TaskDomStatsT._fields_}


def to_handle_access_array(task, handle, mode, concurrency) -> np.ndarray:
    a = np.zeros(len(task), HANDLE_ACCESS_DTYPE)
    a["task"], a["handle"], a["mode"], a["concurrency"] = task, handle, mode, concurrency
    return a


class HandleDomain:
    """A GPU handle dependency domain (the reference's DependencyMap,
    include/xkrt/memory/access/handle/dependency-map.hpp:52-330): ids
    consecutive per submit; inserted empty writes get ids VIRTUAL_ID | k."""

    def __init__(self, max_handles: int = 1 << 16, max_accesses: int = 1 << 22, max_virtual: int = 1 << 20,
                 max_batch: int = 1 << 18, max_edges: int = 1 << 22, stream: int | None = None, device: int = 0):
        self.L = hip_lib()
        _chk(self.L.xkrt_hip_init(device), "xkrt_hip_init")
        cfg = HandleConfigT(max_handles, max_accesses, max_virtual, max_batch, max_edges)
        h = _p()
        _chk(self.L.xkrt_hip_handle_domain_create(C.byref(cfg), stream, C.byref(h)), "xkrt_hip_handle_domain_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.xkrt_hip_handle_domain_destroy(self.h)
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

    def submit(self, acc: np.ndarray) -> int:
        acc = np.ascontiguousarray(acc, HANDLE_ACCESS_DTYPE)
        first = _u64()
        _chk(self.L.xkrt_hip_handle_submit(self.h, acc.ctypes.data, len(acc), C.byref(first)), "xkrt_hip_handle_submit")
        return int(first.value)

    def sync(self):
        ne, nv = _u64(), _u64()
        _chk(self.L.xkrt_hip_handle_sync(self.h, C.byref(ne), C.byref(nv)), "xkrt_hip_handle_sync")
        return int(ne.value), int(nv.value)

    def edges(self) -> np.ndarray:
        n, _ = self.sync()
        out = np.empty((max(n, 1), 2), np.uint32)
        got = _u64()
        _chk(self.L.xkrt_hip_handle_edges_copy(self.h, out.ctypes.data, n, C.byref(got)), "xkrt_hip_handle_edges_copy")
        return out[:int(got.value)]


def batches_by_tasks(task: np.ndarray, tasks_per_batch) -> list[tuple[int, int]]:
    """Split access indices into batches of whole tasks.  tasks_per_batch is an
    int (fixed) or an array of batch sizes in tasks."""
    starts = np.flatnonzero(np.r_[True, np.diff(task.astype(np.int64)) != 0])
    ends = np.r_[starts[1:], len(task)]
    out = []
    if np.isscalar(tasks_per_batch):
        sizes = []
        t = 0
        while t < len(starts):
            sizes.append(min(int(tasks_per_batch), len(starts) - t))
            t += sizes[-1]
    else:
        sizes = [int(x) for x in tasks_per_batch]
    t = 0
    for b in sizes:
        if t >= len(starts):
            break
        e = min(t + b, len(starts))
        out.append((int(starts[t]), int(ends[e - 1])))
        t = e
    return out


def resolve_stream(stream, tasks_per_batch=4096, **domain_kw) -> np.ndarray:
    """Resolve a whole single-domain stream on the GPU, one launch per batch;
    returns (pred, succ) access-index edges (stream positions)."""
    ld, s = int(stream.ld[0]), int(stream.s[0])
    acc = to_access_array(stream.task, stream.mode, stream.addr, stream.m, stream.n)
    with DependencyDomain(ld, s, **domain_kw) as dom:
        first = None
        for (a, b) in batches_by_tasks(stream.task, tasks_per_batch):
            f = dom.submit(acc[a:b])
            if first is None:
                first = f
        e = dom.edges()
    return e.astype(np.int64) - (first or 0)
