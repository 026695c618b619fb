"""Seeded task/access streams (host side, numpy) for the resolver workloads.

A stream is a spawn-ordered list of column-major tile accesses, i.e. the
arguments of the reference's BLAS access constructor
(include/xkrt/memory/access/access.hpp:539-574): ``addr, ld, m, n, s`` and the
access ``mode`` (include/xkrt/memory/access/mode.h:56-62), plus the spawning
``task`` index, the ``slot`` of the access in its task and, for coherency
streams, the fetching device ``dev`` (0 = host, 1 = GPU).

The generators live in ``csrc/streams.c`` (SplitMix64, explicit integer
transforms) so that every consumer sees identical streams.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from ._loader import load_host_lib

MODE_R = 1
MODE_W = 2
MODE_RW = 3
MODE_V = 4
MODE_VW = 6
MODE_VR = 5
MODE_D = 8

_u64 = C.c_uint64
_p = C.c_void_p


def _lib():
    lib = load_host_lib("libxkstreams.so")
    if not getattr(lib, "_xks_typed", False):
        sig = [_u64, _p, _p, _p, _p, _p, _p, _p, _p, _p]
        lib.xks_cholesky.argtypes = [_u64, _u64, _u64, _u64, C.c_uint32] + sig
        lib.xks_cholesky.restype = _u64
        lib.xks_lu.argtypes = [_u64, _u64, _u64, _u64, C.c_uint32] + sig
        lib.xks_lu.restype = _u64
        lib.xks_open_loop.argtypes = [_p, _p, _p, _u64, _p, _p, _p, _p, _p, _p, _p]
        lib.xks_open_loop.restype = C.c_int
        lib.xks_random2d.argtypes = [_u64, _u64, _u64, _u64, _u64, _u64, C.c_double, C.c_double, _u64] + sig
        lib.xks_random2d.restype = _u64
        lib.xks_heat.argtypes = [_u64, _u64, _u64, _u64, _u64, C.c_uint32] + sig
        lib.xks_heat.restype = _u64
        lib.xks_coherency.argtypes = [_u64, _u64, _u64, _u64, _u64, _u64, _u64, _u64,
                                      C.c_double, C.c_double, C.c_double, _u64] + sig
        lib.xks_coherency.restype = _u64
        lib.xks_geometric_batches.argtypes = [_u64, C.c_double, _u64, _u64, _u64, _p]
        lib.xks_geometric_batches.restype = _u64
        lib.xks_clients.argtypes = [_u64, _u64, _u64, _u64, _p, _p]
        lib.xks_clients.restype = None
        lib.xks_matrix_to_rects.argtypes = [_u64, _p, _p, _p, _p, _p, _p, _p]
        lib.xks_matrix_to_rects.restype = None
        lib._xks_typed = True
    return lib


@dataclass
class Stream:
    """Spawn-ordered access records (structure of arrays)."""

    task: np.ndarray  # uint32, non-decreasing
    slot: np.ndarray  # uint16
    mode: np.ndarray  # uint8
    dev: np.ndarray  # uint8
    addr: np.ndarray  # uint64
    ld: np.ndarray  # uint64
    m: np.ndarray  # uint64
    n: np.ndarray  # uint64
    s: np.ndarray  # uint64
    name: str = ""
    batches: np.ndarray | None = field(default=None)  # tasks per submission batch

    def __len__(self) -> int:
        return int(self.task.shape[0])

    @property
    def ntasks(self) -> int:
        if len(self) == 0:
            return 0
        return int(np.count_nonzero(np.diff(self.task.astype(np.int64))) + 1)

    def rects(self):
        """(rects[n, 8] uint64, nrect[n] uint8): access.hpp:205-287 decomposition."""
        return matrix_to_rects(self.addr, self.ld, self.m, self.n, self.s)

    def take(self, idx) -> "Stream":
        return Stream(self.task[idx], self.slot[idx], self.mode[idx], self.dev[idx], self.addr[idx],
                      self.ld[idx], self.m[idx], self.n[idx], self.s[idx], name=self.name)

    def prefix_tasks(self, ntasks: int) -> "Stream":
        """The accesses of the first `ntasks` tasks."""
        if ntasks >= self.ntasks:
            return self
        starts = np.flatnonzero(np.r_[True, np.diff(self.task.astype(np.int64)) != 0])
        return self.take(slice(0, int(starts[ntasks])))

    def task_starts(self) -> np.ndarray:
        """Index of the first access of every task, plus len(self) at the end."""
        if len(self) == 0:
            return np.zeros(1, np.int64)
        st = np.flatnonzero(np.r_[True, np.diff(self.task.astype(np.int64)) != 0])
        return np.r_[st, len(self)].astype(np.int64)


def _alloc(count: int):
    return (np.empty(count, np.uint32), np.empty(count, np.uint16), np.empty(count, np.uint8),
            np.ones(count, np.uint8), np.empty(count, np.uint64), np.empty(count, np.uint64),
            np.empty(count, np.uint64), np.empty(count, np.uint64), np.empty(count, np.uint64))


def _ptrs(arrs):
    return [a.ctypes.data for a in arrs]


def _run(fn, args, name):
    nul = [None] * 9
    count = int(fn(*args, 0, *nul))
    arrs = _alloc(count)
    got = int(fn(*args, count, *_ptrs(arrs)))
    assert got == count
    return Stream(*arrs, name=name)


def cholesky(N: int, nb: int, s: int = 8, base: int = 0, task0: int = 0) -> Stream:
    """Tiled right-looking Cholesky, (N/nb)^2 tiles: POTRF/TRSM/SYRK/GEMM."""
    return _run(_lib().xks_cholesky, [N, nb, s, base, task0], f"cholesky N={N} nb={nb}")


def lu(N: int, nb: int, s: int = 8, base: int = 0, task0: int = 0) -> Stream:
    """Tiled right-looking LU without pivoting: GETRF/TRSM-L/TRSM-U/GEMM."""
    return _run(_lib().xks_lu, [N, nb, s, base, task0], f"lu N={N} nb={nb}")


def random2d(nacc: int, rows: int = 65536, cols: int = 65536, s: int = 4, maxsz: int = 4096,
             seed: int = 0x2B, pr: float = 0.6, pw: float = 0.2, base: int = 0) -> Stream:
    """Single-access tasks with uniform origins and log-uniform sizes in [1, maxsz]."""
    return _run(_lib().xks_random2d, [nacc, rows, cols, s, maxsz, seed, pr, pw, base],
                f"random2d n={nacc} {rows}x{cols}")


def heat(Ng: int = 16384, tile: int = 128, s: int = 4, steps: int = 200, base: int = 0, task0: int = 0) -> Stream:
    """2D heat halo stream: RW own tile + R neighbour halo strips, row-major tiles."""
    return _run(_lib().xks_heat, [Ng, tile, s, steps, base, task0], f"heat {Ng}^2 tile={tile} steps={steps}")


def coherency(nacc: int, ld0: int = 8192, s0: int = 8, ld1: int = 16384, s1: int = 4, cols: int = 8192,
              maxsz: int = 2048, seed: int = 0x4C, pdev_host: float = 0.5, pr: float = 0.5, pw: float = 0.25,
              base: int = 0) -> Stream:
    """Two views of one buffer, host/GPU placement, R/W/RW modes."""
    return _run(_lib().xks_coherency, [nacc, ld0, s0, ld1, s1, cols, maxsz, seed, pdev_host, pr, pw, base],
                f"coherency n={nacc}")


def geometric_batches(total: int, mean: float = 2048.0, maxb: int = 65536, seed: int = 0x3B) -> np.ndarray:
    lib = _lib()
    nb = int(lib.xks_geometric_batches(total, mean, maxb, seed, 0, None))
    out = np.empty(nb, np.uint64)
    lib.xks_geometric_batches(total, mean, maxb, seed, nb, out.ctypes.data)
    return out


def clients(nclients: int = 64, nmin: int = 4, nmax: int = 48, seed: int = 0x5C):
    kind = np.empty(nclients, np.uint8)
    nt = np.empty(nclients, np.uint32)
    _lib().xks_clients(nclients, nmin, nmax, seed, kind.ctypes.data, nt.ctypes.data)
    return kind, nt


def matrix_to_rects(addr, ld, m, n, s):
    addr = np.ascontiguousarray(addr, np.uint64)
    ld = np.ascontiguousarray(ld, np.uint64)
    m = np.ascontiguousarray(m, np.uint64)
    n = np.ascontiguousarray(n, np.uint64)
    s = np.ascontiguousarray(s, np.uint64)
    cnt = addr.shape[0]
    rects = np.empty((cnt, 8), np.uint64)
    nrect = np.empty(cnt, np.uint8)
    _lib().xks_matrix_to_rects(cnt, addr.ctypes.data, ld.ctypes.data, m.ctypes.data, n.ctypes.data,
                               s.ctypes.data, rects.ctypes.data, nrect.ctypes.data)
    return rects, nrect


def concat(streams, renumber: bool = True) -> Stream:
    """Concatenate streams; tasks are renumbered to stay increasing."""
    parts = []
    t0 = 0
    for st in streams:
        task = st.task.astype(np.int64)
        if renumber and len(st):
            _, inv = np.unique(task, return_inverse=True)
            task = inv + t0
            t0 = int(task[-1]) + 1
        parts.append((task.astype(np.uint32), st))
    return Stream(np.concatenate([p[0] for p in parts]),
                  *[np.concatenate([getattr(p[1], f) for p in parts]) for f in
                    ("slot", "mode", "dev", "addr", "ld", "m", "n", "s")])


def stencil1d(nchunks: int = 4096, chunk: int = 1 << 16, steps: int = 50, halo: int = 256,
              base: int = 1 << 40) -> dict:
    """A 1-D blocked stencil over byte segments: nchunks chunks of `chunk`
    bytes from `base`; per sweep, task i reads the last `halo` bytes of chunk
    i - 1 and the first `halo` bytes of chunk i + 1 (when they exist), then
    updates chunk i (R, R, RW segment accesses, in that slot order).
    Returns task, mode, addr, size (numpy arrays)."""
    task, mode, addr, size = [], [], [], []
    i = np.arange(nchunks, dtype=np.uint64)
    left = np.flatnonzero(i > 0)
    right = np.flatnonzero(i < nchunks - 1)
    for s in range(steps):
        t0 = s * nchunks
        parts = [(left, base + left.astype(np.uint64) * chunk - halo, halo, MODE_R, 0),
                 (right, base + (right.astype(np.uint64) + 1) * chunk, halo, MODE_R, 1),
                 (np.arange(nchunks), base + i * chunk, chunk, MODE_RW, 2)]
        tk, md, ad, sz, sl = [], [], [], [], []
        for idx, a, z, m, slot in parts:
            tk.append(t0 + np.asarray(idx, np.int64))
            ad.append(np.asarray(a, np.uint64))
            sz.append(np.full(len(idx), z, np.uint64))
            md.append(np.full(len(idx), m, np.uint8))
            sl.append(np.full(len(idx), slot, np.int64))
        tk, ad, sz, md, sl = (np.concatenate(x) for x in (tk, ad, sz, md, sl))
        order = np.lexsort((sl, tk))
        task.append(tk[order]); mode.append(md[order]); addr.append(ad[order]); size.append(sz[order])
    return dict(task=np.concatenate(task).astype(np.uint32), mode=np.concatenate(mode),
                addr=np.concatenate(addr), size=np.concatenate(size))


def handle_stream(n: int = 1 << 20, nhandles: int = 4096, pcnc: float = 0.3, per_task: int = 3,
                  seed: int = 0x4A) -> dict:
    """OpenMP-style dependencies on handles: tasks of 1..per_task accesses on
    handles drawn uniformly from nhandles (a task's later accesses shifted to
    other handles, so they rarely share one); modes R / W / RW, writes
    concurrent (inoutset) with probability pcnc, else sequential or
    commutative.  Returns task, handle, mode, concurrency."""
    rng = np.random.default_rng(seed)
    k = rng.integers(1, per_task + 1, n)
    task = np.repeat(np.arange(n, dtype=np.uint32), k)[:n]
    handle = (rng.integers(0, nhandles, n).astype(np.uint64) * np.uint64(64) + np.uint64(0x7F00_0000_0000))
    # distinct handles inside a task: shift repeats by the slot
    start = np.r_[0, np.flatnonzero(np.diff(task.astype(np.int64)) != 0) + 1]
    slot = np.arange(n) - np.repeat(start, np.diff(np.r_[start, n]))
    handle = handle + (slot.astype(np.uint64) * np.uint64(64 * 7919)) % np.uint64(64 * nhandles)
    mode = rng.choice(np.array([MODE_R, MODE_W, MODE_RW], np.uint8), n)
    conc = np.where(rng.random(n) < pcnc, 4, rng.choice(np.array([1, 2], np.uint8), n)).astype(np.uint8)
    conc = np.where(mode & MODE_W, conc, 1).astype(np.uint8)
    return dict(task=task, handle=handle, mode=mode, concurrency=conc)
