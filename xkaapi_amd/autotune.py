"""Autotuner for the resolver's bin geometry and LDS capacity class.

The resolver cuts a domain's frame into bins of BX bytes x BY columns and runs
one wave64 per active bin with the bin's state in LDS.  Small bins give more
parallelism but split tiles over several bins (more incidences, more
candidates); large bins hold more fragments and records (a larger, slower LDS
tier, or -ENOSPC).  The right point depends on the stream, so it is measured:

  1. candidates from the sample's geometry: BX near the median tile height in
     bytes, BY near the median tile width in columns, and their neighbours
     (halved and doubled in each dimension);
  2. the capacity class from the densest bin of a launch-sized batch;
  3. each candidate resolves the sample on the device (plain form, HIP-event
     timed); failures (a capacity overflow) disqualify it; the fastest wins.

The batch size per launch of the batched form is chosen the same way from a
list of candidates.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import resolver as R

KCLASS_RECORDS = (256, 512, 1024)   # records per bin and launch of each LDS class
KCLASS_FRAGS = (64, 256, 1024)


def _pow2_floor(v: int) -> int:
    v = max(int(v), 1)
    return 1 << (v.bit_length() - 1)


def _pow2_ceil(v: int) -> int:
    v = max(int(v), 1)
    return 1 << (v - 1).bit_length()


@dataclass
class TuneResult:
    bin_x_bytes: int
    bin_y_cols: int
    kernel: int
    us_per_access: float
    tried: list


def candidates(stream, max_cands: int = 7):
    """Bin extents to try: (BX bytes, BY columns), powers of two."""
    ld, s = int(stream.ld[0]), int(stream.s[0])
    lds = ld * s
    xb = np.minimum(stream.m.astype(np.int64) * s, lds)
    yb = stream.n.astype(np.int64)
    bx = min(max(_pow2_ceil(int(np.median(xb))), 64), _pow2_ceil(lds))
    by = min(max(_pow2_ceil(int(np.median(yb))), 4), 4096)
    out = []
    for fx, fy in ((1, 1), (2, 1), (1, 2), (2, 2), (0.5, 1), (1, 0.5), (4, 4)):
        c = (int(min(max(bx * fx, 64), _pow2_ceil(lds))), int(min(max(by * fy, 4), 4096)))
        if c not in out:
            out.append(c)
    return out[:max_cands]


def capacity_class(stream, bx: int, by: int, tasks_per_batch: int) -> int:
    """Smallest LDS class whose per-bin record capacity holds the densest bin
    of a launch (with 2x headroom), from the rect geometry on the host."""
    r, nr = stream.rects()
    batches = R.batches_by_tasks(stream.task, tasks_per_batch)
    worst = 0
    for a, b in batches[: min(len(batches), 8)]:
        rows = []
        for j in range(2):
            sel = np.nonzero(nr[a:b] > j)[0] + a
            rows.append(r[sel, 4 * j:4 * j + 4].astype(np.int64))
        q = np.concatenate(rows)
        if not len(q):
            continue
        gx0, gx1 = q[:, 0] // bx, (q[:, 1] - 1) // bx
        gy0, gy1 = q[:, 2] // by, (q[:, 3] - 1) // by
        w = gx1 - gx0 + 1
        cnt = w * (gy1 - gy0 + 1)
        idx = np.repeat(np.arange(len(q)), cnt)
        off = np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        gx = gx0[idx] + off % w[idx]
        gy = gy0[idx] - gy0.min() + off // w[idx]
        key = gy * (int(gx.max()) + 1) + gx
        worst = max(worst, int(np.bincount(key).max()))
    for k, cap in enumerate(KCLASS_RECORDS):
        if 2 * worst <= cap:
            return k
    return 2


def measure(stream, bx: int, by: int, kernel: int, tasks_per_batch: int, reps: int = 2) -> float:
    """Seconds per access to resolve the sample, plain form, on the device."""
    acc = R.to_access_array(stream.task, stream.mode, stream.addr, stream.m, stream.n)
    batches = R.batches_by_tasks(stream.task, tasks_per_batch)
    best = float("inf")
    with R.DependencyDomain(int(stream.ld[0]), int(stream.s[0]), bin_x_bytes=bx, bin_y_cols=by, kernel=kernel,
                            max_rects=2 * max(b - a for a, b in batches) + 16) as dom:
        for _ in range(reps):
            dom.reset()
            dom.sync()
            t0 = time.perf_counter()
            for a, b in batches:
                dom.submit(acc[a:b])
            dom.sync()
            best = min(best, time.perf_counter() - t0)
            dom.consume()
    return best / max(len(stream), 1)


# A capacity class above the one the sample needs holds more HBM per bin, and the sample
# (a prefix) can understate what its launches cost later: the heat stream's sample once
# timed class 1 2% faster than class 0, and the whole stream then resolved 30% slower.
# Such a class is taken only when it is faster by this much per class.
CLASS_MARGIN = 0.05


def tune(sample, tasks_per_batch: int = 4096, max_cands: int = 7) -> TuneResult:
    """Pick (BX, BY, class) for streams like `sample` (a prefix of the stream)."""
    tried = []
    best = None
    for bx, by in candidates(sample, max_cands):
        k = capacity_class(sample, bx, by, tasks_per_batch)
        for kc in sorted({k, min(k + 1, 2)}):
            try:
                t = measure(sample, bx, by, kc, tasks_per_batch)
            except R.ResolverError as e:
                tried.append((bx, by, kc, None, str(e)))
                continue
            tried.append((bx, by, kc, t * 1e6, ""))
            score = t * (1.0 + CLASS_MARGIN * (kc - k))
            if best is None or score < best[4]:
                best = (bx, by, kc, t, score)
    if best is None:
        raise RuntimeError(f"no bin configuration resolves the sample: {tried}")
    return TuneResult(best[0], best[1], best[2], best[3] * 1e6, tried)


def tune_merge(pool_factory, run, merges=(1, 4, 16, 64)) -> int:
    """Batches per launch for the batched form: run(pool, merge) -> seconds."""
    best = None
    for m in merges:
        pool = pool_factory()
        try:
            t = run(pool, m)
        except R.ResolverError:
            continue
        finally:
            pool.close()
        if best is None or t < best[1]:
            best = (m, t)
    return best[0] if best else merges[0]
