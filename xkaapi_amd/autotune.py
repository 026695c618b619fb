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
  3. each candidate resolves the sample on the device in launches of the size
     the stream will use (the plain form's batches, or the batched form's
     merged launches), timed by HIP events on the device (host record
     building is not counted: with large launches it would dominate), the
     median of a few passes; failures (a capacity overflow) disqualify it;
     the candidates come back ranked.

The sample is a set of windows spread across the stream (spread_windows), not
its prefix: a prefix can misstate what later launches cost.  The caller can
confirm the finalists on the whole stream (the benchmark does).

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
    ranked: list = None   # (BX, BY, class, us per access) of every configuration that resolved, best first


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


def spread_windows(stream, nwin: int = 4, win_tasks: int = 50_000):
    """nwin runs of win_tasks whole tasks at evenly spaced points of the stream
    (its start, its end and between), concatenated with tasks renumbered: a
    sample whose later windows meet the state the earlier ones left, as the
    stream's later launches do."""
    st = stream.task_starts()
    ntask = len(st) - 1
    if nwin * win_tasks >= ntask:
        return stream
    parts = []
    for k in range(nwin):
        t0 = (ntask - win_tasks) * k // max(nwin - 1, 1)
        parts.append(stream.take(slice(int(st[t0]), int(st[t0 + win_tasks]))))
    from . import streams as S
    return S.concat(parts)


def measure(stream, bx: int, by: int, kernel: int, tasks_per_batch: int, reps: int = 3,
            launch_tasks: int | None = None) -> float:
    """Device seconds per access to resolve the sample in launches of
    launch_tasks tasks (default: tasks_per_batch, the plain form), the median
    of reps passes; each launch is timed from its first kernel's start to its
    last kernel's end."""
    acc = R.to_access_array(stream.task, stream.mode, stream.addr, stream.m, stream.n)
    batches = R.batches_by_tasks(stream.task, launch_tasks or tasks_per_batch)
    times = []
    with R.DependencyDomain(int(stream.ld[0]), int(stream.s[0]), bin_x_bytes=bx, bin_y_cols=by, kernel=kernel,
                            max_rects=3 * max(b - a for a, b in batches) + 16) as dom:
        for _ in range(reps):
            dom.reset()
            dom.sync()
            dom.launch_times(clear=True)
            dom.set_timing(True)
            for a, b in batches:
                dom.submit(acc[a:b])
            dom.sync()
            dom.set_timing(False)
            times.append(float(dom.launch_times(clear=True).sum()) * 1e-3)
            dom.consume()
    return float(np.median(times)) / max(len(stream), 1)


# A capacity class above the one the sample needs holds more HBM per bin, and the sample
# (a prefix) can understate what its launches cost later: the heat stream's sample once
# timed class 1 2% faster than class 0, and the whole stream then resolved 30% slower.
# Such a class is taken only when it is faster by this much per class.
CLASS_MARGIN = 0.05


def tune(sample, tasks_per_batch: int = 4096, max_cands: int = 7, launch_tasks: int | None = None,
         reps: int = 3) -> TuneResult:
    """Pick (BX, BY, class) for streams like `sample` (windows of the stream,
    see spread_windows), timed in launches of launch_tasks tasks (the batched
    form) or of tasks_per_batch (the plain form).  `ranked` lists every
    configuration that resolved the sample, best score first."""
    tried = []
    scored = []
    lt = launch_tasks or tasks_per_batch
    for bx, by in candidates(sample, max_cands):
        k = capacity_class(sample, bx, by, lt)
        for kc in sorted({k, min(k + 1, 2)}):
            try:
                t = measure(sample, bx, by, kc, tasks_per_batch, reps=reps, launch_tasks=launch_tasks)
            except R.ResolverError as e:
                tried.append((bx, by, kc, None, str(e)))
                continue
            tried.append((bx, by, kc, t * 1e6, ""))
            scored.append((t * (1.0 + CLASS_MARGIN * (kc - k)), bx, by, kc, t))
    if not scored:
        raise RuntimeError(f"no bin configuration resolves the sample: {tried}")
    scored.sort()
    _, bx, by, kc, t = scored[0]
    res = TuneResult(bx, by, kc, t * 1e6, tried)
    res.ranked = [(b, c, k, tt * 1e6) for _, b, c, k, tt in scored]
    return res


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
