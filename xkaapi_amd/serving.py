"""Serving loop: many clients submit task streams; the GPU resolver serves them.

Each client owns a dependency domain admitted into one device Pool against a
device-memory budget (a client waits for admission while the budget is
full).  Client batches arrive by an open-loop schedule; the server batches
continuously: whenever fewer than `depth` launches are in flight, every batch
that has arrived is staged into one launch (the batched form), or each batch
gets a launch of its own (the plain per-batch form).  Launches queue on the
pool's stream, so staging the next launch overlaps the device work of the
current ones.  Submit-to-ready latency of a batch = its arrival -> the
completion of the launch that resolved it, as the host observes it; every
task of the batch counts once in the percentiles.

Edges are copied back whenever nothing is in flight (and at the end) and
filed per client by the domain index each edge carries; a client leaves --
its domain's bins return to the pool -- at the first such drain after its
last batch completed.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np

from . import resolver as R
from . import streams as S


@dataclass
class Client:
    cid: int
    kind: str            # "cholesky" or "lu"
    ntiles: int
    stream: object
    acc: np.ndarray
    batches: list        # (start, end) access ranges of whole tasks
    ntasks: list         # tasks per batch
    dom: object = None
    next_batch: int = 0  # next batch to arrive
    completed: int = 0
    left: bool = False
    edges: list = field(default_factory=list)


def make_clients(nclients: int = 64, nmin: int = 4, nmax: int = 48, tile: int = 128, batch_tasks: int = 256,
                 seed: int = 0x5C):
    """Cholesky or LU (50/50) clients with n x n tiles of tile^2 fp64, n ~ U[nmin, nmax]."""
    kinds, ntiles = S.clients(nclients, nmin, nmax, seed)
    out = []
    for c in range(nclients):
        n = int(ntiles[c])
        kind = "cholesky" if kinds[c] == 0 else "lu"
        st = (S.cholesky if kind == "cholesky" else S.lu)(n * tile, tile)
        acc = R.to_access_array(st.task, st.mode, st.addr, st.m, st.n)
        b = R.batches_by_tasks(st.task, batch_tasks)
        nt = [int(np.count_nonzero(np.diff(st.task[a:e].astype(np.int64))) + 1) for a, e in b]
        out.append(Client(c, kind, n, st, acc, b, nt))
    return out


def poisson_schedule(clients, rate: float, seed: int = 7):
    """Open-loop arrivals: a Poisson process of `rate` batches/s; each arrival is
    the next batch of a client drawn uniformly among those with batches left.
    rate = inf: everything arrives at t = 0 (closed loop, saturation)."""
    rng = np.random.default_rng(seed)
    left = {c.cid: len(c.batches) for c in clients}
    active = [c.cid for c in clients if left[c.cid]]
    t = 0.0
    ev = []
    while active:
        if np.isfinite(rate):
            t += rng.exponential(1.0 / rate)
        k = int(rng.integers(len(active)))
        cid = active[k]
        ev.append((t, cid))
        left[cid] -= 1
        if left[cid] == 0:
            active[k] = active[-1]
            active.pop()
    return ev


@dataclass
class RunResult:
    mode: str
    accesses: int
    makespan_s: float
    launches: int
    lat_us: np.ndarray
    weights: np.ndarray
    edges: dict

    @property
    def rate(self):
        return self.accesses / self.makespan_s

    def pct(self, q):
        o = np.argsort(self.lat_us)
        v, w = self.lat_us[o], self.weights[o].astype(np.float64)
        c = np.cumsum(w) / w.sum()
        return float(v[min(int(np.searchsorted(c, q / 100.0)), len(v) - 1)])


def col_range(st):
    ld, s = int(st.ld[0]), int(st.s[0])
    cols = (st.addr // (ld * s)).astype(np.int64)
    return int(cols.min()), int((cols + st.n.astype(np.int64)).max()) + 1


def serve(clients, schedule, mode: str = "batched", budget_bytes: int = 16 << 30, kernel: int = 1,
          graph: bool = True, depth: int = 2, max_rects: int = 1 << 20, bin_x_bytes: int = 1024,
          bin_y_cols: int = 128) -> RunResult:
    """Run one schedule through a fresh pool; mode "batched" or "plain"."""
    for c in clients:
        c.next_batch = c.completed = 0
        c.left = False
        c.dom = None
        c.edges = []
    byid = {c.cid: c for c in clients}
    pool = R.Pool(kernel=kernel, graph=graph and mode == "batched", max_rects=max_rects,
                  budget_bytes=budget_bytes)
    owner = {}              # domain index -> client, for filing edges
    pending = collections.deque()   # (arrival, cid, batch)
    waiting = collections.deque()   # arrived batches of clients not admitted yet
    inflight = collections.deque()  # (ticket, [(arrival, cid, batch)])
    lat, wts = [], []
    nacc = sum(len(c.acc) for c in clients)
    launches = 0

    def drain():
        e, d = pool.edges()
        for idx in np.unique(d):
            owner[int(idx)].edges.append(e[d == idx])
        for c in clients:
            if c.completed == len(c.batches) and c.dom is not None and not c.left:
                owner.pop(c.dom.index, None)
                pool.leave(c.dom)
                c.left = True

    def admit(cid):
        c = byid[cid]
        if c.dom is not None:
            return True
        try:
            c.dom = pool.domain(int(c.stream.ld[0]), int(c.stream.s[0]), col_range(c.stream), bin_x_bytes,
                                bin_y_cols)
        except R.ResolverError as ex:
            if ex.rc == -12:  # -ENOMEM: not admitted now
                return False
            raise
        owner[c.dom.index] = c
        return True

    t0 = time.perf_counter()
    i = 0
    try:
        while i < len(schedule) or pending or waiting or inflight:
            now = time.perf_counter() - t0
            while i < len(schedule) and schedule[i][0] <= now:
                t, cid = schedule[i]
                c = byid[cid]
                pending.append((t, cid, c.next_batch))
                c.next_batch += 1
                i += 1
            if inflight:
                done = pool.poll()
                tnow = time.perf_counter() - t0
                while inflight and inflight[0][0] <= done:
                    for (ta, cid, b) in inflight.popleft()[1]:
                        lat.append((tnow - ta) * 1e6)
                        wts.append(byid[cid].ntasks[b])
                        byid[cid].completed += 1
            if not inflight:
                drain()
                # clients that could not be admitted retry once others have left
                while waiting:
                    pending.appendleft(waiting.pop())
            if pending and len(inflight) < depth:
                batch = []
                hold = collections.deque()
                while pending:
                    ta, cid, b = pending.popleft()
                    if not admit(cid):
                        hold.append((ta, cid, b))
                        continue
                    c = byid[cid]
                    a, e = c.batches[b]
                    try:
                        pool.stage(c.dom, c.acc[a:e])
                    except R.ResolverError as ex:
                        if ex.rc != -28:  # -ENOSPC: the launch is full; the rest waits for the next one
                            raise
                        pending.appendleft((ta, cid, b))
                        break
                    batch.append((ta, cid, b))
                    if mode == "plain":
                        break
                waiting.extend(hold)
                if not batch and not inflight and hold and pool.stats()["domains"] == 0:
                    raise RuntimeError("the device budget cannot hold even one waiting client")
                if batch:
                    inflight.append((pool.launch(), batch))
                    launches += 1
        drain()
        makespan = time.perf_counter() - t0
    finally:
        pool.close()
    edges = {c.cid: (np.concatenate(c.edges) if c.edges else np.zeros((0, 2), np.uint32)) for c in clients}
    return RunResult(mode, nacc, makespan, launches, np.asarray(lat), np.asarray(wts), edges)


def saturation(clients, mode: str = "batched", **kw) -> RunResult:
    """Closed loop: every batch available at t = 0."""
    return serve(clients, poisson_schedule(clients, float("inf")), mode, **kw)
