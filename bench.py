#!/usr/bin/env python
"""Resolver benchmark on one MI355X.

Workload: the plain-GPU configuration -- (a) tiled Cholesky on an 8192^2 fp64
matrix in 128^2 tiles (45,760 tasks, 133,120 accesses) and (b) 1,000,000
seeded random single-access tasks over a 65536^2 fp32 index space -- both in
submission batches of 4096 tasks.  Synthetic streams, generated in-repo.

One step = one pass of the resolver over both streams from an empty domain,
with the rect records already resident in HBM.  Two forms are timed:
  plain   : one launch (count/alloc/fill/resolve/dedup/final) per 4096-task batch
  batched : consecutive batches merged into one launch (up to --merge batches)
`value` is the plain form's accesses/sec (whole job); the batched form, the
per-batch latency percentiles, the roofline of the dominant kernel and the
reference CPU resolver's rate are reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def weighted_pct(values, weights, q):
    o = np.argsort(values)
    v = np.asarray(values)[o]
    w = np.asarray(weights, np.float64)[o]
    c = np.cumsum(w) / w.sum()
    return float(v[min(np.searchsorted(c, q / 100.0), len(v) - 1)])


class Prepared:
    """Device-resident rect records of one stream, per plain batch and per merged
    group, for each candidate number of batches per launch (merges)."""

    def __init__(self, name, st, dom_kw, merges, torch):
        from xkaapi_amd import resolver as R
        self.name, self.st = name, st
        ld, s = int(st.ld[0]), int(st.s[0])
        self.acc = R.to_access_array(st.task, st.mode, st.addr, st.m, st.n)
        self.batches = R.batches_by_tasks(st.task, 4096)
        groups = {m: [(self.batches[i][0], self.batches[min(i + m, len(self.batches)) - 1][1])
                      for i in range(0, len(self.batches), m)] for m in merges}
        # one domain per form, each sized for its own launches
        stream = torch.cuda.current_stream().cuda_stream
        self.doms = {
            "plain": R.DependencyDomain(ld, s, max_rects=2 * max(b - a for a, b in self.batches),
                                        max_edges=8 * len(st) + (1 << 20), stream=stream, **dom_kw),
            "batched": R.DependencyDomain(ld, s, max_rects=2 * max(b - a for g in groups.values() for a, b in g),
                                          max_edges=8 * len(st) + (1 << 20), stream=stream, **dom_kw)}
        self.plain = [self._upload(self.doms["plain"], a, b, torch) for a, b in self.batches]
        self.merged_by = {m: [self._upload(self.doms["batched"], a, b, torch) for a, b in g]
                          for m, g in groups.items()}
        self.merge = merges[0]
        self.tried = {}
        self.ntasks_plain = [int(np.count_nonzero(np.diff(st.task[a:b].astype(np.int64))) + 1) for a, b in self.batches]

    @property
    def merged(self):
        return self.merged_by[self.merge]

    def tune_merge(self, torch):
        """Batches per launch of the batched form: the fastest candidate on one pass."""
        best = None
        for m in self.merged_by:
            self.merge = m
            self.run("batched")
            torch.cuda.synchronize()
            t = time.perf_counter()
            self.run("batched")
            self.doms["batched"].sync()
            dt = time.perf_counter() - t
            self.tried[m] = round(dt * 1e3, 3)
            if best is None or dt < best[1]:
                best = (m, dt)
        self.merge = best[0]
        return self.merge

    def _upload(self, dom, a, b, torch):
        recs, ylo, yhi = dom.build_records(self.acc[a:b], a)
        dev = torch.from_numpy(recs.view(np.int32).reshape(-1)).cuda()
        return dev, len(recs), b - a, ylo, yhi

    def run(self, which):
        dom = self.doms[which]
        dom.reset()
        for dev, nrec, nacc, ylo, yhi in (self.plain if which == "plain" else self.merged):
            dom.submit_device(dev.data_ptr(), nrec, nacc, ylo, yhi)


def host_info():
    """CPU model and the cores this process may use (the GPU box's host)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu": model, "nproc": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}


def load_digests():
    with open(os.path.join(ROOT, "tests", "golden", "digests.json")) as fh:
        return json.load(fh)


def cpu_baseline(chol, rnd_prefix=10_000, chol_target_s=2.0):
    """The reference's own KHP engine (oracle/_ref, built from the reference
    sources) on a bounded sample of the headline workload, one core: the whole
    Cholesky stream (repeated to about chol_target_s) and the first rnd_prefix
    accesses of the random stream (its canonical fragmentation grows too fast
    for more: ~2,500 fragments per access at 10^5 accesses).  value = sample
    accesses / sample seconds, the streams weighted as one Cholesky pass plus
    the random prefix."""
    from oracle import refkhp
    from xkaapi_amd import streams as S
    if not refkhp.available():
        return None
    r, nr = chol.rects()
    tot_ns, reps = 0, 0
    while tot_ns < chol_target_s * 1e9 and reps < 1000:
        ns, _ = refkhp.time_resolve(chol.task, chol.mode, r, nr)
        tot_ns += ns
        reps += 1
    chol_rate = reps * len(chol) / (tot_ns / 1e9)
    rs = S.random2d(rnd_prefix)
    r, nr = rs.rects()
    rns, _ = refkhp.time_resolve(rs.task, rs.mode, r, nr)
    rnd_rate = len(rs) / (rns / 1e9)
    t_sample = len(chol) / chol_rate + rns / 1e9
    return {"value": round((len(chol) + len(rs)) / t_sample, 1), "unit": "accesses/s", "cores": 1,
            "kind": "reference",
            "sample": f"Cholesky 8192^2 nb128 ({len(chol)} accesses, timed x{reps}) + the first {len(rs)} random "
                      f"accesses; LINK+PUT only, reference khp-tree.hpp engine, 1 thread",
            "per_stream": {"cholesky_8192_nb128": round(chol_rate, 1), f"random2d_prefix_{len(rs)}": round(rnd_rate, 1)},
            "host": host_info()}


def cpu_baseline_heat(steps=20):
    """Reference engine on the first `steps` steps of the heat halo stream, one core."""
    from oracle import refkhp
    from xkaapi_amd import streams as S
    if not refkhp.available():
        return None
    st = S.heat(16384, 128, 4, steps)
    r, nr = st.rects()
    ns, ne = refkhp.time_resolve(st.task, st.mode, r, nr)
    return {"value": round(len(st) / (ns / 1e9), 1), "unit": "accesses/s", "cores": 1, "kind": "reference",
            "sample": f"heat 16384^2, first {steps} of 200 steps ({len(st)} accesses, {ne} edges), 1 thread"}


def cpu_baseline_serving(clients):
    """Reference engine, one client domain per core (threads over min(64, cores)
    cores), every client's whole stream: accesses / wall seconds."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import refkhp
    if not refkhp.available():
        return None
    work = [(c.stream.task, c.stream.mode) + c.stream.rects() for c in clients]
    cores = min(64, len(os.sched_getaffinity(0)))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        list(ex.map(lambda w: refkhp.time_resolve(*w), work))
    dt = time.perf_counter() - t0
    n = sum(len(c.stream) for c in clients)
    return {"value": round(n / dt, 1), "unit": "accesses/s", "cores": cores, "kind": "reference",
            "sample": f"{len(clients)} client streams ({n} accesses), one domain per thread, closed loop"}


def roofline_leg(torch, prefix_name="random2d_102400"):
    """Roofline of the dominant kernel, k_resolve, on the random stream: the
    plain form's 4096-task launches over the stream's first 102,400 accesses
    (the oracle resolves no more in reasonable time), with the benchmark's bins
    (2048 B x 256 columns, class 1).  Algorithmic bytes per launch from the
    oracle's counts (tests/golden/digests.json): B_alg = (24 N_acc + 64 sum F_i
    + 12 E) / launches.  Duration: HIP events around each k_resolve tier on the
    launch stream."""
    from xkaapi_amd import resolver as R
    from xkaapi_amd import streams as S
    d = load_digests()[prefix_name]
    st = S.random2d(d["accesses"])
    acc = R.to_access_array(st.task, st.mode, st.addr, st.m, st.n)
    batches = R.batches_by_tasks(st.task, d["batch_tasks"])
    stream = torch.cuda.current_stream().cuda_stream
    dom = R.DependencyDomain(int(st.ld[0]), int(st.s[0]), bin_x_bytes=2048, bin_y_cols=256, kernel=1,
                             max_rects=2 * d["batch_tasks"], max_edges=8 * len(st) + (1 << 20), stream=stream)
    dev = []
    for a, b in batches:
        recs, ylo, yhi = dom.build_records(acc[a:b], a)
        dev.append((torch.from_numpy(recs.view(np.int32).reshape(-1)).cuda(), len(recs), b - a, ylo, yhi))
    kts = []
    for rep in range(3):
        dom.reset()
        dom.set_timing(rep > 0)
        for t, nrec, nacc, ylo, yhi in dev:
            dom.submit_device(t.data_ptr(), nrec, nacc, ylo, yhi)
        edges = dom.sync()
        assert edges == d["edges"], "the roofline pass must find the oracle's edge count"
        if rep > 0:
            kt, nl = dom.kernel_times(clear=True)
            kts.append(kt["k_resolve"] / nl)
        dom.set_timing(False)
    dom.close()
    nl = len(batches)
    b_alg = (24 * d["accesses"] + 64 * sum(d["sum_frags_per_batch"]) + 12 * d["edges"]) / nl
    t_res = min(kts) * 1e-3
    achieved = b_alg / t_res / 1e9
    traffic = prof_us = None
    tp = os.path.join(ROOT, "profiles", "resolver_random2d_prefix_plain_traffic.json")
    if os.path.exists(tp):  # HBM bytes per launch from the committed rocprofv3 counter runs
        t = json.load(open(tp))
        traffic = int(1024 * (t["fetch_kib_per_launch"] * t["fetch_factor"]
                              + t["write_kib_per_launch"] * t["write_factor"]))
        prof_us = t.get("rocprof_avg_us_per_launch")
    out = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
           "kernel": "k_resolve", "avg_us": round(t_res * 1e6, 2), "rocprof_avg_us": prof_us,
           "bytes_per_launch": int(b_alg),
           "units_per_launch": {"accesses": d["accesses"] / nl, "sum_F": sum(d["sum_frags_per_batch"]) / nl,
                                "edges": d["edges"] / nl},
           "workload": f"random2d, first {d['accesses']} accesses in {nl} plain 4096-task launches, "
                       "bins 2048 B x 256 cols, class 1; sum F_i and E from the oracle"}
    if traffic:
        # the hardware view: measured HBM bytes per launch over the same duration
        out["traffic_achieved"] = round(traffic / t_res / 1e9, 3)
        out["traffic_frac"] = round(traffic / t_res / 1e9 / HBM_PEAK_GBS, 6)
        out["note"] = ("B_alg charges 64 B per fragment of the reference's canonical partition (sum F_i); "
                       "the bin painter maps never hold those fragments, so achieved exceeds what HBM "
                       "moves: traffic_frac is the kernel's hardware roofline fraction")
    return out


def e2e_leg(chol, rnd, merge, reps=3):
    """End to end, as a runtime would drive it: per batch the host builds the
    rect records from the accesses and uploads them (xkrt_hip_dep_submit),
    the device resolves, and at the end every edge is copied into a pinned host
    buffer the caller holds.  Timed from the first submit to the last edge on
    the host; both forms."""
    import torch
    from xkaapi_amd import resolver as R
    out = {}
    ecap = 8 * max(len(chol) if chol is not None else 0, len(rnd) if rnd is not None else 0) + (1 << 20)
    pinned = torch.empty((ecap, 2), dtype=torch.int32, pin_memory=True).numpy().view(np.uint32)
    for form in ("plain", "batched"):
        total_s = 0.0
        nacc = 0
        for name, st, kw in (("cholesky_8192_nb128", chol, dict(bin_x_bytes=1024, bin_y_cols=128, kernel=0)),
                             ("random2d_65536", rnd, dict(bin_x_bytes=2048, bin_y_cols=256, kernel=1))):
            if st is None:
                continue
            acc = R.to_access_array(st.task, st.mode, st.addr, st.m, st.n)
            tpb = 4096 if form == "plain" else 4096 * merge[name]
            batches = R.batches_by_tasks(st.task, tpb)
            best = None
            with R.DependencyDomain(int(st.ld[0]), int(st.s[0]), max_rects=2 * max(b - a for a, b in batches),
                                    max_edges=8 * len(st) + (1 << 20), **kw) as dom:
                for _ in range(reps):
                    dom.reset()
                    dom.sync()
                    t0 = time.perf_counter()
                    for a, b in batches:
                        dom.submit(acc[a:b])
                    e = dom.edges(out=pinned)
                    dt = time.perf_counter() - t0
                    best = dt if best is None else min(best, dt)
            total_s += best
            nacc += len(st)
            out.setdefault(form, {})[name] = {"s": round(best, 4), "edges": int(len(e))}
        out[form]["accesses_per_s"] = round(nacc / total_s, 1)
        out[form]["ms_per_step"] = round(total_s * 1e3, 3)
    out["workload"] = ("accesses handed over in host memory; records built (on up to 16 host threads for batches of "
                       "64K accesses or more) and uploaded per launch, edges copied into a pinned host buffer at "
                       "the end; plain: 4096-task launches, batched: the tuned batches per launch")
    return out


def coherency_leg(n_stream=6000, reps=3):
    """Coherency fetch planning on the two-view stream, cut where the reference
    aborts (the first ninth allocation view): GPU planner per tree, one call
    each, against the oracle restatement on one host core."""
    from xkaapi_amd import coherency as K
    from xkaapi_amd import streams as S
    st = S.coherency(n_stream)
    views = [(8192, 8), (16384, 4)]
    subs = []
    for ld, s in views:
        sel = np.nonzero((st.ld == ld) & (st.s == s))[0]
        subs.append((ld, s, sel, st.take(sel)))
    # global cut: the first access at which either tree aborts
    cut = len(st)
    for ld, s, sel, sub in subs:
        with K.MemoryTree(ld, s, (0, 8192)) as t:
            _, _, fail = t.plan_counts(K.to_mcc_array(sub.dev, sub.mode, sub.addr, sub.m, sub.n))
        if fail >= 0:
            cut = min(cut, int(sel[fail]))
    trees = []
    for ld, s, sel, sub in subs:
        k = int(np.searchsorted(sel, cut))
        trees.append((ld, s, K.to_mcc_array(sub.dev[:k], sub.mode[:k], sub.addr[:k], sub.m[:k], sub.n[:k]),
                      sub.take(slice(0, k))))
    gpu_ms, wall_ms, parts, fetches = [], [], 0, 0
    for _ in range(reps):
        g = w = 0.0
        parts = fetches = 0
        for ld, s, acc, _ in trees:
            with K.MemoryTree(ld, s, (0, 8192)) as t:
                t0 = time.perf_counter()
                p, f, fail = t.plan_counts(acc)
                w += time.perf_counter() - t0
                g += t.stats()["last_ms"]
                parts += p
                fetches += f
                assert fail < 0
        gpu_ms.append(g)
        wall_ms.append(w * 1e3)
    out = {"accesses": int(cut), "trees": len(trees), "parts": int(parts), "fetches": int(fetches),
           "gpu_ms": round(min(gpu_ms), 3), "plan_call_ms": round(min(wall_ms), 3),
           "accesses_per_s_gpu": round(cut / (min(gpu_ms) * 1e-3), 1),
           "workload": f"two-view coherency stream (8192^2 buffer, LD 8192/8 B and 16384/4 B), first {cut} "
                       "accesses: up to the reference's ninth-view abort"}
    return out, [sub for _, _, _, sub in trees]


def coherency_cpu_baseline(subs):
    """The oracle's restatement of the reference memory tree on the same trees, one host core."""
    from oracle import khp
    t0 = time.perf_counter()
    n = 0
    for sub in subs:
        r, nr = sub.rects()
        khp.coherency(sub.dev, sub.mode, r, nr, sub.m * sub.n * sub.s)
        n += len(sub)
    dt = time.perf_counter() - t0
    return {"value": round(n / dt, 1), "unit": "accesses/s", "cores": 1, "kind": "port",
            "sample": f"the same {n} accesses, oracle/khp_oracle.c (restated memory-tree.hpp), 1 thread"}


def heat_leg(torch, steps=200, reps=3, finalists=2):
    """Heat-diffusion halo stream, 16384^2 fp32 in 128^2 tiles (16,384 tiles),
    `steps` timesteps; uneven batches (geometric, mean 2048 tasks, max 65536).
    Records resident in HBM; plain = one launch per batch, batched =
    consecutive batches merged into launches of up to 2^20 records.
    Bins and capacity class: the autotuner ranks the candidates on four
    windows spread across the stream, resolved in launches of the batched
    form's size, and the best `finalists` are then timed on the whole stream
    (batched form, median of reps passes); the fastest is used for both forms."""
    from xkaapi_amd import autotune as A
    from xkaapi_amd import resolver as R
    from xkaapi_amd import streams as S
    st = S.heat(16384, 128, 4, steps)
    sizes = S.geometric_batches(st.ntasks, 2048.0, 65536, 0x3B)
    batches = R.batches_by_tasks(st.task, sizes)
    acc = R.to_access_array(st.task, st.mode, st.addr, st.m, st.n)
    cap = 1 << 20
    groups, cur = [], None
    for a, b in batches:  # merged launches of up to cap records (2 per access at most)
        if cur is not None and 2 * (b - cur[0]) <= cap:
            cur = (cur[0], b)
        else:
            if cur is not None:
                groups.append(cur)
            cur = (a, b)
    groups.append(cur)
    t0 = time.time()
    launch_tasks = int(np.median([st.task[b - 1] - st.task[a] + 1 for a, b in groups]))
    tune = A.tune(A.spread_windows(st, nwin=4, win_tasks=launch_tasks), tasks_per_batch=2048, max_cands=4,
                  launch_tasks=launch_tasks)
    log(f"heat: {len(st)} acc / {st.ntasks} tasks / {len(batches)} batches; windows ranked {tune.ranked} "
        f"({time.time()-t0:.1f}s)")

    def run_form(spans, cfg, nreps):
        bx, by, kc = cfg
        maxrec = 2 * max(b - a for a, b in spans)
        dom = R.DependencyDomain(int(st.ld[0]), int(st.s[0]), bin_x_bytes=bx, bin_y_cols=by, kernel=kc,
                                 max_rects=maxrec, max_edges=4 * len(st) + (1 << 20),
                                 stream=torch.cuda.current_stream().cuda_stream)
        dev = []
        for a, b in spans:
            recs, ylo, yhi = dom.build_records(acc[a:b], a)
            dev.append((torch.from_numpy(recs.view(np.int32).reshape(-1)).cuda(), len(recs), b - a, ylo, yhi))
        torch.cuda.synchronize()
        ts, edges = [], 0
        for _ in range(nreps):
            dom.reset()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for d, nrec, nacc, ylo, yhi in dev:
                dom.submit_device(d.data_ptr(), nrec, nacc, ylo, yhi)
            edges = dom.sync()
            ts.append(time.perf_counter() - t1)
        dom.close()
        del dev
        torch.cuda.empty_cache()
        return float(np.median(ts)), min(ts), int(edges)

    # the finalists on the whole stream, batched form
    final = {}
    for bx, by, kc, _ in tune.ranked[:finalists]:
        med, _, _ = run_form(groups, (bx, by, kc), reps)
        final[(bx, by, kc)] = med
    pick = min(final, key=final.get)
    log(f"heat: finalists on the whole stream (median s) {final} -> {pick}")
    out = {}
    for form, spans in (("plain", batches), ("batched", groups)):
        med, best, edges = run_form(spans, pick, reps)
        out[form] = {"launches": len(spans), "s": round(best, 4), "median_s": round(med, 4),
                     "accesses_per_s": round(len(st) / best, 1), "edges": edges}
    out["speedup_batched_vs_plain"] = round(out["batched"]["accesses_per_s"] / out["plain"]["accesses_per_s"], 3)
    out["workload"] = (f"heat 16384^2 fp32, 128^2 tiles, {steps} steps: {len(st)} accesses, {st.ntasks} tasks, "
                       f"{len(batches)} geometric batches (mean 2048, max 65536)")
    out["autotune"] = {"bin_x_bytes": pick[0], "bin_y_cols": pick[1], "kernel": pick[2],
                       "windows_ranked": [list(r[:3]) + [round(r[3], 5)] for r in tune.ranked],
                       "finalists_whole_stream_s": {f"{k[0]}x{k[1]}c{k[2]}": round(v, 4) for k, v in final.items()},
                       "window_tasks": launch_tasks}
    return out


def serving_leg(fracs=(0.1, 0.3, 0.5, 0.7, 0.9), cpu=True):
    """64 clients (Cholesky / LU 50/50, n ~ U[4, 48] tiles of 128^2 fp64), one
    pool with a 16 GB budget, engine defaults (prearmed launches).  Saturation
    (closed loop, median of three runs) of the batched form
    (continuous batching, hipGraph launches) and of the plain form (one launch
    per batch); then open-loop Poisson arrivals at fractions of the batched
    saturation rate, both forms on the same schedule.  The serving engine is
    the library's native host thread; batches are submitted at their arrival
    times.  Latency = scheduled arrival -> the launch that resolved it complete
    and its edges in the client's host buffer, task-weighted; host staging and
    upload included.  Beside it, the reference engine with one client domain
    per core."""
    from xkaapi_amd import serving as V
    cs = V.make_clients(64, 4, 48, 128, 256)
    nb = sum(len(c.batches) for c in cs)
    sat = {m: V.saturation(cs, m, runs=3) for m in ("batched", "plain")}
    rate = nb / sat["batched"].makespan_s
    out = {"clients": len(cs), "accesses": sat["batched"].accesses, "batches": nb, "batch_tasks": 256,
           "saturation": {m: {"accesses_per_s": round(r.rate, 1), "launches": r.launches,
                              "makespan_ms": round(r.makespan_s * 1e3, 2)} for m, r in sat.items()},
           "sweep": []}
    for f in fracs:
        sched = V.poisson_schedule(cs, f * rate, seed=11)
        row = {"load": f, "arrival_batches_per_s": round(f * rate, 1)}
        for m in ("batched", "plain"):
            r = V.serve(cs, sched, m)
            row[m] = {"p50_us": round(r.pct(50), 1), "p99_us": round(r.pct(99), 1),
                      "accesses_per_s": round(r.rate, 1), "launches": r.launches, "gated": r.stats["gated"],
                      "driver_lag_p99_us": round(float(np.percentile(r.lag_us, 99)), 1)}
        out["sweep"].append(row)
        log(f"serving load {f}: {row}")
    # the engine without prearmed launches (each launch enqueued when its batches are staged),
    # on the same schedules
    out["no_prearm"] = {"sweep": []}
    for f in (0.5, 0.7):
        r = V.serve(cs, V.poisson_schedule(cs, f * rate, seed=11), "batched", prearm=False)
        out["no_prearm"]["sweep"].append({"load": f, "accesses_per_s": round(r.rate, 1),
                                          "p50_us": round(r.pct(50), 1), "p99_us": round(r.pct(99), 1),
                                          "launches": r.launches})
    log(f"serving without prearm: {out['no_prearm']}")
    if cpu:
        out["cpu_baseline"] = cpu_baseline_serving(cs)
    return out


PCIE_PEAK_GBS = 64.0  # PCIe Gen5 x16, one direction (MI355X host link)


def segments_leg(reps=3, nchunks=4096, steps=50):
    """Segment domain (the reference's interval tree): a 1-D blocked stencil
    over 4096 chunks of 64 KiB (256 MiB), 50 sweeps; each task reads its
    neighbours' 256-byte halos and updates its chunk (3 segment accesses).
    One sweep per launch, submitted from host memory."""
    from xkaapi_amd import resolver as R
    from xkaapi_amd import streams as S
    st = S.stencil1d(nchunks, 1 << 16, steps, 256)
    n = len(st["task"])
    acc = R.to_access_array(st["task"], st["mode"], st["addr"], st["size"], 0,
                            kind=np.full(n, R.KIND_SEGMENT, np.uint8))
    batches = R.batches_by_tasks(st["task"], nchunks)
    best, dev_ms, edges = None, None, 0
    for _ in range(reps):
        with R.SegmentDomain(1 << 16, 1 << 16, 4, kernel=1, max_rects=1 << 16) as d:
            d.set_timing(True)
            t0 = time.perf_counter()
            for a, b in batches:
                d.submit(acc[a:b])
            edges = d.sync()
            dt = time.perf_counter() - t0
            lt = d.launch_times()
        if best is None or dt < best:
            best, dev_ms = dt, float(np.sum(lt))
    out = {"accesses": n, "tasks": int(st["task"][-1]) + 1, "edges": edges, "launches": len(batches),
           "accesses_per_s": round(n / best, 1), "accesses_per_s_device": round(n / (dev_ms * 1e-3), 1),
           "ms": round(best * 1e3, 3), "device_ms": round(dev_ms, 3),
           "workload": f"1-D stencil over segments: {nchunks} chunks of 64 KiB, {steps} sweeps, one sweep per "
                       "launch (segment domain, 64 KiB columns, 64 KiB x 4-column bins)"}
    # the reference's KHPTree<1> engine on the first sweeps, one core
    k = max(1, min(steps, 5))
    m = int(np.searchsorted(st["task"], k * nchunks))
    seg = np.stack([st["addr"][:m], st["addr"][:m] + st["size"][:m]], axis=1)
    try:
        from oracle import refkhp
        kind, fn = ("reference", refkhp.resolve_intervals) if refkhp.available() else (None, None)
    except Exception:
        kind, fn = None, None
    if fn is None:
        from oracle import khp
        kind, fn = "port", khp.resolve_intervals
    t0 = time.perf_counter()
    fn(st["task"][:m], st["mode"][:m], seg)
    dt = time.perf_counter() - t0
    out["cpu_baseline"] = {"value": round(m / dt, 1), "unit": "accesses/s", "cores": 1, "kind": kind,
                           "sample": f"first {k} sweeps ({m} accesses) through "
                                     + ("the reference's KHPTree<1> with the interval-tree hooks (oracle/_ref)"
                                        if kind == "reference" else "the oracle's restatement")}
    return out


def release_leg():
    """Successor release on the device: the Cholesky DAG 8192^2 fp64 in
    128^2 tiles (45,760 tasks), empty tasks.  The host submits every task in
    4096-task batches into a task domain with release on, then loops: read the
    released tasks, complete them, until all ran.  Two loops: completions only
    enqueued, the loop taking whatever has landed in the queue; and
    completions that wait until their successors are released (as the
    reference's __task_complete, src/task.cc:160-247), so each pass takes
    every task released by the last."""
    from xkaapi_amd import resolver as R
    from xkaapi_amd import streams as S
    from oracle import khp
    st = S.cholesky(8192, 128)
    ntask = int(st.task[-1]) + 1
    acc = R.to_task_access_array(st.task, st.mode, np.zeros(len(st), np.uint8), st.addr, st.m, st.n, st.ld, st.s)
    batches = R.batches_by_tasks(st.task, 4096)
    def run(wait):
        best = None
        for _ in range(3):
            with R.TaskDomain(kernel=1, release=True, max_tasks=1 << 16, max_accesses=1 << 18,
                              max_succ=1 << 20) as t:
                t0 = time.perf_counter()
                for a, b in batches:
                    t.submit(acc[a:b])
                t.sync()
                t1 = time.perf_counter()
                done, waves, deadline = 0, 0, t1 + 60.0
                while done < ntask:
                    r = t.ready()
                    if len(r):
                        t.complete(r, wait=wait)
                        done += len(r)
                        waves += 1
                    elif time.perf_counter() > deadline:
                        raise RuntimeError(f"release: {done} of {ntask} tasks ran")
                t.sync()
                t2 = time.perf_counter()
                stt = t.stats()
            if stt["completed"] != ntask:
                raise RuntimeError(f"release: {stt['completed']} of {ntask} tasks completed")
            if best is None or t2 - t0 < best["total_s"]:
                best = {"total_s": t2 - t0, "resolve_s": t1 - t0, "run_s": t2 - t1, "waves": waves, "stats": stt}
        return best

    best, waited = run(False), run(True)
    out = {"tasks": ntask, "accesses": len(st), "edges_applied": best["stats"]["succ_nodes"],
           "tasks_per_s": round(ntask / best["total_s"], 1), "ms": round(best["total_s"] * 1e3, 3),
           "resolve_and_apply_ms": round(best["resolve_s"] * 1e3, 3), "run_ms": round(best["run_s"] * 1e3, 3),
           "waves": best["waves"],
           "waiting": {"tasks_per_s": round(ntask / waited["total_s"], 1), "ms": round(waited["total_s"] * 1e3, 3),
                       "run_ms": round(waited["run_s"] * 1e3, 3), "waves": waited["waves"]},
           "workload": "cholesky 8192^2 nb128 DAG, empty tasks: submit all (4096-task batches, edges applied "
                       "on the device), then read released tasks and complete them until all ran (waiting: "
                       "each completion returns once its successors are released)"}
    # the reference's algorithm on one core: apply every edge, commit, run FIFO
    r, nr = st.rects()
    e = np.asarray(khp.resolve(st.task, st.mode, r, nr), np.uint32)
    t0 = time.perf_counter()
    ran, _ = khp.release_run(st.task, st.mode, e, ntask)
    dt = time.perf_counter() - t0
    assert ran == ntask
    out["cpu_baseline"] = {"value": round(ntask / dt, 1), "unit": "tasks/s", "cores": 1, "kind": "port",
                           "sample": f"the same DAG ({len(e)} edges given), oracle/khp_oracle.c xko_release_run: "
                                     "edges applied, tasks committed and run FIFO as the reference's "
                                     "__task_precedes / __task_commit / __task_complete, 1 thread"}
    return out


def handles_leg(n=1 << 20, nh=4096):
    """Handle dependency domain: 2^20 accesses on 4096 handles, tasks of 1..3
    accesses, 30% of writes concurrent (inoutset); 16384-task batches."""
    from xkaapi_amd import resolver as R
    from xkaapi_amd import streams as S
    from oracle import handles as H
    st = S.handle_stream(n, nh)
    acc = R.to_handle_access_array(st["task"], st["handle"], st["mode"], st["concurrency"])
    batches = R.batches_by_tasks(st["task"], 16384)
    best = None
    for _ in range(3):
        with R.HandleDomain(max_handles=nh * 4, max_accesses=n + 16, max_virtual=n, max_batch=1 << 16,
                            max_edges=1 << 23) as d:
            t0 = time.perf_counter()
            for a, b in batches:
                d.submit(acc[a:b])
            ne, nv = d.sync()
            dt = time.perf_counter() - t0
        if best is None or dt < best[0]:
            best = (dt, ne, nv)
    out = {"accesses": n, "handles": nh, "edges": best[1], "empty_writes": best[2],
           "accesses_per_s": round(n / best[0], 1), "ms": round(best[0] * 1e3, 3), "launches": len(batches),
           "workload": f"{n} handle accesses on {nh} handles, 1..3 per task, 30% of writes concurrent, "
                       "16384-task batches from host memory"}
    m = 100_000
    t0 = time.perf_counter()
    H.resolve(st["task"][:m], st["handle"][:m], st["mode"][:m], st["concurrency"][:m])
    dt = time.perf_counter() - t0
    out["cpu_baseline"] = {"value": round(m / dt, 1), "unit": "accesses/s", "cores": 1, "kind": "port",
                           "sample": f"first {m} accesses through oracle/handles.py (the dependency map restated "
                                     "in Python), 1 thread"}
    return out


def fetch_exec_leg(torch, tile=1024):
    """Fetch-plan execution: GPU reads of every 1024^2 tile of an 8192^2 fp64
    matrix in pinned host memory (512 MiB), planned in one call on an
    asynchronous coherency tree; the H2D fetch rows executed by one
    xkrt_hip_mcc_exec kernel into the chunk arena.  Beside it: one DMA copy
    of the same bytes (what per-fetch hipMemcpy2DAsync calls are bound by)."""
    from xkaapi_amd import coherency as K
    N, s = 8192, 8
    host = torch.empty(N * N * s, dtype=torch.uint8, pin_memory=True)
    arena = torch.empty(N * N * s + (64 << 20), dtype=torch.uint8, device="cuda")
    nt = N // tile
    ii, jj = np.meshgrid(np.arange(nt), np.arange(nt), indexing="ij")
    addr = ((jj.ravel() * tile * N + ii.ravel() * tile) * s).astype(np.uint64)
    k = len(addr)
    acc = K.to_mcc_array(np.ones(k, np.uint8), np.ones(k, np.uint8), addr, np.full(k, tile, np.uint64),
                         np.full(k, tile, np.uint64))
    stream = torch.cuda.current_stream()
    best = None
    for _ in range(3):
        with K.MemoryTree(N, s, (0, N), async_completion=True) as t:
            t.exec_bind(host.data_ptr(), arena.data_ptr(), arena.numel())
            plan = t.plan(acc)
            base = t.stats()["last_fetch_base"]
            ids = np.arange(base, base + len(plan.fetches), dtype=np.uint64)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            t.exec(ids, stream.cuda_stream)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1)
            t.complete(ids)
        best = ms if best is None else min(best, ms)
    nbytes = N * N * s
    dma = None
    dev = arena[:nbytes]
    for _ in range(3):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        dev.copy_(host, non_blocking=True)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        dma = ms if dma is None else min(dma, ms)
    gbs = nbytes / (best * 1e-3) / 1e9
    out = {"fetches": int(len(plan.fetches)), "bytes": nbytes, "ms": round(best, 3), "GB_per_s": round(gbs, 2),
           "roofline": {"bound": "pcie", "achieved": round(gbs, 2), "peak": PCIE_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / PCIE_PEAK_GBS, 4)},
           "dma_copy_same_bytes": {"ms": round(dma, 3), "GB_per_s": round(nbytes / (dma * 1e-3) / 1e9, 2)},
           "workload": f"{k} GPU reads of {tile}^2 fp64 tiles of an {N}^2 matrix in pinned host memory, one plan "
                       "call, every H2D fetch row executed by one kernel"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--random-n", type=int, default=1_000_000)
    ap.add_argument("--merge", type=int, default=0, help="batches per launch in the batched form (0: tuned)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("all", "cholesky", "random"), default="all",
                    help="restrict to one stream (profiling runs)")
    ap.add_argument("--forms", choices=("both", "plain", "batched"), default="both")
    ap.add_argument("--no-coherency", action="store_true")
    ap.add_argument("--no-heat", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-e2e", action="store_true")
    ap.add_argument("--heat-steps", type=int, default=200)
    ap.add_argument("--no-rows", action="store_true", help="skip the segment / release / handle / fetch legs")
    ap.add_argument("--calibrate-mib", type=int, default=0,
                    help="also run one streaming copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported (run --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    from xkaapi_amd import streams as S

    t0 = time.time()
    chol = S.cholesky(8192, 128)
    rnd = S.random2d(args.random_n) if args.workload != "cholesky" else None
    log(f"streams: cholesky {len(chol)} acc / {chol.ntasks} tasks ({time.time()-t0:.1f}s)")
    preps = []
    merges = [args.merge] if args.merge else [16, 32, 64, 128]  # batches per launch: given, or tuned
    if args.workload in ("all", "cholesky"):
        preps.append(Prepared("cholesky_8192_nb128", chol, dict(bin_x_bytes=1024, bin_y_cols=128, kernel=0),
                              merges, torch))
    if rnd is not None:
        preps.append(Prepared("random2d_65536", rnd, dict(bin_x_bytes=2048, bin_y_cols=256, kernel=1), merges,
                              torch))
    torch.cuda.synchronize()
    log(f"records resident in HBM ({time.time()-t0:.1f}s)")

    forms = ("plain", "batched") if args.forms == "both" else (args.forms,)
    # correctness gate on the Cholesky stream: the closed-form edge count
    if preps[0].name.startswith("cholesky"):
        p0 = preps[0]
        for form in forms:
            p0.run(form)
            assert p0.doms[form].sync() == 131040, "Cholesky 8192/128 must have 131,040 edges"

    for _ in range(args.warmup):
        for p in preps:
            for form in forms:
                p.run(form)
    torch.cuda.synchronize()
    if "batched" in forms and len(merges) > 1:  # the autotuner's pick of batches per launch, per stream
        for p in preps:
            log(f"{p.name}: batches per launch {p.tune_merge(torch)} (ms per pass: {p.tried})")
    for p in preps:
        for d in p.doms.values():
            d.sync()
            d.kernel_times(clear=True)
            d.launch_times(clear=True)
            d.set_timing(False)

    if args.calibrate_mib:
        from xkaapi_amd import resolver as R
        nb = args.calibrate_mib << 20
        src = torch.empty(nb // 4, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        R.device_copy(dst.data_ptr(), src.data_ptr(), nb, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_mib} MiB")

    nacc = sum(len(p.st) for p in preps)
    res = {}
    for form in forms:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            for p in preps:
                p.run(form)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t1) / args.steps
        edges = {p.name: p.doms[form].sync() for p in preps}
        # per-kernel and per-launch times from one more pass with HIP-event timing, which
        # slows launches (an event pair per kernel), so the timed steps above run without it
        for p in preps:
            p.doms[form].set_timing(True)
            p.run(form)
        kt = {p.name: p.doms[form].kernel_times(clear=True) for p in preps}
        lt = {p.name: p.doms[form].launch_times(clear=True) for p in preps}
        for p in preps:
            assert p.doms[form].sync() == edges[p.name], "the instrumented pass must find the same edges"
            p.doms[form].set_timing(False)
        res[form] = dict(ms=dt * 1e3, rate=nacc / dt, kt=kt, lt=lt, edges=edges)
        log(f"{form}: {dt*1e3:.2f} ms/step, {nacc/dt:.4g} acc/s, edges {edges}")
        for name, (k, n) in kt.items():
            log(f"   {name}: launches {n} " + " ".join(f"{kk}={vv/max(n,1)*1e3:.1f}us" for kk, vv in k.items()
                                                    if kk.startswith("k_")))

    if "plain" not in res:
        res["plain"] = res["batched"]
    if "batched" not in res:
        res["batched"] = res["plain"]
    # per-batch latency (plain form): launch start -> edges in HBM, task weighted
    lat_us, wts = [], []
    for p in preps:
        l = res["plain"]["lt"][p.name] * 1e3
        nb = len(p.batches)
        l = l.reshape(args.steps, nb) if len(l) == args.steps * nb else l[:nb][None, :]
        for row in l:
            lat_us.extend(row.tolist())
            wts.extend(p.ntasks_plain)
    p50 = weighted_pct(lat_us, wts, 50)
    p99 = weighted_pct(lat_us, wts, 99)

    # k_resolve on the Cholesky stream, plain form: algorithmic bytes 24 N_acc + 64 sum F_i + 12 E;
    # tile-aligned accesses overlap exactly one fragment of the reference's partition, so
    # sum F_i = N_acc.  (The headline roofline is the random stream's, below: it dominates the step.)
    roof_chol = None
    if "cholesky_8192_nb128" in res["plain"]["kt"]:
        kt, nl = res["plain"]["kt"]["cholesky_8192_nb128"]
        n_a = len(chol)
        e = res["plain"]["edges"]["cholesky_8192_nb128"]
        per_pass = len(next(p for p in preps if p.name == "cholesky_8192_nb128").batches)  # launches of one pass
        b_alg = (24 * n_a + 64 * n_a + 12 * e) / per_pass  # one pass resolves every access and edge once
        t_res = kt["k_resolve"] / nl * 1e-3  # nl timed launches (all passes)
        achieved = b_alg / t_res / 1e9
        traffic = prof_us = None
        tp = os.path.join(ROOT, "profiles", "resolver_cholesky8192_plain_traffic.json")
        if os.path.exists(tp):  # HBM bytes per launch from the committed rocprofv3 counter runs
            t = json.load(open(tp))
            traffic = int(1024 * (t["fetch_kib_per_launch"] * t["fetch_factor"]
                                  + t["write_kib_per_launch"] * t["write_factor"]))
            # the same kernels' duration in that profile's kernel trace; the start / stop events
            # above add the dispatch of each kernel (~2.3 us per kernel on this part)
            prof_us = t.get("rocprof_avg_us_per_launch")
        roof_chol = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                     "kernel": "k_resolve", "avg_us": round(t_res * 1e6, 2), "rocprof_avg_us": prof_us,
                     "bytes_per_launch": int(b_alg),
                     "workload": "cholesky 8192^2 nb128, 4096-task launches (plain form)"}
    # the dominant kernel of the step: k_resolve of the random stream's plain launches
    roof = roofline_leg(torch) if rnd is not None else roof_chol
    log(f"roofline: {roof}")
    e2e = None
    if not args.no_e2e and rnd is not None:
        e2e = e2e_leg(chol, rnd, {p.name: p.merge for p in preps})
        log(f"end to end: {e2e}")

    cpu = None if args.no_cpu else cpu_baseline(chol)
    if cpu:
        log(f"cpu baseline: {cpu}")
    coh = None
    if not args.no_coherency:
        coh, coh_subs = coherency_leg()
        if not args.no_cpu:
            coh["cpu_baseline"] = coherency_cpu_baseline(coh_subs)
        log(f"coherency: {coh}")
    heat = None if args.no_heat else heat_leg(torch, args.heat_steps)
    if heat:
        if not args.no_cpu:
            heat["cpu_baseline"] = cpu_baseline_heat()
        log(f"heat: {heat}")
    serving = None if args.no_serving else serving_leg(cpu=not args.no_cpu)
    if serving:
        log(f"serving: {serving}")
    rows = None
    if not args.no_rows:
        rows = {}
        for name, fn in (("segments", segments_leg), ("release", release_leg), ("handles", handles_leg),
                         ("fetch_exec", lambda: fetch_exec_leg(torch))):
            rows[name] = fn()
            log(f"{name}: {rows[name]}")
    out = {
        "metric": "resolved task accesses/sec + p99 submit-to-ready µs, batched vs plain GPU",
        "value": round(res["plain"]["rate"], 1),
        "unit": "accesses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["plain"]["ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "plain GPU resolver, one 4096-task batch per launch: "
                               + " + ".join(p.name for p in preps),
                   "accesses": nacc, "batch_tasks": 4096, "parallelism": "single GPU"},
        "batched": {"value": round(res["batched"]["rate"], 1), "ms_per_step": round(res["batched"]["ms"], 3),
                    "batches_per_launch": {p.name: p.merge for p in preps},
                    "batches_per_launch_tried_ms": {p.name: p.tried for p in preps},
                    "speedup_vs_plain": round(res["batched"]["rate"] / res["plain"]["rate"], 3)},
        "p50_batch_latency_us": round(p50, 2),
        "p99_batch_latency_us": round(p99, 2),
        "edges": res["plain"]["edges"],
        "roofline": roof,
        "roofline_cholesky": roof_chol,
        "end_to_end": e2e,
        "cpu_baseline": cpu,
        "coherency": coh,
        "heat": heat,
        "serving": serving,
        "rows": rows,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
