"""ORACLE (test infrastructure): brute-force per-point semantics on tiny spaces.

No tree and no fragments: every byte offset x of every column y carries its
own state.  This pins the closed-form semantics of the reference for
single-rect accesses:

Resolver (dependency-tree.hpp:229-256, 311-355; task.hpp:494-512;
src/task/task.cc:285-299), per point x: LW(x) last writer, RS(x) readers
since LW.  For task T with accesses a_0..a_{m-1}:
  LINK a_i (slot order): P = U_x ((W in mode and RS(x)) ? RS(x) : {LW(x)} - {none});
                         edge p -> a_i for p in P unless p already has one to T
  PUT a_i (slot order):  W in mode -> LW = a_i, RS = [];  mode == R -> RS += a_i

Coherency (memory-tree.hpp, one tree, host = device 0, one GPU = device 1,
synchronous completion, no eviction), per point x: coh(x) device bits,
views(x) chronological chunk list on the GPU, vcoh(x) coherent chunks.
"""
from __future__ import annotations

import numpy as np

M_R, M_W = 1, 2


def _points(rect_row, nrect):
    """List of (x0, x1, y0, y1) rects of one access."""
    out = []
    for r in range(int(nrect)):
        q = rect_row[4 * r:4 * r + 4]
        out.append(tuple(int(v) for v in q))
    return out


def resolve(task, mode, rects, nrect, xdim, ydim, put=None, done=None):
    """Edges as a set of (pred, succ) access indices.  put[a] False: the
    access links but its PUT changes nothing (a segment seen by a BLAS domain
    whose own rects meet no node of it).  done[a]: the accesses p < done[a]
    belong to completed tasks when a is LINKed; the state keeps them, but
    __task_precedes (include/xkrt/task/task.hpp:452-487) makes no edge from
    a completed predecessor and __access_link records no successor."""
    n = len(task)
    LW = np.full((xdim, ydim), -1, np.int64)
    RS = [[[] for _ in range(ydim)] for _ in range(xdim)]
    edges = set()
    linked = {}  # pred -> last task linked to
    i = 0
    while i < n:
        j = i
        while j < n and task[j] == task[i]:
            j += 1
        T = int(task[i])
        for a in range(i, j):
            w = bool(mode[a] & M_W)
            preds = set()
            for (x0, x1, y0, y1) in _points(rects[a], nrect[a]):
                for x in range(x0, x1):
                    for y in range(y0, y1):
                        rs = RS[x][y]
                        if w and rs:
                            preds.update(rs)
                        elif LW[x, y] >= 0:
                            preds.add(int(LW[x, y]))
            for p in sorted(preds):
                if done is not None and p < done[a]:
                    continue
                if linked.get(p) != T:
                    linked[p] = T
                    edges.add((p, a))
        for a in range(i, j):
            if put is not None and not put[a]:
                continue
            for (x0, x1, y0, y1) in _points(rects[a], nrect[a]):
                for x in range(x0, x1):
                    for y in range(y0, y1):
                        if mode[a] & M_W:
                            LW[x, y] = a
                            RS[x][y] = []
                        elif mode[a] == M_R:
                            RS[x][y] = RS[x][y] + [a]
        i = j
    return edges


def coherency(dev, mode, rects, nrect, chunk_bytes, xdim, ydim, want_state=False):
    """Per access: (coh_before map {(x,y): bits}, fetch map {(x,y): (src_dev,
    src_chunk, dst_dev, dst_chunk)}), or the index of the access that would
    exceed 8 allocation views (then the lists stop before it).  With
    want_state, the per-point coherency bits at the end come third."""
    n = len(dev)
    coh = np.zeros((xdim, ydim), np.int64)
    views = [[[] for _ in range(ydim)] for _ in range(xdim)]
    vcoh = [[set() for _ in range(ydim)] for _ in range(xdim)]
    chunk_size = []
    gpu_seen = False
    out = []
    for a in range(n):
        d = int(dev[a])
        m = int(mode[a])
        pts = [(x, y) for (x0, x1, y0, y1) in _points(rects[a], nrect[a])
               for x in range(x0, x1) for y in range(y0, y1)]
        if d == 0 and not gpu_seen:
            out.append(({}, {}))
            continue
        before = {p: int(coh[p]) for p in pts}
        fetch = {}
        if d == 0:
            for p in pts:
                c = int(coh[p])
                if c == 0 or (c & 1):
                    continue
                vs = views[p[0]][p[1]]
                best = 0
                for k in range(1, len(vs)):
                    if chunk_size[vs[best]] < chunk_size[vs[k]]:
                        best = k
                fetch[p] = (1, vs[best], 0, -1)
            for p in fetch:
                coh[p] |= 1
            out.append((before, fetch))
            continue
        gpu_seen = True
        common = None
        for p in pts:
            s = set(views[p[0]][p[1]])
            common = s if common is None else (common & s)
        if common:
            chunk = min(common)  # chunks are numbered in allocation order
        else:
            for p in pts:
                if len(views[p[0]][p[1]]) >= 8:
                    return (out, a, coh) if want_state else (out, a)
            chunk = len(chunk_size)
            chunk_size.append(int(chunk_bytes[a]))
            for p in pts:
                views[p[0]][p[1]].append(chunk)
        if m & M_R:
            for p in pts:
                vc = vcoh[p[0]][p[1]]
                if chunk in vc:
                    continue
                if coh[p] & 2:
                    # lowest-index coherent view = oldest coherent chunk
                    src_chunk = min(vc)
                    fetch[p] = (1, src_chunk, 1, chunk)
                else:
                    fetch[p] = (0, -1, 1, chunk)
        if m & M_W:
            for p in pts:
                vcoh[p[0]][p[1]] = {chunk}
                coh[p] = 2
        for p in fetch:
            vcoh[p[0]][p[1]].add(chunk)
            coh[p] |= 2
        out.append((before, fetch))
    return (out, -1, coh) if want_state else (out, -1)


def coherency_events(dev, mode, rects, nrect, chunk_bytes, ev_type, ev_arg, log, xdim, ydim,
                     forget_fetch_clear=False):
    """Replays an event schedule point by point with fetches in flight
    (memory-tree.hpp:1110-1215, 1471-1667, 955-1061, 2067-2127) and returns
    per event the expected per-point outcome, to check an event log of
    khp.coherency_events:
      ("plan", a, before {p: (coh, fetching)}, fetch {p: (src dev, src chunk, dst dev, dst chunk)},
       await {p: chunk}, forward {p: (src chunk, dst chunk)})
      ("done", fired {p: set of (kind, access, dst chunk)})      kind 4 released, 5 forward issued
    Registrations are kept per point and per view; a completion fires those of
    the points it covers.  The fetches' rects and ids come from the log (kinds 1
    and 5); their per-point content is what is checked.  forget_fetch_clear
    plants a mistake (device fetching bits are never cleared) for the tests."""
    n = len(dev)
    coh = np.zeros((xdim, ydim), np.int64)
    fch = np.zeros((xdim, ydim), np.int64)
    views = [[[] for _ in range(ydim)] for _ in range(xdim)]
    vcoh = [[set() for _ in range(ydim)] for _ in range(xdim)]
    vfet = [[set() for _ in range(ydim)] for _ in range(xdim)]
    regs = [[{} for _ in range(ydim)] for _ in range(xdim)]
    chunk_size = []
    fetches = {}  # id -> (x0, x1, y0, y1, dst dev, dst chunk), from the log
    inflight = []
    for row in log:
        if row[1] in (1, 5):
            fetches[int(row[11])] = (int(row[3]), int(row[4]), int(row[5]), int(row[6]), int(row[9]), int(row[10]))
    issued_by_event = {}
    for row in log:
        if row[1] in (1, 5):
            issued_by_event.setdefault(int(row[0]), []).append(int(row[11]))
    gpu_seen = False
    out = []
    a = 0
    for ev in range(len(ev_type)):
        if ev_type[ev] == 1:
            if not inflight:
                out.append(("done", {}))
                continue
            fid = sorted(inflight)[int(ev_arg[ev]) % len(inflight)]
            inflight.remove(fid)
            x0, x1, y0, y1, dd, c = fetches[fid]
            fired = {}
            for x in range(x0, x1):
                for y in range(y0, y1):
                    if dd == 1:
                        if c in views[x][y]:
                            vcoh[x][y].add(c)
                            vfet[x][y].discard(c)
                        for key in regs[x][y].pop(c, []):
                            fired.setdefault((x, y), set()).add(key)
                        coh[x, y] |= 2
                        if not vfet[x][y] and not forget_fetch_clear:
                            fch[x, y] &= ~2
                    else:
                        coh[x, y] |= 1
                        fch[x, y] &= ~1
            out.append(("done", fired))
            inflight.extend(issued_by_event.get(ev, []))
            continue
        if a >= n:
            out.append(("plan", None, {}, {}, {}, {}))
            continue
        d, m = int(dev[a]), int(mode[a])
        pts = [(x, y) for (x0, x1, y0, y1) in _points(rects[a], nrect[a])
               for x in range(x0, x1) for y in range(y0, y1)]
        if d == 0 and not gpu_seen:
            out.append(("plan", a, {}, {}, {}, {}))
            a += 1
            continue
        before = {p: (int(coh[p]), int(fch[p])) for p in pts}
        fetch, awaits, fwd = {}, {}, {}
        if d == 0:
            for p in pts:
                if coh[p] == 0 or (coh[p] & 1) or (fch[p] & 1):
                    continue
                fch[p] |= 1
                vs = views[p[0]][p[1]]
                best = 0
                for k in range(1, len(vs)):
                    if chunk_size[vs[best]] < chunk_size[vs[k]]:
                        best = k
                fetch[p] = (1, vs[best], 0, -1)
        else:
            gpu_seen = True
            common = None
            for p in pts:
                s = set(views[p[0]][p[1]])
                common = s if common is None else (common & s)
            if common:
                chunk = min(common)
            else:
                for p in pts:
                    if len(views[p[0]][p[1]]) >= 8:
                        out.append(("fail", a))
                        return out
                chunk = len(chunk_size)
                chunk_size.append(int(chunk_bytes[a]))
                for p in pts:
                    views[p[0]][p[1]].append(chunk)
            if m & M_R:
                for p in pts:
                    vc, vf = vcoh[p[0]][p[1]], vfet[p[0]][p[1]]
                    if chunk in vc:
                        continue
                    if chunk in vf:
                        awaits[p] = chunk
                        regs[p[0]][p[1]].setdefault(chunk, []).append((4, a, -1))
                        continue
                    if coh[p] & 2:
                        fetch[p] = (1, min(vc), 1, chunk)
                    elif fch[p] & 2:
                        src = min(vf)
                        fwd[p] = (src, chunk)
                        regs[p[0]][p[1]].setdefault(src, []).append((5, a, chunk))
                    else:
                        fetch[p] = (0, -1, 1, chunk)
                    vf.add(chunk)
                    fch[p] |= 2
            if m & M_W:
                for p in pts:
                    vcoh[p[0]][p[1]] = {chunk}
                    coh[p] = 2
        out.append(("plan", a, before, fetch, awaits, fwd))
        inflight.extend(issued_by_event.get(ev, []))
        a += 1
    return out
