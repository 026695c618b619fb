"""ORACLE (test infrastructure): the handle dependency map of the reference,
restated for the GPU handle domain's tests
(include/xkrt/memory/access/handle/dependency-map.hpp:52-330).

Per handle: the concurrent writes since the last sequential write
(last_conc_writes), the sequential reads since (last_seq_reads) and the last
sequential write.  LINK (191-282), in this order:
  * reads recorded and the access writes: a concurrent write inserts an empty
    write X (insert_empty_write, 94-135); any other write depends on every
    read (and on nothing else by transitivity);
  * concurrent writes recorded and the access is not a concurrent write: a
    writer depends on every one of them (transitively on nothing else); any
    other access inserts an empty write X;
  * the last sequential write, unless a transitive edge was set.
PUT (284-326): a concurrent write joins the concurrent writes; any other write
clears both lists and becomes the last sequential write; a read (mode & R)
clears the concurrent writes and joins the reads.
The empty write X is a task of its own with one access (mode VW, sequential):
it is LINKed, then PUT at once when it got a predecessor (its commit guard
tests that, 118-131).  Edges are deduplicated as __access_precedes does
(include/xkrt/task/task.hpp:494-512): skipped when the predecessor's last
successor belongs to the same task.  Tasks never complete here, so every edge
is made (link_or_pop never pops, 137-170).

Access ids: the given ones; the k-th empty write gets id VIRT | k, task
VIRT | k.  Pinned by tests/test_oracle_handles.py.
"""
from __future__ import annotations

import numpy as np

VIRT = 1 << 31
MODE_R, MODE_W, MODE_VW = 1, 2, 6
SEQ, COM, CNC = 1, 2, 4


class HandleMap:
    def __init__(self, mistakes=()):
        self.nodes = {}      # handle -> [conc writes], [seq reads], last seq write (or None)
        self.last = {}       # pred id -> task of its last successor
        self.task = {}       # id -> task
        self.edges = []
        self.nvirt = 0
        self.mistakes = set(mistakes)

    def _precedes(self, p, s):
        if "no_dedup" not in self.mistakes and self.last.get(p) == self.task[s]:
            return
        self.last[p] = self.task[s]
        self.edges.append((p, s))

    def _insert_empty_write(self, h):
        x = VIRT | self.nvirt
        self.nvirt += 1
        self.task[x] = x
        n0 = len(self.edges)
        self._link(x, h, MODE_VW, SEQ)
        if len(self.edges) > n0:  # a live predecessor: the empty write stays
            self._put(x, h, MODE_VW, SEQ)

    def _link(self, a, h, mode, conc):
        node = self.nodes.get(h)
        if node is None:
            return
        trans = False
        if node[1] and (mode & MODE_W):
            if conc == CNC and "no_empty_write" not in self.mistakes:
                self._insert_empty_write(h)
            else:
                for p in list(node[1]):
                    self._precedes(p, a)
                trans = True
        if node[0] and conc != CNC:
            if mode & MODE_W:
                for p in list(node[0]):
                    self._precedes(p, a)
                trans = True
            else:
                self._insert_empty_write(h)
        if node[2] is not None and not trans:
            self._precedes(node[2], a)

    def _put(self, a, h, mode, conc):
        node = self.nodes.setdefault(h, [[], [], None])
        if mode & MODE_W:
            if conc == CNC:
                node[0].append(a)
            else:
                node[1].clear()
                node[0].clear()
                node[2] = a
        elif mode & MODE_R:
            node[0].clear()
            node[1].append(a)

    def resolve(self, ids, task, handle, mode, conc):
        """Accesses in spawn order (a task's accesses contiguous): LINK all of a
        task's, then PUT them (src/task/task.cc:285-299)."""
        n = len(ids)
        for i in range(n):
            self.task[int(ids[i])] = int(task[i])
        i = 0
        while i < n:
            j = i
            while j < n and task[j] == task[i]:
                j += 1
            for k in range(i, j):
                self._link(int(ids[k]), int(handle[k]), int(mode[k]), int(conc[k]))
            for k in range(i, j):
                self._put(int(ids[k]), int(handle[k]), int(mode[k]), int(conc[k]))
            i = j
        return self.edges


def resolve(task, handle, mode, conc, mistakes=()):
    """Edges (pred id, succ id) of a handle stream with ids 0..n-1; and the number of empty writes."""
    m = HandleMap(mistakes)
    e = m.resolve(np.arange(len(task)), task, handle, mode, conc)
    return np.asarray(e, np.int64).reshape(-1, 2), m.nvirt
