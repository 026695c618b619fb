"""ORACLE (test infrastructure): the runtime's successor release, restated
from the reference for the device release engine's tests.

  __task_precedes   include/xkrt/task/task.hpp:453-486: an edge pred -> succ
                    whose predecessor task is not COMPLETED adds 1 to the
                    successor task's wait counter and puts the successor
                    access on the predecessor access's list
  __task_commit     task.hpp:535-551: the counter starts at 1 (task.hpp:129),
                    the commit takes that 1 away; 0 -> ready
  __task_complete   src/task.cc:160-247: COMPLETED, then for every access not
                    in mode D (199-201), every successor on its list loses 1;
                    0 -> ready

The edges given are already deduplicated per (pred access, succ task)
(__access_precedes, task.hpp:494-512).  A launch applies the edges of the
tasks it resolved, then commits those tasks; the link happens at the launch,
so "completed" is tested there.

Pinned by tests/test_oracle_release.py against a closed form over sets, with
a planted mistake.
"""
from __future__ import annotations

import numpy as np

MODE_D = 8


VIRT = 1 << 31  # edge end: the k-th empty write a handle map inserted (oracle/handles.py)


class Release:
    """Program tasks 0 .. ntask - 1 (accesses 0 .. nacc - 1); an inserted empty
    write VIRT | k is task ntask + k with access nacc + k (mode VW): an empty
    task, completed as soon as it is released (it is never handed out)."""

    def __init__(self, acc_task, acc_mode, nvirt=0):
        self.acc_task = np.asarray(acc_task, np.int64)
        self.acc_mode = np.asarray(acc_mode, np.int64)
        self.nacc = len(self.acc_task)
        self.ntask = int(self.acc_task.max()) + 1 if len(self.acc_task) else 0
        self.acc_task = np.r_[self.acc_task, self.ntask + np.arange(nvirt)]
        self.acc_mode = np.r_[self.acc_mode, np.full(nvirt, 6)]
        n = self.ntask + nvirt
        self.wc = [1] * n               # the commit guard
        self.state = ["alloc"] * n      # alloc -> ready -> done
        self.succ = {}                  # pred access -> [succ task]

    def _acc(self, x):
        x = int(x)
        return self.nacc + (x - VIRT) if x >= VIRT else x

    def _ready(self, t, out):
        if t >= self.ntask:  # an empty write: completes at once
            self.state[t] = "ready"
            out.extend(self.complete([t]))
            return
        self.state[t] = "ready"
        out.append(t)

    def launch(self, edges, t0, t1, skip_completed=True, virt=(0, 0)):
        """Apply a launch's edges, commit the empty writes virt = [k0, k1) it inserted, then
        its tasks [t0, t1); returns the released program tasks.
        skip_completed / honor_detached (complete) False plant mistakes for the tests."""
        out = []
        for p, s in np.asarray(edges, np.int64).reshape(-1, 2):
            p, s = self._acc(p), self._acc(s)
            tp, ts = int(self.acc_task[p]), int(self.acc_task[s])
            if skip_completed and self.state[tp] == "done":
                continue
            self.wc[ts] += 1
            self.succ.setdefault(int(p), []).append(ts)
        for k in range(virt[0], virt[1]):
            t = self.ntask + k
            self.wc[t] -= 1
            if self.wc[t] == 0:
                self._ready(t, out)
        for t in range(t0, t1):
            self.wc[t] -= 1
            if self.wc[t] == 0:
                self._ready(t, out)
        return out

    def complete(self, tasks, honor_detached=True):
        """Complete released tasks; returns the tasks they release."""
        out = []
        for t in tasks:
            t = int(t)
            assert self.state[t] == "ready", t
            self.state[t] = "done"
            for a in ([self.nacc + t - self.ntask] if t >= self.ntask else np.flatnonzero(self.acc_task[:self.nacc] == t)):
                if honor_detached and self.acc_mode[a] & MODE_D:
                    continue
                for s in self.succ.get(int(a), []):
                    self.wc[s] -= 1
                    if self.wc[s] == 0:
                        self._ready(s, out)
        return out


def released_closed_form(edges, acc_task, acc_mode, launched_at, done_at, now):
    """The tasks released after event `now`, from sets: a task launched at or
    before `now` is released unless some edge into it that counted -- its
    predecessor task not completed before the successor's launch -- comes from
    a detached access or from a task not completed by `now`.
    launched_at[t] / done_at[t]: event index of t's launch / completion (inf
    when none)."""
    acc_task = np.asarray(acc_task, np.int64)
    acc_mode = np.asarray(acc_mode, np.int64)
    launched_at = np.asarray(launched_at, np.float64)
    done_at = np.asarray(done_at, np.float64)
    blocked = np.zeros(len(launched_at), bool)
    for p, s in np.asarray(edges, np.int64).reshape(-1, 2):
        tp, ts = acc_task[p], acc_task[s]
        counted = not done_at[tp] < launched_at[ts]
        if counted and ((acc_mode[p] & MODE_D) or not done_at[tp] <= now):
            blocked[ts] = True
    return set(np.flatnonzero((launched_at <= now) & ~blocked).tolist())


def released_closed_form_virtual(edges, acc_task, acc_mode, launched_at, done_at, now, virt_launched_at):
    """released_closed_form with inserted empty writes: edges may name VIRT | k, the
    k-th empty write, launched at virt_launched_at[k]; an empty write completes as
    soon as every counted predecessor has (its predecessors are program tasks), and
    then counts as done for its successors."""
    acc_task = np.asarray(acc_task, np.int64)
    acc_mode = np.asarray(acc_mode, np.int64)
    ntask = len(launched_at)
    vl = np.asarray(virt_launched_at, np.float64)
    vdone = vl.copy()
    e = np.asarray(edges, np.int64).reshape(-1, 2)
    done = np.asarray(done_at, np.float64)
    for p, s in e:
        if s >= VIRT:
            k = s - VIRT
            tp = acc_task[p]
            counted = not done[tp] < vl[k]
            if counted:
                vdone[k] = np.inf if (acc_mode[p] & MODE_D) else max(vdone[k], done[tp])
    vdone = np.where(vdone <= now, vdone, np.inf)
    blocked = np.zeros(ntask, bool)
    launched = np.asarray(launched_at, np.float64)
    for p, s in e:
        if s >= VIRT:
            continue
        ts = acc_task[s]
        if p >= VIRT:
            dp, det = vdone[p - VIRT], False
        else:
            dp, det = done[acc_task[p]], bool(acc_mode[p] & MODE_D)
        counted = not dp < launched[ts]
        if counted and (det or not dp <= now):
            blocked[ts] = True
    return set(np.flatnonzero((launched <= now) & ~blocked).tolist())
