"""The serving workload's native open-loop driver (xks_open_loop in
libxkstreams.so), on the CPU with stand-in server entry points: batches go
out in schedule order, not before their arrival times, each client is closed
right after its last batch, and the first failing call stops the loop."""
import ctypes as C
import time

import numpy as np

from xkaapi_amd import resolver as R
from xkaapi_amd import streams as S

SUBMIT = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64))
CLOSE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32)


def run(t_s, clients, sizes, last, fail_at=-1):
    log = []
    t0 = [None]

    def submit(srv, client, acc, n, req):
        now = time.perf_counter()  # on entry: the callback's own work (a first call is slow) comes after
        if t0[0] is None:
            t0[0] = now
        k = sum(1 for e in log if e[0] == "s")
        if k == fail_at:
            return -5
        a = np.ctypeslib.as_array(C.cast(acc, C.POINTER(C.c_uint8)), shape=(n * R.ACCESS_DTYPE.itemsize,))
        first = int(a.view(R.ACCESS_DTYPE)[0]["task"]) if n else -1
        log.append(("s", int(client), int(n), first, now))
        req[0] = k
        return 0

    def close(srv, client):
        log.append(("c", int(client)))
        return 0

    fs, fc = SUBMIT(submit), CLOSE(close)
    batches = []
    for i, n in enumerate(sizes):
        b = np.zeros(n, R.ACCESS_DTYPE)
        b["task"] = 1000 * i
        batches.append(b)
    n = len(t_s)
    ptrs = (C.c_void_p * max(n, 1))(*[b.ctypes.data for b in batches])
    t = np.asarray(t_s, np.float64)
    cl = np.asarray(clients, np.uint32)
    cnt = np.asarray(sizes, np.uint64)
    ls = np.asarray(last, np.uint8)
    lag = np.zeros(max(n, 1), np.float64)
    req = np.zeros(max(n, 1), np.uint64)
    start = time.perf_counter()
    rc = S._lib().xks_open_loop(None, C.cast(fs, C.c_void_p), C.cast(fc, C.c_void_p), n, t.ctypes.data,
                                cl.ctypes.data, ptrs, cnt.ctypes.data, ls.ctypes.data, lag.ctypes.data,
                                req.ctypes.data)
    return rc, log, lag[:n], req[:n], start


def test_schedule_order_times_and_closes():
    t_s = [0.0, 0.002, 0.002, 0.005, 0.009]
    clients = [3, 1, 3, 1, 3]
    sizes = [4, 1, 2, 7, 3]
    last = [0, 0, 0, 1, 1]
    rc, log, lag, req, start = run(t_s, clients, sizes, last)
    assert rc == 0
    subs = [e for e in log if e[0] == "s"]
    assert [(e[1], e[2], e[3]) for e in subs] == [(3, 4, 0), (1, 1, 1000), (3, 2, 2000), (1, 7, 3000), (3, 3, 4000)]
    # each close right after its client's last batch
    assert [e[0] if e[0] == "s" else ("c", e[1]) for e in log] == ["s", "s", "s", "s", ("c", 1), "s", ("c", 3)]
    # no batch before its arrival time (measured from the first submit, which the loop makes at once)
    t_first = subs[0][4]
    for e, t in zip(subs, t_s):
        assert e[4] - t_first >= t - 1e-4
    assert (lag >= 0).all() and req.tolist() == list(range(5))


def test_first_failing_call_stops_the_loop():
    rc, log, _, _, _ = run([0.0, 0.0, 0.0], [0, 0, 0], [1, 1, 1], [0, 0, 1], fail_at=1)
    assert rc == -5
    assert [e[0] for e in log] == ["s"]  # the failing second submit logged nothing, the third never came


def test_empty_schedule():
    rc, log, _, _, _ = run([], [], [], [])
    assert rc == 0 and log == []
