"""Host-side parts of the autotuner (no GPU): candidate bin extents and the
LDS capacity class from the stream's geometry."""
import numpy as np

from xkaapi_amd import autotune as A
from xkaapi_amd import resolver as R
from xkaapi_amd import streams as S


def test_autotune_candidates_and_capacity():
    # candidate bins follow the tile geometry; the class follows the densest bin of a launch
    st = S.cholesky(2048, 128)
    c = A.candidates(st)
    assert c[0] == (1024, 128)  # one 128 x 128 fp64 tile per bin
    assert all((bx & (bx - 1)) == 0 and (by & (by - 1)) == 0 for bx, by in c)
    # whole stream in one launch: the diagonal tiles get many records
    k_all = A.capacity_class(st, 1024, 128, st.ntasks)
    k_small = A.capacity_class(st, 1024, 128, 16)
    assert k_small <= k_all
    # brute force of the densest bin for a small launch
    r, nr = st.rects()
    a, b = R.batches_by_tasks(st.task, 64)[0]
    cnt = {}
    for i in range(a, b):
        x0, x1, y0, y1 = (int(v) for v in r[i, :4])
        for gy in range(y0 // 128, (y1 - 1) // 128 + 1):
            for gx in range(x0 // 1024, (x1 - 1) // 1024 + 1):
                cnt[(gx, gy)] = cnt.get((gx, gy), 0) + 1
    worst = max(cnt.values())
    want = next(k for k, cap in enumerate(A.KCLASS_RECORDS) if 2 * worst <= cap) if 2 * worst <= 1024 else 2
    assert A.capacity_class(st, 1024, 128, 64) == want


def test_spread_windows_cover_the_stream():
    # windows of whole tasks at the stream's start, end and evenly between, tasks renumbered
    st = S.heat(1024, 128, 4, 20)
    starts = st.task_starts()
    w = A.spread_windows(st, nwin=4, win_tasks=100)
    assert w.ntasks == 400 and (np.diff(w.task.astype(np.int64)) >= 0).all()
    ntask = len(starts) - 1
    for k in range(4):
        t0 = (ntask - 100) * k // 3
        a, b = int(starts[t0]), int(starts[t0 + 100])
        got = w.take(slice(*[int(x) for x in w.task_starts()[[100 * k, 100 * (k + 1)]]]))
        assert np.array_equal(got.addr, st.addr[a:b]) and np.array_equal(got.mode, st.mode[a:b])
    assert A.spread_windows(st, nwin=4, win_tasks=ntask) is st  # a short stream is its own sample
