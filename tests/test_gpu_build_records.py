"""Host record building (xkrt_hip_dep_build_records) on batches large enough to be
built in parallel chunks: the records equal the closed form of access.hpp:205-287
(xkaapi_amd/streams.c, an implementation of its own), and a batch's error is the
one of its earliest failing access, as one pass in order finds it."""
import numpy as np
import pytest

from xkaapi_amd import resolver as R
from xkaapi_amd import streams as S

pytestmark = pytest.mark.gpu

EINVAL, ERANGE, ENOTSUP = 22, 34, 95


def expected(st, first_id, dom):
    """Records from the stream's own rect decomposition; frame base 0 (every column < 2^30)."""
    r, nr = st.rects()
    tl = np.r_[0, np.cumsum(np.diff(st.task.astype(np.int64)) != 0)]
    rows = []
    for j in range(2):
        sel = np.flatnonzero(nr > j)
        a = np.zeros(len(sel), R.RECT_DTYPE)
        a["x0"], a["x1"] = r[sel, 4 * j], r[sel, 4 * j + 1]
        a["y0"], a["y1"] = r[sel, 4 * j + 2], r[sel, 4 * j + 3]
        a["acc"] = first_id + sel
        a["tm"] = (tl[sel] << 8) | st.mode[sel]
        a["dom"] = dom
        rows.append((sel, j, a))
    key = np.concatenate([2 * s + j for s, j, _ in rows])
    return np.concatenate([a for _, _, a in rows])[np.argsort(key, kind="stable")]


def build(st, first_id=7):
    acc = R.to_access_array(st.task, st.mode, st.addr, st.m, st.n)
    with R.DependencyDomain(int(st.ld[0]), int(st.s[0]), max_rects=1 << 20) as d:
        out, ylo, yhi = d.build_records(acc, first_id)
        return out, ylo, yhi, d.index


def wrapping(n, seed, ld=4096, s=4):
    """Narrow tiles, some wrapping past the column's end (two rects): nothing wide
    precedes a wrapped tile, so none is refused."""
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 4, n)
    task = np.repeat(np.arange(n, dtype=np.uint32), k)[:n]
    m = rng.integers(1, 5, n).astype(np.uint64)
    row = rng.integers(0, ld, n).astype(np.uint64)
    col = rng.integers(0, 50000, n).astype(np.uint64)
    return S.Stream(task, np.zeros(n, np.uint16), rng.choice(np.array([1, 2, 3], np.uint8), n),
                    np.ones(n, np.uint8), (col * ld + row) * s, np.full(n, ld, np.uint64), m,
                    rng.integers(1, 4, n).astype(np.uint64), np.full(n, s, np.uint64))


@pytest.mark.parametrize("name", ["random", "heat", "wrapping"])
def test_large_batch_equals_closed_form(name):
    st = {"random": lambda: S.random2d(300_000),
          "heat": lambda: S.heat(2048, 128, 4, 64),
          "wrapping": lambda: wrapping(250_000, 3)}[name]()
    print(f"build records {name}: {len(st)} accesses", flush=True)
    assert len(st) >= 1 << 16  # the parallel path
    out, ylo, yhi, dom = build(st)
    want = expected(st, 7, dom)
    assert len(out) == len(want)
    if name == "wrapping":
        assert len(out) > len(st)  # wrapped tiles give two records
    for f in ("x0", "x1", "y0", "y1", "acc", "tm", "dom"):
        assert np.array_equal(out[f], want[f]), f
    assert ylo == int(want["y0"].min()) and yhi == int(want["y1"].max())


def refused(acc, ld=4096, s=4):
    with R.DependencyDomain(ld, s, max_rects=1 << 20) as d:
        with pytest.raises(R.ResolverError) as ei:
            d.build_records(acc, 0)
        return ei.value.rc


def narrow(n, ld=4096, s=4):
    st = wrapping(n, 5, ld, s)
    acc = R.to_access_array(st.task, st.mode, st.addr, st.m, st.n)
    # no wrap anywhere: every tile stays in its column
    acc["addr"] = (acc["addr"] // (ld * s)) * (ld * s) + 64 * s
    return acc


def wide(acc, i, ld=4096, s=4):
    # a tile of ld - 1 rows from the column's start: wider in dim 0 than any wrapped tile's gap
    acc["addr"][i] = (acc["addr"][i] // (ld * s)) * (ld * s)
    acc["m"][i] = ld - 1


def wrapped(acc, i, ld=4096, s=4):
    # a tile of 8 rows starting 4 rows before the column's end: it wraps into the next column
    acc["addr"][i] = (acc["addr"][i] // (ld * s)) * (ld * s) + (ld - 4) * s
    acc["m"][i] = 8


def test_first_error_wins_across_chunks():
    n = 200_000
    # a wide rect early (chunk 0) makes a wrapped tile far later unsupported: the refusal needs
    # the state every chunk before it leaves
    acc = narrow(n)
    wide(acc, 10)
    wrapped(acc, 190_000)
    assert refused(acc) == -ENOTSUP
    # the same wrapped tile after an invalid access: the earlier error wins
    acc["m"][120_000] = 0
    assert refused(acc) == -EINVAL
    # the invalid access after the refused one: the refusal wins
    acc = narrow(n)
    wide(acc, 10)
    wrapped(acc, 90_000)
    acc["m"][150_000] = 0
    assert refused(acc) == -ENOTSUP
    # without the wide rect, the wrapped tiles are accepted
    acc = narrow(n)
    wrapped(acc, 90_000)
    wrapped(acc, 190_000)
    with R.DependencyDomain(4096, 4, max_rects=1 << 20) as d:
        out, _, _ = d.build_records(acc, 0)
    assert len(out) == n + 2
