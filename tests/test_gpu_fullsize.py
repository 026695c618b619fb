"""GPU parity at the workloads' real sizes, run the way the benchmark runs them.

* heat halo stream 16384^2 fp32, 128^2 tiles, 200 steps (16.3M accesses):
  bins and capacity class from the autotuner, geometric batches (mean 2048
  tasks, max 65536), plain (one launch per batch) and batched (launches of up
  to 2^20 records) -- against the oracle's edge digest;
* the first 102,400 random single-access tasks with the benchmark's bins
  (2048 B x 256 columns, capacity class 1), plain and batched -- against the
  oracle's digest, which the reference's own engine reproduces;
* 64 serving clients (Cholesky / LU, n in [4, 48]) through the serving engine
  with its defaults (prearmed launches, 32,768-rect cap, 16 GB budget): the
  closed loop in both forms and open-loop arrivals at 70% of the batched
  saturation rate -- every client against the oracle.

Digests (edge count and SHA-256 of the sorted pred << 32 | succ keys) come from
tests/golden/make_digests.py."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import khp
from xkaapi_amd import autotune as A
from xkaapi_amd import resolver as R
from xkaapi_amd import serving as V
from xkaapi_amd import streams as S

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
_DIGESTS = {}


def digests():
    if not _DIGESTS:
        with open(os.path.join(HERE, "golden", "digests.json")) as fh:
            _DIGESTS.update(json.load(fh))
    return _DIGESTS


def digest(e):
    e = np.asarray(e, np.uint64).reshape(-1, 2)
    k = np.sort((e[:, 0] << np.uint64(32)) | e[:, 1])
    return int(len(k)), hashlib.sha256(k.astype("<u8").tobytes()).hexdigest()


def want(name):
    d = digests()[name]
    return d["edges"], d["sha256"]


def merged_groups(batches, cap_records):
    """Consecutive batches merged into launches of at most cap_records records
    (2 per access at most), as the benchmark's batched heat form."""
    groups, cur = [], None
    for a, b in batches:
        if cur is not None and 2 * (b - cur[0]) <= cap_records:
            cur = (cur[0], b)
        else:
            if cur is not None:
                groups.append(cur)
            cur = (a, b)
    groups.append(cur)
    return groups


def resolve_device(st, spans, **dom_kw):
    """One launch per span from device-resident records (xkrt_hip_dep_submit_device)."""
    import torch
    acc = R.to_access_array(st.task, st.mode, st.addr, st.m, st.n)
    maxrec = 2 * max(b - a for a, b in spans)
    # the domain launches on torch's stream, behind the uploads; every upload stays alive
    # until the edges are out (a freed block could be refilled before a launch reads it)
    stream = torch.cuda.current_stream().cuda_stream
    with R.DependencyDomain(int(st.ld[0]), int(st.s[0]), max_rects=maxrec, max_edges=4 * len(st) + (1 << 20),
                            stream=stream, **dom_kw) as dom:
        keep = []
        for a, b in spans:
            recs, ylo, yhi = dom.build_records(acc[a:b], a)
            keep.append(torch.from_numpy(recs.view(np.int32).reshape(-1)).cuda())
            dom.submit_device(keep[-1].data_ptr(), len(recs), b - a, ylo, yhi)
        e = dom.edges()
        del keep
    return e


@pytest.fixture(scope="module")
def heat():
    st = S.heat(16384, 128, 4, 200)
    sizes = S.geometric_batches(st.ntasks, 2048.0, 65536, 0x3B)
    return st, R.batches_by_tasks(st.task, sizes)


@pytest.fixture(scope="module")
def heat_tune(heat):
    st, _ = heat
    return A.tune(st.take(slice(0, 300_000)), tasks_per_batch=2048, max_cands=4)


@pytest.mark.parametrize("form", ["plain", "batched"])
def test_heat_full_stream(heat, heat_tune, form):
    st, batches = heat
    assert len(st) == digests()["heat_16384_t128_s200"]["accesses"]
    spans = batches if form == "plain" else merged_groups(batches, 1 << 20)
    if form == "batched":
        assert max(b - a for a, b in spans) > 1 << 18  # launches of several 100K records
    e = resolve_device(st, spans, bin_x_bytes=heat_tune.bin_x_bytes, bin_y_cols=heat_tune.bin_y_cols,
                       kernel=heat_tune.kernel)
    assert digest(e) == want("heat_16384_t128_s200"), (form, heat_tune)


@pytest.mark.parametrize("tasks_per_launch", [4096, 102400])
def test_random2d_prefix_102400(tasks_per_launch):
    d = digests()["random2d_102400"]
    st = S.random2d(d["accesses"])
    e = R.resolve_stream(st, tasks_per_batch=tasks_per_launch, bin_x_bytes=2048, bin_y_cols=256, kernel=1,
                         max_rects=2 * tasks_per_launch)
    assert digest(e) == want("random2d_102400")


def test_cholesky_8192_digest():
    st = S.cholesky(8192, 128)
    e = R.resolve_stream(st, tasks_per_batch=4096, bin_x_bytes=1024, bin_y_cols=128, kernel=0)
    assert digest(e) == want("cholesky_8192_nb128")


@pytest.fixture(scope="module")
def clients64():
    cs = V.make_clients(64, 4, 48, 128, 256)
    for c in cs:
        r, nr = c.stream.rects()
        c.want = digest(khp.resolve(c.stream.task, c.stream.mode, r, nr))
    return cs


def check_all(res, cs):
    for c in cs:
        got = np.asarray(res.edges[c.cid], np.int64).reshape(-1, 2)
        assert digest(got) == c.want, (res.mode, c.cid, c.kind, c.ntiles)


def test_serving_64_clients(clients64):
    cs = clients64
    sat = V.saturation(cs, "batched")
    check_all(sat, cs)
    plain = V.saturation(cs, "plain")
    check_all(plain, cs)
    rate = sum(len(c.batches) for c in cs) / sat.makespan_s
    res = V.serve(cs, V.poisson_schedule(cs, 0.7 * rate, seed=11), "batched")
    check_all(res, cs)
    assert res.stats["error"] == 0
