"""Fetch-plan execution on the GPU with real data: a pinned host matrix and a
device arena of chunks, a coherency tree with fetches in flight, the fetches
executed by xkrt_hip_mcc_exec in random batches and completed out of order.

* After every batch, every byte of the host matrix and of every placed chunk
  equals a host simulation that applies the same fetch rows (bit-exact copy
  execution, H2D, D2D and D2H, forwards included).
* Once a window's fetches are done, every GPU read sees the data of the last
  write of each element, through its own chunk (the plan moves the right
  bytes): tasks write a signature into their chunk after their reads.

Host accesses only read here, each planned alone: the reference moves data to
the host from the largest GPU view, coherent or not (memory-tree.hpp:1155-1163),
so a host read can see an old GPU copy (the simulation check covers those
bytes), and planned beside GPU reads it could copy from a view that one of
their fetches is still writing."""
import numpy as np
import pytest

from oracle import khp
from xkaapi_amd import coherency as K
from xkaapi_amd import streams as S

pytestmark = pytest.mark.gpu

LD, COLS, SZ = 64, 48, 4
INIT = 1_000_000
ARENA = 16 << 20


def stream(seed, n=300):
    rng = np.random.default_rng(seed)
    kind = rng.choice(4, n, p=[0.45, 0.15, 0.15, 0.25])  # GPU R, GPU W, GPU RW, host R
    mode = np.array([1, 2, 3, 1], np.uint8)[kind]
    dev = (kind < 3).astype(np.uint8)
    row = rng.integers(0, LD, n).astype(np.uint64)
    col = rng.integers(0, COLS, n).astype(np.uint64)
    m = np.minimum(rng.integers(1, 25, n).astype(np.uint64), LD - row)
    nn = np.minimum(rng.integers(1, 17, n).astype(np.uint64), COLS - col)
    st = S.Stream(np.arange(n, dtype=np.uint32), np.zeros(n, np.uint16), mode, dev, (col * LD + row) * SZ,
                  np.full(n, LD, np.uint64), m, nn, np.full(n, SZ, np.uint64))
    r, nr = st.rects()
    fail = khp.coherency(st.dev, st.mode, r, nr, st.m * st.n * st.s).fail_at  # stop before the ninth view
    return st if fail < 0 else st.take(slice(0, fail))


class Sim:
    """The same memory on the host: the matrix (COLS x LD elements, column-major) and one
    array per chunk over the rect its allocating access covers."""

    def __init__(self, host):
        self.host = host.copy()
        self.chunks = {}  # chunk -> (rect in elements, array (cols, rows))

    def chunk(self, t, c):
        if c not in self.chunks:
            _, _, rect = t.chunk_ptr(c)
            u0, u1, y0, y1 = int(rect[0]) // SZ, int(rect[1]) // SZ, int(rect[2]), int(rect[3])
            self.chunks[c] = ((u0, u1, y0, y1), np.zeros((y1 - y0, u1 - u0), np.int32))
        return self.chunks[c]

    def region(self, t, dev, c, u0, u1, y0, y1):
        if dev == 0:
            return self.host[y0:y1, u0:u1]
        (cu0, _, cy0, _), a = self.chunk(t, c)
        return a[y0 - cy0:y1 - cy0, u0 - cu0:u1 - cu0]

    def copy(self, t, row):
        x0, x1, y0, y1, sd, sc, dd, dc = (int(v) for v in row)
        u0, u1 = x0 // SZ, x1 // SZ
        self.region(t, dd, dc, u0, u1, y0, y1)[...] = self.region(t, sd, sc, u0, u1, y0, y1).copy()


def device_chunk(t, arena_np, arena_base, c, rect):
    ptr, pitch, _ = t.chunk_ptr(c)
    u0, u1, y0, y1 = rect
    off = ptr - arena_base
    return arena_np[off:off + pitch * (y1 - y0)].view(np.int32).reshape(y1 - y0, pitch // SZ)


@pytest.mark.parametrize("seed", range(4))
def test_fetch_execution_moves_the_planned_bytes(seed):
    import torch
    st = stream(seed)
    n = len(st)
    rng = np.random.default_rng(1000 + seed)
    host_t = torch.empty(COLS * LD, dtype=torch.int32, pin_memory=True)
    host = host_t.numpy().reshape(COLS, LD)
    host[...] = INIT + np.arange(COLS * LD, dtype=np.int32).reshape(COLS, LD)
    truth = host.copy()
    arena = torch.zeros(ARENA, dtype=torch.uint8, device="cuda")
    a32 = arena.view(torch.int32)
    sim = Sim(host)
    acc = K.to_mcc_array(st.dev, st.mode, st.addr, st.m, st.n)
    cstream = torch.cuda.current_stream().cuda_stream
    seen_fwd = seen_await = 0

    def placed(t):
        """The arena up to the end of the last placed chunk, on the host."""
        end = 0
        for c, ((u0, u1, y0, y1), _) in sim.chunks.items():
            ptr, pitch, _ = t.chunk_ptr(c)
            end = max(end, ptr - arena.data_ptr() + pitch * (y1 - y0))
        return arena[:end].cpu().numpy()

    def check_memory(t):
        torch.cuda.synchronize()
        assert np.array_equal(host, sim.host)
        arena_np = placed(t)
        for c, (rect, arr) in sim.chunks.items():
            assert np.array_equal(device_chunk(t, arena_np, arena.data_ptr(), c, rect), arr), c

    with K.MemoryTree(LD, SZ, (0, COLS), async_completion=True) as t:
        t.exec_bind(host_t.data_ptr(), arena.data_ptr(), ARENA)
        i = 0
        while i < n:
            # a window: a run of GPU reads planned together (their fetches are in flight together:
            # awaits, forwards), or one writer, or one host read.  A host read takes the largest GPU
            # view as its source even when that view is not coherent, so planned beside GPU reads
            # it could copy from a view another fetch of the window is still writing.
            j = i + 1
            if st.dev[i] == 1 and st.mode[i] == 1:
                while j < n and j - i < 6 and st.dev[j] == 1 and st.mode[j] == 1:
                    j += 1
            plan = t.plan(acc[i:j])
            chunks = t.access_chunks(j - i)
            off, regs = t.awaits(j - i)
            seen_await += int((regs[:, 4] == 2).sum()) if len(regs) else 0
            base = t.stats()["last_fetch_base"]
            rows = {base + k: plan.fetches[k] for k in range(len(plan.fetches))}
            while rows:
                ids = [int(x) for x in rng.choice(sorted(rows), min(len(rows), int(rng.integers(1, 4))),
                                                  replace=False)]
                t.exec(ids, cstream)
                torch.cuda.synchronize()
                for k in ids:
                    sim.copy(t, rows.pop(k))
                _, _, _, fwd = t.complete(ids)
                for f in fwd:  # a forward: from the view just fetched to another view
                    rows[int(f[0])] = f[1:]
                    seen_fwd += 1
                check_memory(t)
            # every read of the window sees the last write, through its own chunk
            for k in range(i, j):
                if st.dev[k] == 1:
                    sim.chunk(t, int(chunks[k - i]))
            arena_np = placed(t)
            for k in range(i, j):
                u0 = int(st.addr[k]) // SZ % LD
                y0 = int(st.addr[k]) // SZ // LD
                u1, y1 = u0 + int(st.m[k]), y0 + int(st.n[k])
                if st.dev[k] == 1 and (st.mode[k] & 1):
                    rect, _ = sim.chunk(t, int(chunks[k - i]))
                    got = device_chunk(t, arena_np, arena.data_ptr(), int(chunks[k - i]), rect)
                    cu0, _, cy0, _ = rect
                    assert np.array_equal(got[y0 - cy0:y1 - cy0, u0 - cu0:u1 - cu0], truth[y0:y1, u0:u1]), (seed, k)
            # the writer's task writes its signature through its chunk
            if st.mode[i] & 2:
                c = int(chunks[0])
                (cu0, cu1, cy0, cy1), arr = sim.chunk(t, c)
                ptr, pitch, _ = t.chunk_ptr(c)
                u0 = int(st.addr[i]) // SZ % LD
                y0 = int(st.addr[i]) // SZ // LD
                u1, y1 = u0 + int(st.m[i]), y0 + int(st.n[i])
                sig = np.int32(2_000_000 + i)
                yy, uu = np.meshgrid(np.arange(y0, y1), np.arange(u0, u1), indexing="ij")
                idx = ((ptr - arena.data_ptr()) + (yy - cy0) * pitch + (uu - cu0) * SZ) // SZ
                a32[torch.from_numpy(idx.reshape(-1)).cuda()] = int(sig)
                arr[y0 - cy0:y1 - cy0, u0 - cu0:u1 - cu0] = sig
                truth[y0:y1, u0:u1] = sig
                check_memory(t)
            i = j
        assert t.stats()["in_flight"] == 0
    assert seen_fwd + seen_await > 0 or seed > 0
