// Successor release on the device: what the reference does on the host when
// an edge is linked, when a task is committed and when it completes.
//
//   __task_precedes (include/xkrt/task/task.hpp:453-486): an edge whose
//     predecessor task has not completed adds 1 to the successor task's wait
//     counter and puts the successor on the predecessor access's list;
//   __task_commit (task.hpp:535-551): the counter starts at 1, a guard that
//     the commit drops once the task's accesses are linked -- here once the
//     launch that resolved the task has had its edges applied;
//   __task_complete (src/task.cc:160-247): the task turns COMPLETED, and every
//     successor on the lists of its accesses -- except detached (D mode)
//     accesses, whose task fulfils them itself -- loses 1; a counter reaching
//     0 releases its task.
//
// Data layout in HBM (one engine per task domain, capacities fixed at
// creation): per task a wait counter, a state word and its first access id;
// per access its task (bit 31: detached) and the head of its successor list.
// A list is a chain of segments, one per launch that linked successors to the
// access: {first node, node count, next segment}, the nodes (one successor
// task each, one per resolved edge) contiguous.  A completion walks a segment
// with a wave's lanes, so a list of tens of successors costs a few dependent
// loads rather than one per successor.  The edges of a launch are applied in
// three passes: count the successors per predecessor access (listing each
// such access once), give each listed access a segment, fill the segments.
// Released tasks go to a queue in pinned host memory, one slot per task,
// written with system-scope release stores: the host reads them as they land,
// with no stream synchronisation.
//
// The kernels are integer and latency bound (a few dependent loads per edge
// or successor); they run on the pool's stream after each launch, so edges
// never leave the device.

#include <hip/hip_runtime.h>

#include <errno.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <vector>

#include "release.h"

namespace xkrt_release {

namespace {

constexpr uint32_t EMPTY = 0xFFFFFFFFu;
constexpr uint32_t ST_WAIT = 0, ST_READY = 1, ST_DONE = 2;
constexpr uint32_t DETACHED = 1u << 31;
constexpr uint32_t VIRT = 0x80000000u;  // edge end: the k-th empty write a handle map inserted
constexpr uint8_t MODE_D = 8;  // ACCESS_MODE_D (include/xkrt/memory/access/mode.h:62)
enum : uint32_t { RERR_NODES = 1, RERR_STATE = 2, RERR_RANGE = 4 };

struct Dev {
    uint32_t *wc;        // per task: wait counter (starts at 1: the commit guard)
    uint32_t *state;     // per task: ST_*
    uint32_t *first;     // per task + 1: first access id (first[t + 1] ends task t)
    uint32_t *acc_task;  // per access: task | DETACHED
    uint32_t *head;      // per access: first segment of its successor list, EMPTY when none
    uint4 *seg;          // per segment: {first node, nodes, next segment, -}
    uint32_t *node;      // per node: a successor task
    uint32_t *deg;       // per access: successors linked to it by the launch being applied
    uint32_t *sbase;     // per access: first node of its segment in that launch
    uint32_t *fill;      // per access: nodes of that segment filled
    uint32_t *tl;        // the accesses given successors by that launch, each once
    uint32_t *ready;     // device address of the pinned release queue
    unsigned int *fin;   // device address of a pinned word: completions whose successors are released
    unsigned int *ctr;   // [0] nodes used, [1] queue slots used, [2] error bits, [3] completions,
                         // [4] empty writes whose guard was dropped, [5] completions processed,
                         // [6] segments used, [7] entries of tl
    uint32_t max_tasks, max_acc, max_nodes, max_virtual;
    // tasks [max_tasks, max_tasks + max_virtual) and accesses [max_acc, max_acc + max_virtual) are
    // the empty writes handle maps insert (an empty task with one access each)
};

// a program task to the ready queue
__device__ void release_task(const Dev &r, uint32_t t)
{
    r.state[t] = ST_READY;
    const uint32_t slot = atomicAdd(&r.ctr[1], 1u);
    if (slot < r.max_tasks)
        __hip_atomic_store(&r.ready[slot], t, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    else
        atomicOr(&r.ctr[2], RERR_RANGE);
}

// a released task: a program task goes to the queue; an inserted empty write has nothing to run
// and completes at once (its successors are program tasks: an empty write is only ever linked to
// reads, concurrent writes or program accesses, never to another empty write).  One lane walks
// the empty write's list.
__device__ void release(const Dev &r, uint32_t t)
{
    if (t < r.max_tasks) {
        release_task(r, t);
        return;
    }
    r.state[t] = ST_DONE;
    for (uint32_t g = r.head[r.max_acc + (t - r.max_tasks)]; g != EMPTY;) {
        const uint4 sg = r.seg[g];
        for (uint32_t j = 0; j < sg.y; ++j) {
            const uint32_t s = r.node[sg.x + j];
            if (s >= r.max_tasks) {
                atomicOr(&r.ctr[2], RERR_RANGE);
                continue;
            }
            if (atomicSub(&r.wc[s], 1u) == 1u) release_task(r, s);
        }
        g = sg.z;
    }
}

// an edge end: its access (list head index) and task
__device__ bool edge_end(const Dev &r, uint32_t id, uint32_t &acc, uint32_t &task)
{
    if (id & VIRT) {
        const uint32_t k = id & ~VIRT;
        if (k >= r.max_virtual) return false;
        acc = r.max_acc + k;
        task = r.max_tasks + k;
        return true;
    }
    if (id >= r.max_acc) return false;
    acc = id;
    task = r.acc_task[id] & ~DETACHED;
    return task < r.max_tasks;
}

// edges of the stripes before stripe s (each clamped to the stripe's room)
__device__ uint64_t stripe_prefix(const unsigned long long *scount, uint32_t s, uint64_t cap)
{
    uint64_t p = 0;
    for (uint32_t j = 0; j < s; ++j) p += min((uint64_t)scount[j], cap);
    return p;
}

// The edges of a launch: blockIdx.y 0 takes the pool's buffer (count on the device), y = 1 + s
// the handle map's stripe s.  f(edge, valid) for each wave-wide group of 64, grid-stride: every
// lane of a wave calls f the same number of times (a lane past the end with valid false), so f
// may use the wave's ballots.
struct EdgeSrc {
    const uint2 *edges;
    const unsigned long long *count;
    const uint2 *hedges;
    const unsigned long long *hcount;
    uint32_t hstripes;
    uint64_t hcap;
};

template <typename F>
__device__ void for_each_edge(const EdgeSrc &src, F f)
{
    const uint2 *e;
    uint64_t n;
    if (blockIdx.y == 0) {
        e = src.edges;
        n = *src.count;
    } else {
        const uint32_t st = blockIdx.y - 1;
        e = src.hedges + st * src.hcap;
        n = min((uint64_t)src.hcount[st], src.hcap);
    }
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;  // a multiple of 64
    for (uint64_t w0 = blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x & ~63u); w0 < n; w0 += stride) {
        const uint64_t i = w0 + (threadIdx.x & 63u);
        f(i < n ? e[i] : make_uint2(0, 0), i < n);
    }
}

// the ends of an edge that links (its predecessor task has not completed); false otherwise
__device__ bool linking(const Dev &r, uint2 e, uint32_t &pa, uint32_t &ts)
{
    uint32_t tp, sa;
    if (!edge_end(r, e.x, pa, tp) || !edge_end(r, e.y, sa, ts)) {
        atomicOr(&r.ctr[2], RERR_RANGE);
        return false;
    }
    return r.state[tp] != ST_DONE;  // a completed predecessor adds no edge
}

__device__ __forceinline__ uint32_t lane_below(uint64_t m)
{
    return (uint32_t)__popcll(m & ((1ull << (threadIdx.x & 63u)) - 1ull));
}

// __task_precedes, pass 1: the successor's wait counter, and the successors per predecessor
// access (the first edge of an access lists it; one list-counter update per wave).  Nodes for
// every edge must fit, or nothing of the launch is applied.
__global__ void k_apply_count(Dev r, EdgeSrc src)
{
    const uint64_t n = *src.count + (src.hcount ? stripe_prefix(src.hcount, src.hstripes, src.hcap) : 0);
    if (r.ctr[2] & RERR_NODES) return;
    if (r.ctr[0] + n > r.max_nodes) {
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) atomicOr(&r.ctr[2], RERR_NODES);
        return;
    }
    for_each_edge(src, [&](uint2 e, bool valid) {
        uint32_t pa = 0, ts;
        bool first = false;
        if (valid && linking(r, e, pa, ts)) {
            atomicAdd(&r.wc[ts], 1u);
            first = atomicAdd(&r.deg[pa], 1u) == 0u;
        }
        const uint64_t m = __ballot(first);
        if (!m) return;
        uint32_t b = 0;
        if ((threadIdx.x & 63u) == 0) b = atomicAdd(&r.ctr[7], (uint32_t)__popcll(m));
        b = __shfl(b, 0, 64);
        if (first) r.tl[b + lane_below(m)] = pa;
    });
}

// pass 2: each listed access gets a segment of deg nodes at the head of its list (the nodes and
// segments of a wave's accesses taken with one update of each counter)
__global__ void k_apply_link(Dev r)
{
    if (r.ctr[2] & RERR_NODES) return;
    const uint32_t m = r.ctr[7], lane = threadIdx.x & 63u;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t w0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); w0 < m; w0 += stride) {
        const uint32_t j = w0 + lane;
        const uint32_t pa = j < m ? r.tl[j] : 0u;
        const uint32_t d = j < m ? r.deg[pa] : 0u;
        uint32_t incl = d;  // inclusive prefix of the wave's node counts
        for (uint32_t o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o, 64);
            if (lane >= o) incl += v;
        }
        const uint32_t tot = __shfl(incl, 63, 64), cnt = min(64u, m - w0);
        uint32_t nb = 0, gb = 0;
        if (lane == 0) {
            nb = atomicAdd(&r.ctr[0], tot);
            gb = atomicAdd(&r.ctr[6], cnt);  // segments <= nodes: within max_nodes
        }
        nb = __shfl(nb, 0, 64);
        gb = __shfl(gb, 0, 64);
        if (j < m) {
            const uint32_t base = nb + incl - d, g = gb + lane;
            r.seg[g] = make_uint4(base, d, r.head[pa], 0u);
            r.head[pa] = g;
            r.sbase[pa] = base;
            r.fill[pa] = 0;
            r.deg[pa] = 0;  // ready for the next launch
        }
    }
}

// pass 3: the successors into their segments (the order within a segment is free)
__global__ void k_apply_fill(Dev r, EdgeSrc src)
{
    if (r.ctr[2] & RERR_NODES) return;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) r.ctr[7] = 0;  // read by pass 2 only
    for_each_edge(src, [&](uint2 e, bool valid) {
        uint32_t pa, ts;
        if (!valid || !linking(r, e, pa, ts)) return;
        r.node[r.sbase[pa] + atomicAdd(&r.fill[pa], 1u)] = ts;
    });
}

__global__ void k_fill(uint32_t *p, uint64_t n, uint32_t v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// the commit guard of the launch's tasks
__global__ void k_guard(Dev r, uint32_t t0, uint32_t t1)
{
    // edges not applied (no room): the launch's tasks keep their guard -- none runs before its
    // predecessors; the domain reports -ENOSPC
    if (r.ctr[2] & RERR_NODES) return;
    const uint32_t t = t0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= t1) return;
    if (atomicSub(&r.wc[t], 1u) == 1u) release(r, t);
}

// the guards of the empty writes inserted since the last call (their count on the device)
__global__ void k_guard_virtual(Dev r, const unsigned int *nvirtual)
{
    if (r.ctr[2] & RERR_NODES) return;
    const uint32_t v0 = r.ctr[4], v1 = min(*nvirtual, r.max_virtual);
    for (uint32_t k = v0 + threadIdx.x; k < v1; k += blockDim.x)
        if (atomicSub(&r.wc[r.max_tasks + k], 1u) == 1u) release(r, r.max_tasks + k);
    __syncthreads();
    if (threadIdx.x == 0 && v1 > v0) r.ctr[4] = v1;
}

// one wave per completed task: the lanes walk each segment of its accesses' successor lists.
// Then the completions are counted as processed, block by block, and the block that brings the
// count to `target` (the completions enqueued up to this launch) publishes it in pinned host
// memory -- after every block's releases are visible there -- for a host that waits for them.
constexpr uint32_t CWAVES = 4;  // k_complete: waves (tasks) per block

__global__ void __launch_bounds__(64 * CWAVES) k_complete(Dev r, const uint32_t *tasks, uint32_t n, uint32_t target)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t i = blockIdx.x * CWAVES + (threadIdx.x >> 6);
    if (i < n) {
        const uint32_t t = tasks[i];
        uint32_t ok = 0;
        if (lane == 0) {
            ok = t < r.max_tasks && atomicCAS(&r.state[t], ST_READY, ST_DONE) == ST_READY;
            if (ok) atomicAdd(&r.ctr[3], 1u);
            else atomicOr(&r.ctr[2], RERR_STATE);  // not released, or completed twice
        }
        if (__shfl(ok, 0, 64)) {
            for (uint32_t a = r.first[t]; a < r.first[t + 1]; ++a) {
                if (r.acc_task[a] & DETACHED) continue;
                for (uint32_t g = r.head[a]; g != EMPTY;) {
                    const uint4 sg = r.seg[g];
                    for (uint32_t j = lane; j < sg.y; j += 64) {
                        const uint32_t s = r.node[sg.x + j];
                        if (atomicSub(&r.wc[s], 1u) == 1u) release(r, s);
                    }
                    g = sg.z;
                }
            }
        }
    }
    __threadfence_system();  // this thread's queue stores reach host memory first
    __syncthreads();
    if (threadIdx.x == 0 && blockIdx.x * CWAVES < n) {
        const uint32_t mine = min(n - blockIdx.x * CWAVES, CWAVES);
        const uint32_t old = __hip_atomic_fetch_add(&r.ctr[5], mine, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old + mine == target) __hip_atomic_store(r.fin, target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

int hip_rc(hipError_t e) { return e == hipSuccess ? 0 : (e == hipErrorOutOfMemory ? -ENOMEM : -EIO); }

#define RCHK(x)                        \
    do {                               \
        const int rc_ = hip_rc(x);     \
        if (rc_) return rc_;           \
    } while (0)

} // namespace

struct Engine {
    hipStream_t stream = nullptr;
    Dev d{};
    uint32_t *h_acc_task = nullptr;  // pinned mirrors: append-only sources of the uploads
    uint32_t *h_first = nullptr;
    uint32_t *h_done = nullptr;      // completion lists, append-only
    uint32_t *d_done = nullptr;      // their device address
    uint32_t *h_ready = nullptr;     // the release queue (host address)
    unsigned int *h_fin = nullptr;   // completions processed (written by k_complete)
    unsigned int *h_ctr = nullptr;
    uint64_t ntasks = 0, nacc = 0, done_used = 0, read = 0;
};

int create(void *stream, uint64_t max_tasks, uint64_t max_acc, uint64_t max_nodes, uint64_t max_virtual,
           Engine **out)
{
    if (!out || !max_tasks || !max_acc || !max_nodes) return -EINVAL;
    if (max_tasks + max_virtual >= DETACHED || max_acc + max_virtual >= VIRT || max_nodes >= EMPTY) return -ERANGE;
    *out = nullptr;
    Engine *e = new Engine();
    e->stream = (hipStream_t)stream;
    Dev &d = e->d;
    d.max_tasks = (uint32_t)max_tasks;
    d.max_acc = (uint32_t)max_acc;
    d.max_nodes = (uint32_t)max_nodes;
    d.max_virtual = (uint32_t)max_virtual;
    const uint64_t nt = max_tasks + max_virtual, na = max_acc + max_virtual;
    auto fail = [&](int rc) {
        destroy(e);
        return rc;
    };
    int rc = 0;
    if ((rc = hip_rc(hipMalloc(&d.wc, 4 * nt))) || (rc = hip_rc(hipMalloc(&d.state, 4 * nt))) ||
        (rc = hip_rc(hipMalloc(&d.first, 4 * (max_tasks + 1)))) ||
        (rc = hip_rc(hipMalloc(&d.acc_task, 4 * max_acc))) || (rc = hip_rc(hipMalloc(&d.head, 4 * na))) ||
        (rc = hip_rc(hipMalloc(&d.seg, 16 * max_nodes))) || (rc = hip_rc(hipMalloc(&d.node, 4 * max_nodes))) ||
        (rc = hip_rc(hipMalloc(&d.deg, 4 * na))) || (rc = hip_rc(hipMalloc(&d.sbase, 4 * na))) ||
        (rc = hip_rc(hipMalloc(&d.fill, 4 * na))) || (rc = hip_rc(hipMalloc(&d.tl, 4 * na))) ||
        (rc = hip_rc(hipMalloc(&d.ctr, 64))))
        return fail(rc);
    if ((rc = hip_rc(hipHostMalloc(&e->h_acc_task, 4 * max_acc, hipHostMallocDefault))) ||
        (rc = hip_rc(hipHostMalloc(&e->h_first, 4 * (max_tasks + 1), hipHostMallocDefault))) ||
        (rc = hip_rc(hipHostMalloc(&e->h_done, 4 * max_tasks, hipHostMallocMapped))) ||
        (rc = hip_rc(hipHostMalloc(&e->h_ready, 4 * max_tasks, hipHostMallocCoherent | hipHostMallocMapped))) ||
        (rc = hip_rc(hipHostMalloc(&e->h_fin, 64, hipHostMallocCoherent | hipHostMallocMapped))) ||
        (rc = hip_rc(hipHostMalloc(&e->h_ctr, 64, hipHostMallocDefault))))
        return fail(rc);
    memset(e->h_ready, 0xFF, 4 * max_tasks);
    memset(e->h_fin, 0, 64);
    if ((rc = hip_rc(hipHostGetDevicePointer((void **)&d.ready, e->h_ready, 0)))) return fail(rc);
    if ((rc = hip_rc(hipHostGetDevicePointer((void **)&d.fin, e->h_fin, 0)))) return fail(rc);
    if ((rc = hip_rc(hipHostGetDevicePointer((void **)&e->d_done, e->h_done, 0)))) return fail(rc);
    const dim3 g(1024), b(256);
    hipLaunchKernelGGL(k_fill, g, b, 0, e->stream, d.wc, nt, 1u);
    hipLaunchKernelGGL(k_fill, g, b, 0, e->stream, d.state, nt, ST_WAIT);
    hipLaunchKernelGGL(k_fill, g, b, 0, e->stream, d.head, na, EMPTY);
    hipLaunchKernelGGL(k_fill, g, b, 0, e->stream, d.deg, na, 0u);
    hipLaunchKernelGGL(k_fill, g, b, 0, e->stream, (uint32_t *)d.ctr, 16, 0u);
    if ((rc = hip_rc(hipGetLastError())) || (rc = hip_rc(hipStreamSynchronize(e->stream)))) return fail(rc);
    *out = e;
    return 0;
}

void destroy(Engine *e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    Dev &d = e->d;
    for (void *p : {(void *)d.wc, (void *)d.state, (void *)d.first, (void *)d.acc_task, (void *)d.head, (void *)d.seg,
                    (void *)d.node, (void *)d.deg, (void *)d.sbase, (void *)d.fill, (void *)d.tl, (void *)d.ctr})
        if (p) (void)hipFree(p);
    for (void *p : {(void *)e->h_acc_task, (void *)e->h_first, (void *)e->h_done, (void *)e->h_ready, (void *)e->h_fin,
                    (void *)e->h_ctr})
        if (p) (void)hipHostFree(p);
    delete e;
}

int add(Engine *e, uint64_t acc0, const uint32_t *task, const uint8_t *mode, uint64_t n)
{
    if (!e || (!n) || !task || !mode) return n ? -EINVAL : 0;
    if (acc0 != e->nacc) return -EINVAL;
    if (task[0] != e->ntasks) return -EINVAL;  // tasks continue the previous batch
    for (uint64_t i = 1; i < n; ++i)
        if (task[i] != task[i - 1] && task[i] != task[i - 1] + 1) return -EINVAL;
    const uint64_t t_end = (uint64_t)task[n - 1] + 1;
    if (t_end > e->d.max_tasks || acc0 + n > e->d.max_acc) return -ENOSPC;
    for (uint64_t i = 0; i < n; ++i) {
        e->h_acc_task[acc0 + i] = task[i] | ((mode[i] & MODE_D) ? DETACHED : 0u);
        if (i == 0 || task[i] != task[i - 1]) e->h_first[task[i]] = (uint32_t)(acc0 + i);
    }
    e->h_first[t_end] = (uint32_t)(acc0 + n);
    const uint64_t t0 = e->ntasks;
    RCHK(hipMemcpyAsync(e->d.acc_task + acc0, e->h_acc_task + acc0, 4 * n, hipMemcpyHostToDevice, e->stream));
    RCHK(hipMemcpyAsync(e->d.first + t0, e->h_first + t0, 4 * (t_end + 1 - t0), hipMemcpyHostToDevice, e->stream));
    e->ntasks = t_end;
    e->nacc = acc0 + n;
    return 0;
}

int apply(Engine *e, const uint32_t *edges, const uint64_t *count, const HandleEdges *h, uint32_t t0, uint32_t t1)
{
    if (!e || !edges || !count || t0 > t1 || t1 > e->ntasks) return -EINVAL;
    EdgeSrc src{(const uint2 *)edges, (const unsigned long long *)count, nullptr, nullptr, 0, 0};
    if (h) {
        src.hedges = (const uint2 *)h->edges;
        src.hcount = (const unsigned long long *)h->count;
        src.hstripes = h->stripes;
        src.hcap = h->stripe_cap;
    }
    const dim3 eg(h ? 128 : 512, 1 + (h ? h->stripes : 0));
    hipLaunchKernelGGL(k_apply_count, eg, dim3(256), 0, e->stream, e->d, src);
    hipLaunchKernelGGL(k_apply_link, dim3(256), dim3(256), 0, e->stream, e->d);
    hipLaunchKernelGGL(k_apply_fill, eg, dim3(256), 0, e->stream, e->d, src);
    if (h)
        hipLaunchKernelGGL(k_guard_virtual, dim3(1), dim3(256), 0, e->stream, e->d,
                           (const unsigned int *)h->nvirtual);
    hipLaunchKernelGGL(k_guard, dim3(t1 > t0 ? (t1 - t0 + 255) / 256 : 1), dim3(256), 0, e->stream, e->d, t0, t1);
    return hip_rc(hipGetLastError());
}

int complete(Engine *e, const uint32_t *tasks, uint64_t n)
{
    if (!e || (n && !tasks)) return -EINVAL;
    if (n == 0) return 0;
    if (e->done_used + n > e->d.max_tasks) return -ENOSPC;  // each task completes once
    uint32_t *src = e->h_done + e->done_used;
    memcpy(src, tasks, 4 * n);
    // the kernel reads the list in place from pinned memory (never rewritten)
    uint32_t *dsrc = e->d_done + e->done_used;
    e->done_used += n;
    hipLaunchKernelGGL(k_complete, dim3((uint32_t)((n + CWAVES - 1) / CWAVES)), dim3(64 * CWAVES), 0, e->stream, e->d,
                       dsrc, (uint32_t)n, (uint32_t)e->done_used);
    return hip_rc(hipGetLastError());
}

int complete_wait(Engine *e, const uint32_t *tasks, uint64_t n)
{
    const int rc = complete(e, tasks, n);
    if (rc) return rc;
    // every completion enqueued so far has released its successors once k_complete says so; the
    // wait polls pinned memory (no stream synchronisation) and is bounded
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t spin = 0; __atomic_load_n(e->h_fin, __ATOMIC_ACQUIRE) < e->done_used; ++spin) {
        if ((spin & 4095u) != 4095u) continue;
        const auto waited = std::chrono::steady_clock::now() - t0;
        if (waited > std::chrono::seconds(10)) return -ETIMEDOUT;
        if (waited > std::chrono::milliseconds(1)) {  // a long wait: has the stream failed?
            const hipError_t q = hipStreamQuery(e->stream);
            if (q != hipSuccess && q != hipErrorNotReady) return hip_rc(q);
        }
    }
    return 0;
}

int ready(Engine *e, uint32_t *out, uint64_t cap, uint64_t *n)
{
    if (!e || !n || (cap && !out)) return -EINVAL;
    uint64_t k = 0;
    while (k < cap && e->read < e->d.max_tasks) {
        const uint32_t t = __atomic_load_n(e->h_ready + e->read, __ATOMIC_ACQUIRE);
        if (t == EMPTY) break;
        out[k++] = t;
        e->read++;
    }
    *n = k;
    return 0;
}

int check(Engine *e)
{
    if (!e) return -EINVAL;
    RCHK(hipMemcpyAsync(e->h_ctr, e->d.ctr, 64, hipMemcpyDeviceToHost, e->stream));
    RCHK(hipStreamSynchronize(e->stream));
    const uint32_t err = e->h_ctr[2];
    if (err & RERR_STATE) return -EINVAL;
    if (err & (RERR_NODES | RERR_RANGE)) return -ENOSPC;
    return 0;
}

int counts(Engine *e, Counts *out)
{
    if (!e || !out) return -EINVAL;
    const int rc = check(e);
    out->tasks = e->ntasks;
    out->accesses = e->nacc;
    out->nodes = e->h_ctr[0];
    out->released = e->h_ctr[1];
    out->read = e->read;
    out->completed = e->h_ctr[3];
    return rc;
}

} // namespace xkrt_release
