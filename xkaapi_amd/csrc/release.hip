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
// per access its task (bit 31: detached) and the head of its successor list;
// per resolved edge one list node {successor task, next node}.  Released
// tasks go to a queue in pinned host memory, one slot per task, written with
// system-scope release stores: the host reads them as they land, with no
// stream synchronisation.
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
    uint32_t *head;      // per access: first successor node, EMPTY when none
    uint2 *node;         // per node: {successor task, next node}
    uint32_t *ready;     // device address of the pinned release queue
    unsigned int *ctr;   // [0] nodes used, [1] queue slots used, [2] error bits, [3] completions,
                         // [4] empty writes whose guard was dropped
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
// reads, concurrent writes or program accesses, never to another empty write)
__device__ void release(const Dev &r, uint32_t t)
{
    if (t < r.max_tasks) {
        release_task(r, t);
        return;
    }
    r.state[t] = ST_DONE;
    for (uint32_t k = r.head[r.max_acc + (t - r.max_tasks)]; k != EMPTY; k = r.node[k].y) {
        const uint32_t s = r.node[k].x;
        if (s >= r.max_tasks) {
            atomicOr(&r.ctr[2], RERR_RANGE);
            continue;
        }
        if (atomicSub(&r.wc[s], 1u) == 1u) release_task(r, s);
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

// __task_precedes for one edge, its list node given
__device__ void apply_edge(const Dev &r, uint2 e, uint32_t node)
{
    uint32_t pa, tp, sa, ts;
    if (!edge_end(r, e.x, pa, tp) || !edge_end(r, e.y, sa, ts)) {
        atomicOr(&r.ctr[2], RERR_RANGE);
        return;
    }
    if (r.state[tp] == ST_DONE) return;  // a completed predecessor adds no edge
    atomicAdd(&r.wc[ts], 1u);
    r.node[node].x = ts;
    r.node[node].y = atomicExch(&r.head[pa], node);
}

// edges of the stripes before stripe s (each clamped to the stripe's room)
__device__ uint64_t stripe_prefix(const unsigned long long *scount, uint32_t s, uint64_t cap)
{
    uint64_t p = 0;
    for (uint32_t j = 0; j < s; ++j) p += min((uint64_t)scount[j], cap);
    return p;
}

__global__ void k_fill(uint32_t *p, uint64_t n, uint32_t v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// one thread per unconsumed edge (grid-stride: the count is read on the device).  Edge i takes
// list node base + i (base = nodes used before the launch; k_guard advances it), so no thread
// contends for a shared node counter; an edge from a completed task leaves its node unused.
__global__ void k_apply(Dev r, const uint2 *edges, const unsigned long long *count,
                        const unsigned long long *hcount, uint32_t hstripes, uint64_t hcap)
{
    const uint64_t n = *count;
    const uint64_t base = r.ctr[0];
    const uint64_t nh = hcount ? stripe_prefix(hcount, hstripes, hcap) : 0;
    if (base + n + nh > r.max_nodes) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&r.ctr[2], RERR_NODES);
        return;
    }
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        apply_edge(r, edges[i], (uint32_t)(base + i));
}

// the handle map's edges of the launch, in stripes (stripe s = blockIdx.y): they take the nodes
// after the pool's edges
__global__ void k_apply_striped(Dev r, const unsigned long long *count, const uint2 *hedges,
                                const unsigned long long *hcount, uint32_t hstripes, uint64_t hcap)
{
    const uint64_t n = *count;
    const uint64_t base = r.ctr[0];
    if (base + n + stripe_prefix(hcount, hstripes, hcap) > r.max_nodes) return;  // k_apply flagged it
    const uint32_t st = blockIdx.y;
    const uint64_t m = min((uint64_t)hcount[st], hcap);
    const uint64_t off = base + n + stripe_prefix(hcount, st, hcap);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
        apply_edge(r, hedges[st * hcap + i], (uint32_t)(off + i));
}

// the commit guard of the launch's tasks; the first thread moves the node base past the
// launch's edges (the apply kernels have finished: stream order)
__global__ void k_guard(Dev r, uint32_t t0, uint32_t t1, const unsigned long long *count,
                        const unsigned long long *hcount, uint32_t hstripes, uint64_t hcap)
{
    // edges not applied (no room): the launch's tasks keep their guard -- none runs before its
    // predecessors; the domain reports -ENOSPC
    if (r.ctr[2] & RERR_NODES) return;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        r.ctr[0] += (unsigned int)(*count + (hcount ? stripe_prefix(hcount, hstripes, hcap) : 0));
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

// one thread per completed task: its accesses' successor lists
__global__ void k_complete(Dev r, const uint32_t *tasks, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = tasks[i];
    if (t >= r.max_tasks || atomicCAS(&r.state[t], ST_READY, ST_DONE) != ST_READY) {
        atomicOr(&r.ctr[2], RERR_STATE);  // not released, or completed twice
        return;
    }
    atomicAdd(&r.ctr[3], 1u);
    for (uint32_t a = r.first[t]; a < r.first[t + 1]; ++a) {
        if (r.acc_task[a] & DETACHED) continue;
        for (uint32_t k = r.head[a]; k != EMPTY; k = r.node[k].y) {
            const uint32_t s = r.node[k].x;
            if (atomicSub(&r.wc[s], 1u) == 1u) release(r, s);
        }
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
        (rc = hip_rc(hipMalloc(&d.node, 8 * max_nodes))) || (rc = hip_rc(hipMalloc(&d.ctr, 64))))
        return fail(rc);
    if ((rc = hip_rc(hipHostMalloc(&e->h_acc_task, 4 * max_acc, hipHostMallocDefault))) ||
        (rc = hip_rc(hipHostMalloc(&e->h_first, 4 * (max_tasks + 1), hipHostMallocDefault))) ||
        (rc = hip_rc(hipHostMalloc(&e->h_done, 4 * max_tasks, hipHostMallocMapped))) ||
        (rc = hip_rc(hipHostMalloc(&e->h_ready, 4 * max_tasks, hipHostMallocCoherent | hipHostMallocMapped))) ||
        (rc = hip_rc(hipHostMalloc(&e->h_ctr, 64, hipHostMallocDefault))))
        return fail(rc);
    memset(e->h_ready, 0xFF, 4 * max_tasks);
    if ((rc = hip_rc(hipHostGetDevicePointer((void **)&d.ready, e->h_ready, 0)))) return fail(rc);
    if ((rc = hip_rc(hipHostGetDevicePointer((void **)&e->d_done, e->h_done, 0)))) return fail(rc);
    const dim3 g(1024), b(256);
    hipLaunchKernelGGL(k_fill, g, b, 0, e->stream, d.wc, nt, 1u);
    hipLaunchKernelGGL(k_fill, g, b, 0, e->stream, d.state, nt, ST_WAIT);
    hipLaunchKernelGGL(k_fill, g, b, 0, e->stream, d.head, na, EMPTY);
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
    for (void *p : {(void *)d.wc, (void *)d.state, (void *)d.first, (void *)d.acc_task, (void *)d.head, (void *)d.node,
                    (void *)d.ctr})
        if (p) (void)hipFree(p);
    for (void *p : {(void *)e->h_acc_task, (void *)e->h_first, (void *)e->h_done, (void *)e->h_ready, (void *)e->h_ctr})
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
    const unsigned long long *hc = h ? (const unsigned long long *)h->count : nullptr;
    const uint32_t hs = h ? h->stripes : 0;
    const uint64_t hcap = h ? h->stripe_cap : 0;
    const unsigned long long *cnt = (const unsigned long long *)count;
    hipLaunchKernelGGL(k_apply, dim3(512), dim3(256), 0, e->stream, e->d, (const uint2 *)edges, cnt, hc, hs, hcap);
    if (h) {
        hipLaunchKernelGGL(k_apply_striped, dim3(8, hs), dim3(256), 0, e->stream, e->d, cnt,
                           (const uint2 *)h->edges, hc, hs, hcap);
        hipLaunchKernelGGL(k_guard_virtual, dim3(1), dim3(256), 0, e->stream, e->d,
                           (const unsigned int *)h->nvirtual);
    }
    hipLaunchKernelGGL(k_guard, dim3(t1 > t0 ? (t1 - t0 + 255) / 256 : 1), dim3(256), 0, e->stream, e->d, t0, t1,
                       cnt, hc, hs, hcap);
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
    hipLaunchKernelGGL(k_complete, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, e->stream, e->d, dsrc, (uint32_t)n);
    return hip_rc(hipGetLastError());
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
