// MI355X handle dependency domain: the reference's DependencyMap
// (include/xkrt/memory/access/handle/dependency-map.hpp:52-330) -- OpenMP
// style dependencies on opaque handles, with concurrent writes
// (inoutset) -- resolved on the device.
//
// Data layout in HBM, per domain:
//   * an open-addressing table of handles (u64 keys, linear probing), one
//     node per slot: the last sequential write, the heads of the sequential
//     reads and concurrent writes lists, and whether an access was put there
//     (a handle nothing was put on yet links to nothing);
//   * list entries {access, next}, indexed by access id: an access enters at
//     most one list once (an inserted empty write none), so entry a is
//     access a's and no counter is shared;
//   * per access (program and empty writes) the task of its last successor,
//     for __access_precedes' dedup (include/xkrt/task/task.hpp:494-512);
//   * the edge buffer, (pred id, succ id) pairs; an empty write inserted by
//     the map (insert_empty_write, dependency-map.hpp:94-135) has id
//     0x80000000 | k and is a task of its own.
//
// One submit = one launch of three steps:
//   k_h_slot   each record finds (or claims) its handle's slot; sort key =
//              slot << 32 | record index
//   sort       hipCUB radix sort of the keys: each handle's records in spawn
//              order, contiguous
//   k_h_replay one thread per handle present (the first record of its run):
//              the node state in registers, the handle's records replayed --
//              LINK of every access of a task, then PUT (src/task/task.cc:
//              285-299) -- and written back; its edges go to one of 64
//              stripes of the edge buffer.
// Handles are independent, so the work is one chain per handle; a handle's
// own chain is sequential, as it is in the reference.  The kernels are
// latency bound (a handful of dependent loads per record).

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <errno.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>

#include "../../include/xkrt_hip.h"

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint32_t VIRT = 0x80000000u;
constexpr unsigned long long EMPTY_KEY = ~0ull;
constexpr uint32_t MODE_R = 1, MODE_W = 2, MODE_VW = 6;
constexpr uint32_t CNC = 4;
enum : uint32_t { HERR_TABLE = 1, HERR_ENTRIES = 2, HERR_VIRTUAL = 4, HERR_EDGES = 8 };
// the edge buffer is cut into stripes, a handle's edges going to stripe (slot mod ESTRIPES): the
// runs of a launch append concurrently, and one shared count serialised them
constexpr uint32_t ESTRIPES = 64;

struct HRec {
    unsigned long long handle;
    uint32_t acc, task;
    uint32_t mc;  // mode | concurrency << 8
    uint32_t pad;
};

struct Node {
    uint32_t lsw, rhead, chead, exists;
};

struct HDev {
    unsigned long long *keys;  // table
    Node *node;
    uint2 *entry;              // {access, next}
    uint32_t *last_prog;       // per program access: task of its last successor
    uint32_t *last_virt;       // per empty write
    uint2 *edges;
    unsigned int *ctr;         // [1] empty writes, [2] errors
    unsigned long long *scount;  // edges per stripe
    uint32_t mask, max_entries, max_virtual, max_prog;
    unsigned long long stripe_cap;  // edges per stripe: max_edges / ESTRIPES
};

__device__ uint32_t hash64(unsigned long long h)
{
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return (uint32_t)h;
}

__global__ void k_h_slot(HDev d, const HRec *rec, uint32_t n, unsigned long long *keys_out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long h = rec[i].handle;
    uint32_t s = hash64(h) & d.mask;
    for (uint32_t probe = 0; probe <= d.mask; ++probe, s = (s + 1) & d.mask) {
        const unsigned long long k = d.keys[s];
        if (k == h) break;
        if (k == EMPTY_KEY) {
            const unsigned long long prev = atomicCAS(&d.keys[s], EMPTY_KEY, h);
            if (prev == EMPTY_KEY || prev == h) break;
        }
        if (probe == d.mask) {
            atomicOr(&d.ctr[2], HERR_TABLE);
            s = NONE;
            break;
        }
    }
    keys_out[i] = s == NONE ? ~0ull : ((unsigned long long)s << 32 | i);
}

struct Replay {
    const HDev &d;
    Node st;
    uint32_t emitted;
    uint32_t stripe;  // the edge stripe this handle appends to

    __device__ uint32_t &last(uint32_t id) { return id & VIRT ? d.last_virt[id & ~VIRT] : d.last_prog[id]; }

    // __access_precedes: skipped when the predecessor's last successor is of the same task
    __device__ void precedes(uint32_t p, uint32_t s, uint32_t stask)
    {
        uint32_t &l = last(p);
        if (l == stask) return;
        l = stask;
        emitted++;
        const unsigned long long k = atomicAdd(&d.scount[stripe], 1ull);
        if (k < d.stripe_cap)
            d.edges[stripe * d.stripe_cap + k] = make_uint2(p, s);
        else
            atomicOr(&d.ctr[2], HERR_EDGES);
    }

    __device__ void link_list(uint32_t head, uint32_t s, uint32_t stask)
    {
        for (uint32_t e = head; e != NONE; e = d.entry[e].y) precedes(d.entry[e].x, s, stask);
    }

    // a program access enters at most one list, once: its entry is its own id (no shared counter)
    __device__ void push(uint32_t &head, uint32_t a)
    {
        if (a >= d.max_entries) {
            atomicOr(&d.ctr[2], HERR_ENTRIES);
            return;
        }
        d.entry[a] = make_uint2(a, head);
        head = a;
    }

    __device__ void put(uint32_t a, uint32_t mode, uint32_t conc)
    {
        st.exists = 1;
        if (mode & MODE_W) {
            if (conc == CNC) {
                push(st.chead, a);
            } else {
                st.rhead = NONE;
                st.chead = NONE;
                st.lsw = a;
            }
        } else if (mode & MODE_R) {
            st.chead = NONE;
            push(st.rhead, a);
        }
    }

    // an empty sequential write (mode VW), a task of its own: linked -- as any sequential write:
    // every read, every concurrent write, else the last sequential write -- then put when it got
    // a predecessor (its commit guard, dependency-map.hpp:118-131)
    __device__ void empty_write()
    {
        const uint32_t k = atomicAdd(&d.ctr[1], 1u);
        if (k >= d.max_virtual) {
            atomicOr(&d.ctr[2], HERR_VIRTUAL);
            return;
        }
        const uint32_t x = VIRT | k;
        const uint32_t e0 = emitted;
        bool trans = false;
        if (st.rhead != NONE) {
            link_list(st.rhead, x, x);
            trans = true;
        }
        if (st.chead != NONE) {
            link_list(st.chead, x, x);
            trans = true;
        }
        if (st.lsw != NONE && !trans) precedes(st.lsw, x, x);
        if (emitted != e0) put(x, MODE_VW, 1);
    }

    __device__ void link(uint32_t a, uint32_t task, uint32_t mode, uint32_t conc)
    {
        if (!st.exists) return;
        bool trans = false;
        if (st.rhead != NONE && (mode & MODE_W)) {
            if (conc == CNC) {
                empty_write();
            } else {
                link_list(st.rhead, a, task);
                trans = true;
            }
        }
        if (st.chead != NONE && conc != CNC) {
            if (mode & MODE_W) {
                link_list(st.chead, a, task);
                trans = true;
            } else {
                empty_write();
            }
        }
        if (st.lsw != NONE && !trans) precedes(st.lsw, a, task);
    }
};

// one thread per sorted record; the first record of each handle's run replays the run.  The
// runs are chains of dependent accesses, so a launch lasts as long as its longest run: all of
// them start together.
__global__ void k_h_replay(HDev d, const HRec *rec, const unsigned long long *keys, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long k0 = keys[i];
    const uint32_t slot = (uint32_t)(k0 >> 32);
    if (k0 == ~0ull) return;
    if (i > 0 && (uint32_t)(keys[i - 1] >> 32) == slot) return;
    Replay r{d, d.node[slot], 0, slot & (ESTRIPES - 1)};
    uint32_t j = i;
    while (j < n && (uint32_t)(keys[j] >> 32) == slot && keys[j] != ~0ull) {
        const uint32_t task = rec[(uint32_t)keys[j]].task;
        uint32_t e = j;
        while (e < n && (uint32_t)(keys[e] >> 32) == slot && keys[e] != ~0ull && rec[(uint32_t)keys[e]].task == task)
            ++e;
        for (uint32_t q = j; q < e; ++q) {
            const HRec &x = rec[(uint32_t)keys[q]];
            r.link(x.acc, x.task, x.mc & 0xFF, x.mc >> 8);
        }
        for (uint32_t q = j; q < e; ++q) {
            const HRec &x = rec[(uint32_t)keys[q]];
            r.put(x.acc, x.mc & 0xFF, x.mc >> 8);
        }
        j = e;
    }
    d.node[slot] = r.st;
}

__global__ void k_h_fill_u32(uint32_t *p, uint64_t n, uint32_t v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_h_init_nodes(Node *node, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        node[i] = Node{NONE, NONE, NONE, 0u};
}

__global__ void k_h_fill_u64(unsigned long long *p, uint64_t n, unsigned long long v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

int hrc(hipError_t e) { return e == hipSuccess ? 0 : (e == hipErrorOutOfMemory ? -ENOMEM : -EIO); }

} // namespace

struct xkrt_hip_handle_domain {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    HDev d{};
    xkrt_hip_handle_config_t cfg{};
    HRec *rec = nullptr, *rec_host = nullptr;  // launch records: device, pinned staging
    unsigned long long *keys = nullptr, *keys_sorted = nullptr;
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int sort_bits = 0;
    hipEvent_t staged = nullptr;
    bool staged_pending = false;
    unsigned int *ctr_host = nullptr;  // pinned copy of the counters
    unsigned long long *scount_host = nullptr;
    uint64_t next_id = 0, launches = 0;
    uint32_t last_task = 0;
    bool any_task = false;
    bool any_id = false;  // caller ids given: the last one
    uint32_t last_id = 0;
};

extern "C" {

int xkrt_hip_handle_domain_create(const xkrt_hip_handle_config_t *cfg, void *stream, xkrt_hip_handle_domain_t *out)
{
    if (!out) return -EINVAL;
    *out = nullptr;
    xkrt_hip_handle_config_t c{};
    if (cfg) c = *cfg;
    if (!c.max_handles) c.max_handles = 1u << 16;
    if (!c.max_accesses) c.max_accesses = 1u << 22;
    if (!c.max_virtual) c.max_virtual = 1u << 20;
    if (!c.max_batch) c.max_batch = 1u << 18;
    if (!c.max_edges) c.max_edges = 1u << 22;
    if (c.max_accesses >= VIRT || c.max_virtual >= VIRT || c.max_handles > (1u << 30) || c.max_batch >= VIRT)
        return -ERANGE;
    xkrt_hip_handle_domain *h = new xkrt_hip_handle_domain();
    h->cfg = c;
    auto fail = [&](int rc) {
        xkrt_hip_handle_domain_destroy(h);
        return rc;
    };
    int rc = 0;
    if (stream) {
        h->stream = (hipStream_t)stream;
    } else {
        if ((rc = hrc(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)))) return fail(rc);
        h->own_stream = true;
    }
    uint32_t cap = 1;
    while (cap < 2 * c.max_handles) cap <<= 1;  // load factor at most 1/2
    HDev &d = h->d;
    d.mask = cap - 1;
    d.max_entries = (uint32_t)c.max_accesses;
    d.max_virtual = (uint32_t)c.max_virtual;
    d.max_prog = (uint32_t)c.max_accesses;
    d.stripe_cap = c.max_edges / ESTRIPES;
    int bits = 1;
    while ((1ull << bits) < cap) ++bits;
    h->sort_bits = 32 + bits;
    if ((rc = hrc(hipMalloc(&d.keys, 8ull * cap))) || (rc = hrc(hipMalloc(&d.node, sizeof(Node) * cap))) ||
        (rc = hrc(hipMalloc(&d.entry, 8ull * c.max_accesses))) ||
        (rc = hrc(hipMalloc(&d.last_prog, 4ull * c.max_accesses))) ||
        (rc = hrc(hipMalloc(&d.last_virt, 4ull * c.max_virtual))) ||
        (rc = hrc(hipMalloc(&d.edges, 8ull * c.max_edges))) || (rc = hrc(hipMalloc(&d.ctr, 64))) ||
        (rc = hrc(hipMalloc(&d.scount, 8ull * ESTRIPES))) ||
        (rc = hrc(hipHostMalloc(&h->scount_host, 8ull * ESTRIPES, hipHostMallocDefault))) ||
        (rc = hrc(hipMalloc(&h->rec, sizeof(HRec) * c.max_batch))) ||
        (rc = hrc(hipMalloc(&h->keys, 8ull * c.max_batch))) || (rc = hrc(hipMalloc(&h->keys_sorted, 8ull * c.max_batch))) ||
        (rc = hrc(hipHostMalloc(&h->rec_host, sizeof(HRec) * c.max_batch, hipHostMallocDefault))) ||
        (rc = hrc(hipHostMalloc(&h->ctr_host, 64, hipHostMallocDefault))) || (rc = hrc(hipEventCreate(&h->staged))))
        return fail(rc);
    if ((rc = hrc(hipcub::DeviceRadixSort::SortKeys(nullptr, h->sort_tmp_bytes, h->keys, h->keys_sorted,
                                                     (int)c.max_batch, 0, h->sort_bits, h->stream))))
        return fail(rc);
    if ((rc = hrc(hipMalloc(&h->sort_tmp, h->sort_tmp_bytes ? h->sort_tmp_bytes : 16)))) return fail(rc);
    const dim3 g(1024), b(256);
    hipLaunchKernelGGL(k_h_fill_u64, g, b, 0, h->stream, d.keys, (uint64_t)cap, EMPTY_KEY);
    hipLaunchKernelGGL(k_h_init_nodes, g, b, 0, h->stream, d.node, (uint64_t)cap);
    hipLaunchKernelGGL(k_h_fill_u32, g, b, 0, h->stream, d.last_prog, (uint64_t)c.max_accesses, NONE);
    hipLaunchKernelGGL(k_h_fill_u32, g, b, 0, h->stream, d.last_virt, (uint64_t)c.max_virtual, NONE);
    hipLaunchKernelGGL(k_h_fill_u32, g, b, 0, h->stream, (uint32_t *)d.ctr, 16, 0u);
    hipLaunchKernelGGL(k_h_fill_u32, g, b, 0, h->stream, (uint32_t *)d.scount, 2 * ESTRIPES, 0u);
    if ((rc = hrc(hipGetLastError())) || (rc = hrc(hipStreamSynchronize(h->stream)))) return fail(rc);
    *out = h;
    return 0;
}

int xkrt_hip_handle_domain_destroy(xkrt_hip_handle_domain_t h)
{
    if (!h) return -EINVAL;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    HDev &d = h->d;
    for (void *p : {(void *)d.keys, (void *)d.node, (void *)d.entry, (void *)d.last_prog, (void *)d.last_virt,
                    (void *)d.edges, (void *)d.ctr, (void *)d.scount, (void *)h->rec, (void *)h->keys, (void *)h->keys_sorted,
                    h->sort_tmp})
        if (p) (void)hipFree(p);
    if (h->rec_host) (void)hipHostFree(h->rec_host);
    if (h->ctr_host) (void)hipHostFree(h->ctr_host);
    if (h->scount_host) (void)hipHostFree(h->scount_host);
    if (h->staged) (void)hipEventDestroy(h->staged);
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

} // extern "C"

namespace {

int submit(xkrt_hip_handle_domain_t h, const xkrt_hip_handle_access_t *acc, const uint32_t *ids, uint64_t n,
           uint64_t *first_id)
{
    if (!h || (!acc && n)) return -EINVAL;
    if (first_id) *first_id = h->next_id;
    if (n == 0) return 0;
    if (n > h->cfg.max_batch) return -EINVAL;
    if (ids) {  // caller ids: strictly increasing across submits, below max_accesses
        for (uint64_t i = 0; i < n; ++i)
            if ((i ? ids[i] <= ids[i - 1] : (h->any_id && ids[0] <= h->last_id)) || ids[i] >= h->cfg.max_accesses)
                return ids[i] >= h->cfg.max_accesses ? -ENOSPC : -EINVAL;
    } else if (h->next_id + n > h->cfg.max_accesses) {
        return -ENOSPC;
    }
    for (uint64_t i = 0; i < n; ++i) {
        const xkrt_hip_handle_access_t &q = acc[i];
        const uint32_t prev = i ? acc[i - 1].task : h->last_task;
        if ((i || h->any_task) && q.task < prev) return -EINVAL;
        if (q.task >= VIRT || q.handle == EMPTY_KEY) return -EINVAL;
        if (q.concurrency != XKRT_HIP_CONC_SEQUENTIAL && q.concurrency != XKRT_HIP_CONC_COMMUTATIVE &&
            q.concurrency != XKRT_HIP_CONC_CONCURRENT)
            return -EINVAL;
    }
    if (h->staged_pending) {  // the staging buffer must be free
        if (hipEventSynchronize(h->staged) != hipSuccess) return -EIO;
        h->staged_pending = false;
    }
    for (uint64_t i = 0; i < n; ++i) {
        HRec &r = h->rec_host[i];
        r.handle = acc[i].handle;
        r.acc = ids ? ids[i] : (uint32_t)(h->next_id + i);
        r.task = acc[i].task;
        r.mc = acc[i].mode | (uint32_t)acc[i].concurrency << 8;
        r.pad = 0;
    }
    int rc = hrc(hipMemcpyAsync(h->rec, h->rec_host, sizeof(HRec) * n, hipMemcpyHostToDevice, h->stream));
    if (!rc) rc = hrc(hipEventRecord(h->staged, h->stream));
    if (rc) return rc;
    h->staged_pending = true;
    const uint32_t nb = (uint32_t)((n + 255) / 256);
    hipLaunchKernelGGL(k_h_slot, dim3(nb), dim3(256), 0, h->stream, h->d, h->rec, (uint32_t)n, h->keys);
    rc = hrc(hipcub::DeviceRadixSort::SortKeys(h->sort_tmp, h->sort_tmp_bytes, h->keys, h->keys_sorted, (int)n, 0,
                                                h->sort_bits, h->stream));
    if (rc) return rc;
    hipLaunchKernelGGL(k_h_replay, dim3(nb), dim3(256), 0, h->stream, h->d, h->rec, h->keys_sorted, (uint32_t)n);
    rc = hrc(hipGetLastError());
    if (rc) return rc;
    h->next_id += n;
    h->any_task = true;
    h->last_task = acc[n - 1].task;
    if (ids) {
        h->any_id = true;
        h->last_id = ids[n - 1];
    }
    h->launches++;
    return 0;
}

} // namespace

extern "C" {

int xkrt_hip_handle_submit(xkrt_hip_handle_domain_t h, const xkrt_hip_handle_access_t *acc, uint64_t n,
                           uint64_t *first_id)
{
    return submit(h, acc, nullptr, n, first_id);
}

int xkrt_hip_handle_submit_ids(xkrt_hip_handle_domain_t h, const xkrt_hip_handle_access_t *acc, const uint32_t *ids,
                               uint64_t n)
{
    if (!ids && n) return -EINVAL;
    return submit(h, acc, ids, n, nullptr);
}

int xkrt_hip_handle_edges_device(xkrt_hip_handle_domain_t h, const uint32_t **edges, const uint64_t **stripe_counts,
                                 uint32_t *nstripes, uint64_t *stripe_cap, const uint32_t **nvirtual)
{
    if (!h || !edges || !stripe_counts || !nstripes || !stripe_cap || !nvirtual) return -EINVAL;
    *edges = (const uint32_t *)h->d.edges;
    *stripe_counts = (const uint64_t *)h->d.scount;
    *nstripes = ESTRIPES;
    *stripe_cap = h->d.stripe_cap;
    *nvirtual = (const uint32_t *)&h->d.ctr[1];
    return 0;
}

int xkrt_hip_handle_edges_consume(xkrt_hip_handle_domain_t h)
{
    if (!h) return -EINVAL;
    hipLaunchKernelGGL(k_h_fill_u32, dim3(1), dim3(64), 0, h->stream, (uint32_t *)h->d.scount, 2 * ESTRIPES, 0u);
    return hrc(hipGetLastError());
}

int xkrt_hip_handle_sync(xkrt_hip_handle_domain_t h, uint64_t *nedges, uint64_t *nvirtual)
{
    if (!h) return -EINVAL;
    int rc = hrc(hipMemcpyAsync(h->ctr_host, h->d.ctr, 64, hipMemcpyDeviceToHost, h->stream));
    if (!rc) rc = hrc(hipMemcpyAsync(h->scount_host, h->d.scount, 8 * ESTRIPES, hipMemcpyDeviceToHost, h->stream));
    if (!rc) rc = hrc(hipStreamSynchronize(h->stream));
    if (rc) return rc;
    unsigned long long ne = 0;
    for (uint32_t k = 0; k < ESTRIPES; ++k) ne += std::min<unsigned long long>(h->scount_host[k], h->d.stripe_cap);
    if (nedges) *nedges = ne;
    if (nvirtual) *nvirtual = h->ctr_host[1];
    const uint32_t err = h->ctr_host[2];
    if (err) return -ENOSPC;
    return 0;
}

int xkrt_hip_handle_edges_copy(xkrt_hip_handle_domain_t h, uint32_t *edges, uint64_t cap, uint64_t *n)
{
    if (!h || !n || (cap && !edges)) return -EINVAL;
    uint64_t ne = 0;
    int rc = xkrt_hip_handle_sync(h, &ne, nullptr);
    *n = ne;
    if (rc) return rc;
    if (ne > cap) return -ENOSPC;
    uint64_t off = 0;
    for (uint32_t k = 0; k < ESTRIPES; ++k) {  // the stripes, one after the other
        const uint64_t c = std::min<unsigned long long>(h->scount_host[k], h->d.stripe_cap);
        if (!c) continue;
        rc = hrc(hipMemcpyAsync(edges + 2 * off, h->d.edges + k * h->d.stripe_cap, 8 * c, hipMemcpyDeviceToHost,
                                h->stream));
        if (rc) return rc;
        off += c;
    }
    hipLaunchKernelGGL(k_h_fill_u32, dim3(1), dim3(64), 0, h->stream, (uint32_t *)h->d.scount, 2 * ESTRIPES, 0u);
    return hrc(hipStreamSynchronize(h->stream));
}

} // extern "C"
