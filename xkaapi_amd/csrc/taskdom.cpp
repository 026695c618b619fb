// Task domains: one task domain's dependency resolution over segments and
// BLAS tiles of any (ld, sizeof_type), the dispatch of task_dependency_resolve
// (src/task/task.cc:204-299) on top of a pool of GPU-resident domains.
//
//   * the segment domain is created at the first segment (task.cc:240-259);
//   * a BLAS domain per (ld, sizeof_type) at its first tile, looked up by a
//     linear scan as task_get_dependency_domain_blas_matrix does (task.cc:
//     153-201); a new one gets every earlier segment replayed (task.cc:
//     177-198), including those of the tile's own task that come before it:
//     their PUT reaches every tree existing at the PUT pass;
//   * a segment LINKs into the segment domain and into every BLAS domain that
//     exists when it is linked (task.cc:240-259): in program order, the BLAS
//     domains created by earlier accesses;
//   * a batch's per-domain parts are staged as one unit and resolved by one
//     launch; every domain's records carry the program-wide access ids, so
//     the edges of all domains name program accesses;
//   * with release on, the edges of each launch are applied on the device
//     (release.hip): wait counters, successor lists and a ready queue, with
//     the commit guard of the launch's tasks dropped once its edges are in.
//
// Host code: the device work goes through the pool's C ABI and the release
// engine.

#include "../../include/xkrt_hip.h"
#include "release.h"

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstring>
#include <vector>

struct xkrt_hip_taskdom {
    xkrt_hip_pool_t pool = nullptr;
    xkrt_hip_taskdom_config_t cfg{};
    xkrt_hip_dep_domain_t seg = nullptr;
    struct Blas {
        uint64_t ld, s;
        xkrt_hip_dep_domain_t d;
    };
    std::vector<Blas> blas;
    // every segment enqueued: what a new BLAS domain gets replayed
    std::vector<xkrt_hip_access_t> segs;
    uint64_t next_id = 0, nseg = 0, launches = 0, edges_out = 0;
    bool any_task = false;
    uint32_t last_task = 0;
    int err = 0;  // a domain's device error: sticky
    // per-submit work lists, by pool domain index
    std::vector<std::vector<xkrt_hip_access_t>> part;
    std::vector<std::vector<uint32_t>> part_ids;
    std::vector<uint32_t> e32, edom;
    // handle accesses: their domain, and the program id of each of its accesses
    xkrt_hip_handle_domain_t hdom = nullptr;
    std::vector<uint32_t> hids;    // release off: program id of each handle-domain access
    std::vector<uint32_t> hbatch;  // this batch's handle accesses' program ids
    std::vector<xkrt_hip_handle_access_t> hpart;
    std::vector<uint32_t> he32;
    // release on the device
    xkrt_release::Engine *rel = nullptr;
    uint64_t next_task = 0;              // tasks are numbered 0, 1, ... in submit order
    uint64_t open_t0 = 0, open_t1 = 0;   // tasks staged into the open launch
    std::vector<uint32_t> btask;
    std::vector<uint8_t> bmode;
};

namespace {

// lifetime capacity of a task domain's handle map (accesses, and inserted empty writes), and
// the handle accesses one batch may hold
constexpr uint64_t HANDLE_MAX_ACCESSES = 1ull << 22;
constexpr uint64_t HANDLE_MAX_BATCH = 1ull << 18;

int dom_index(xkrt_hip_dep_domain_t d)
{
    uint32_t i = 0;
    xkrt_hip_dep_index(d, &i);
    return (int)i;
}

xkrt_hip_access_t plain(const xkrt_hip_task_access_t &q)
{
    xkrt_hip_access_t a{};
    a.addr = q.addr;
    a.m = q.m;
    a.n = q.n;
    a.task = q.task;
    a.mode = q.mode;
    a.kind = q.kind;
    return a;
}

// a new domain of the pool: room in the per-index tables
void admit(xkrt_hip_taskdom *t, xkrt_hip_dep_domain_t d)
{
    const size_t i = (size_t)dom_index(d);
    if (t->part.size() <= i) {
        t->part.resize(i + 1);
        t->part_ids.resize(i + 1);
    }
}

// launches the open launch; with release on, its edges are applied and
// consumed on the device, then its tasks' commit guards dropped
int launch(xkrt_hip_taskdom *t)
{
    uint64_t ticket = 0;
    int rc = xkrt_hip_pool_launch(t->pool, &ticket);
    if (rc) return rc;
    t->launches++;
    if (t->rel) {
        const uint32_t *edges = nullptr, *doms = nullptr;
        const uint64_t *count = nullptr;
        rc = xkrt_hip_pool_edges_device(t->pool, &edges, &doms, &count);
        xkrt_release::HandleEdges he{};
        if (!rc && t->hdom)
            rc = xkrt_hip_handle_edges_device(t->hdom, &he.edges, &he.count, &he.stripes, &he.stripe_cap,
                                              &he.nvirtual);
        if (!rc)
            rc = xkrt_release::apply(t->rel, edges, count, t->hdom ? &he : nullptr, (uint32_t)t->open_t0,
                                     (uint32_t)t->open_t1);
        if (!rc) rc = xkrt_hip_pool_edges_consume(t->pool);
        if (!rc && t->hdom) rc = xkrt_hip_handle_edges_consume(t->hdom);
    }
    t->open_t0 = t->open_t1;
    return rc;
}

} // namespace

extern "C" {

int xkrt_hip_taskdom_create(const xkrt_hip_taskdom_config_t *cfg, void *stream, xkrt_hip_taskdom_t *out)
{
    if (!out) return -EINVAL;
    *out = nullptr;
    xkrt_hip_taskdom_config_t c{};
    if (cfg) c = *cfg;
    if (c.kernel > 2) return -EINVAL;
    if (!c.max_rects) c.max_rects = 1ull << 20;
    if (!c.max_edges) c.max_edges = 1ull << 22;
    if (!c.seg_col_bytes) c.seg_col_bytes = 1ull << 16;
    if (c.seg_col_bytes > 0xFFFFFFFFull) return -ERANGE;
    if (!c.seg_bin_x_bytes) c.seg_bin_x_bytes = (uint32_t)std::min<uint64_t>(c.seg_col_bytes, 4096);
    if (!c.seg_bin_y_cols) c.seg_bin_y_cols = 1;
    if (!c.blas_bin_x_bytes) c.blas_bin_x_bytes = 1024;
    if (!c.blas_bin_y_cols) c.blas_bin_y_cols = 128;
    xkrt_hip_pool_config_t pc{};
    pc.kernel = c.kernel;
    pc.max_rects = c.max_rects;
    pc.max_edges = c.max_edges;
    pc.budget_bytes = c.budget_bytes;
    if (c.release) {
        if (!c.max_tasks) c.max_tasks = 1ull << 20;
        if (!c.max_accesses) c.max_accesses = 4 * c.max_tasks;
        if (!c.max_succ) c.max_succ = 2 * c.max_accesses;
        if (!c.max_virtual) c.max_virtual = c.max_tasks;
    }
    xkrt_hip_pool_t p = nullptr;
    int rc = xkrt_hip_pool_create(&pc, stream, &p);
    if (rc) return rc;
    xkrt_hip_taskdom *t = new xkrt_hip_taskdom();
    t->pool = p;
    t->cfg = c;
    if (c.release) {
        rc = xkrt_release::create(xkrt_hip_pool_stream(p), c.max_tasks, c.max_accesses, c.max_succ, c.max_virtual,
                                  &t->rel);
        if (rc) {
            xkrt_hip_pool_destroy(p);
            delete t;
            return rc;
        }
    }
    *out = t;
    return 0;
}

int xkrt_hip_taskdom_destroy(xkrt_hip_taskdom_t t)
{
    if (!t) return -EINVAL;
    xkrt_release::destroy(t->rel);
    if (t->hdom) xkrt_hip_handle_domain_destroy(t->hdom);
    const int rc = xkrt_hip_pool_destroy(t->pool);  // releases its domains
    delete t;
    return rc;
}

int xkrt_hip_taskdom_submit(xkrt_hip_taskdom_t t, const xkrt_hip_task_access_t *acc, uint64_t n, uint64_t *first_id)
{
    if (!t || (!acc && n)) return -EINVAL;
    if (first_id) *first_id = t->next_id;
    if (t->err) return -ECANCELED;
    if (n == 0) return 0;
    // validation first: a refused batch leaves nothing behind
    if (t->next_id + n > 0xFFFFFFFFull) return -ERANGE;  // program access ids are 32-bit
    uint64_t nh = 0;  // handle accesses
    for (uint64_t i = 0; i < n; ++i) {
        const xkrt_hip_task_access_t &q = acc[i];
        const uint32_t prev = i ? acc[i - 1].task : t->last_task;
        if ((i || t->any_task) && q.task < prev) return -EINVAL;
        // released on the device, tasks are numbered 0, 1, ... in submit order
        if (t->rel && (i ? q.task - prev > 1 : q.task != t->next_task)) return -EINVAL;
        if (q.kind == XKRT_HIP_KIND_HANDLE) {
            if (q.addr == UINT64_MAX || q.task >= 0x80000000u) return -EINVAL;
            if (q.concurrency > XKRT_HIP_CONC_CONCURRENT || q.concurrency == 3) return -EINVAL;
            ++nh;
            continue;
        }
        if (q.kind != XKRT_HIP_KIND_TILE && q.kind != XKRT_HIP_KIND_SEGMENT) return -EINVAL;
        if (q.m == 0) return -EINVAL;
        if (q.kind == XKRT_HIP_KIND_TILE && (q.n == 0 || q.ld == 0 || q.sizeof_type == 0)) return -EINVAL;
    }
    if (nh > HANDLE_MAX_BATCH) return -EINVAL;
    if (nh && !t->rel && t->hids.size() + nh > HANDLE_MAX_ACCESSES) return -ENOSPC;
    if (t->rel && ((uint64_t)acc[n - 1].task + 1 > t->cfg.max_tasks || t->next_id + n > t->cfg.max_accesses))
        return -ENOSPC;
    // domains this batch creates (destroyed again if the batch is refused)
    std::vector<xkrt_hip_dep_domain_t> created;
    const size_t nblas0 = t->blas.size();
    const size_t nsegs0 = t->segs.size();
    auto undo = [&](int rc) {
        for (auto &v : t->part) v.clear();
        for (auto &v : t->part_ids) v.clear();
        t->blas.resize(nblas0);
        t->segs.resize(nsegs0);
        for (xkrt_hip_dep_domain_t d : created) {
            if (d == t->seg) t->seg = nullptr;
            xkrt_hip_dep_domain_destroy(d);
        }
        return rc;
    };
    for (auto &v : t->part) v.clear();
    for (auto &v : t->part_ids) v.clear();
    for (uint64_t i = 0; i < n; ++i) {
        const xkrt_hip_task_access_t &q = acc[i];
        const uint32_t id = (uint32_t)(t->next_id + i);
        const xkrt_hip_access_t a = plain(q);
        if (q.kind == XKRT_HIP_KIND_HANDLE) continue;  // submitted below, once the rest is staged
        if (q.kind == XKRT_HIP_KIND_SEGMENT) {
            if (!t->seg) {
                int rc = xkrt_hip_pool_seg_domain_create(t->pool, t->cfg.seg_col_bytes, t->cfg.seg_bin_x_bytes,
                                                         t->cfg.seg_bin_y_cols, 0, 0, &t->seg);
                if (rc) return undo(rc);
                created.push_back(t->seg);
                admit(t, t->seg);
            }
            const int si = dom_index(t->seg);
            t->part[si].push_back(a);
            t->part_ids[si].push_back(id);
            for (auto &b : t->blas) {  // LINK-only in every BLAS domain that exists now
                const int bi = dom_index(b.d);
                t->part[bi].push_back(a);
                t->part_ids[bi].push_back(id);
            }
            t->segs.push_back(a);
        } else {
            auto it = std::find_if(t->blas.begin(), t->blas.end(),
                                   [&](const xkrt_hip_taskdom::Blas &b) { return b.ld == q.ld && b.s == q.sizeof_type; });
            if (it == t->blas.end()) {
                if (t->blas.size() >= 255) return undo(-EINVAL);
                xkrt_hip_dep_domain_t d = nullptr;
                int rc = xkrt_hip_pool_domain_create(t->pool, q.ld, q.sizeof_type, t->cfg.blas_bin_x_bytes,
                                                     t->cfg.blas_bin_y_cols, 0, 0, &d);
                if (rc) return undo(rc);
                created.push_back(d);
                admit(t, d);
                // every earlier segment is PUT into the new tree with no LINK
                rc = xkrt_hip_dep_replay_segments(d, t->segs.data(), t->segs.size());
                if (rc) return undo(rc);
                t->blas.push_back({q.ld, q.sizeof_type, d});
                it = t->blas.end() - 1;
            }
            const int bi = dom_index(it->d);
            t->part[bi].push_back(a);
            t->part_ids[bi].push_back(id);
        }
    }
    // stage every domain's part as one unit; launch the open launch first if it is full
    std::vector<xkrt_hip_dep_domain_t> doms;
    std::vector<const xkrt_hip_access_t *> ptr;
    std::vector<const uint32_t *> idp;
    std::vector<uint64_t> cnt, first;
    auto add = [&](xkrt_hip_dep_domain_t d) {
        if (!d) return;
        const int i = dom_index(d);
        if (t->part[i].empty()) return;
        doms.push_back(d);
        ptr.push_back(t->part[i].data());
        idp.push_back(t->part_ids[i].data());
        cnt.push_back(t->part[i].size());
    };
    add(t->seg);
    for (auto &b : t->blas) add(b.d);
    first.resize(doms.size());
    int rc = xkrt_hip_pool_stage_group(t->pool, (uint32_t)doms.size(), doms.data(), ptr.data(), idp.data(), cnt.data(),
                                       first.data());
    if (rc == -ENOSPC) {
        rc = launch(t);
        if (!rc) {
            rc = xkrt_hip_pool_stage_group(t->pool, (uint32_t)doms.size(), doms.data(), ptr.data(), idp.data(),
                                           cnt.data(), first.data());
            if (rc == -ENOSPC) rc = -EINVAL;  // larger than one launch
        }
    }
    if (rc) return undo(rc);
    if (nh) {
        if (!t->hdom) {
            xkrt_hip_handle_config_t hc{};
            hc.max_batch = HANDLE_MAX_BATCH;
            if (t->rel) {
                // program ids, and its edges applied on the device: on the pool's stream, so that
                // a launch's apply follows the handle launches of its batches
                hc.max_accesses = t->cfg.max_accesses;
                hc.max_virtual = t->cfg.max_virtual;
                hc.max_edges = std::max<uint64_t>(1ull << 22, 2 * t->cfg.max_rects);
                rc = xkrt_hip_handle_domain_create(&hc, xkrt_hip_pool_stream(t->pool), &t->hdom);
            } else {
                hc.max_accesses = HANDLE_MAX_ACCESSES;
                hc.max_virtual = HANDLE_MAX_ACCESSES;
                rc = xkrt_hip_handle_domain_create(&hc, nullptr, &t->hdom);
            }
            if (rc) {
                t->err = rc;  // the rest of the batch is staged already
                return rc;
            }
        }
        t->hpart.clear();
        t->hbatch.clear();
        for (uint64_t i = 0; i < n; ++i) {
            if (acc[i].kind != XKRT_HIP_KIND_HANDLE) continue;
            xkrt_hip_handle_access_t h{};
            h.handle = acc[i].addr;
            h.task = acc[i].task;
            h.mode = acc[i].mode;
            h.concurrency = acc[i].concurrency ? acc[i].concurrency : (uint8_t)XKRT_HIP_CONC_SEQUENTIAL;
            t->hpart.push_back(h);
            t->hbatch.push_back((uint32_t)(t->next_id + i));
        }
        if (t->rel) {
            rc = xkrt_hip_handle_submit_ids(t->hdom, t->hpart.data(), t->hbatch.data(), t->hpart.size());
        } else {
            t->hids.insert(t->hids.end(), t->hbatch.begin(), t->hbatch.end());
            rc = xkrt_hip_handle_submit(t->hdom, t->hpart.data(), t->hpart.size(), nullptr);
        }
        if (rc) {
            t->err = rc;
            return rc;
        }
    }
    if (t->rel) {
        t->btask.resize(n);
        t->bmode.resize(n);
        for (uint64_t i = 0; i < n; ++i) {
            t->btask[i] = acc[i].task;
            t->bmode[i] = acc[i].mode;
        }
        rc = xkrt_release::add(t->rel, t->next_id, t->btask.data(), t->bmode.data(), n);
        if (rc) {
            t->err = rc;  // staged but not registered: nothing of the task domain can be trusted
            return rc;
        }
        t->next_task = (uint64_t)acc[n - 1].task + 1;
        t->open_t1 = t->next_task;
    }
    for (uint64_t i = 0; i < n; ++i) t->nseg += acc[i].kind == XKRT_HIP_KIND_SEGMENT;
    t->next_id += n;
    t->any_task = true;
    t->last_task = acc[n - 1].task;
    return 0;
}

int xkrt_hip_taskdom_sync(xkrt_hip_taskdom_t t, uint64_t *nedges)
{
    if (!t) return -EINVAL;
    uint64_t nr = 0, na = 0, nb = 0;
    int rc = xkrt_hip_pool_open(t->pool, &nr, &na, &nb);
    if (!rc && nb) rc = launch(t);
    uint64_t ne = 0;
    if (!rc) rc = xkrt_hip_pool_sync(t->pool, &ne);
    if (nedges) *nedges = ne;
    if (rc) return rc;
    std::vector<xkrt_hip_dep_domain_t> all;
    if (t->seg) all.push_back(t->seg);
    for (auto &b : t->blas) all.push_back(b.d);
    for (xkrt_hip_dep_domain_t d : all) {
        const int e = xkrt_hip_dep_sync(d, nullptr);
        if (e && !t->err) t->err = e;
    }
    if (t->rel && !t->err) t->err = xkrt_release::check(t->rel);
    if (t->hdom) {
        uint64_t hn = 0;
        const int e = xkrt_hip_handle_sync(t->hdom, &hn, nullptr);
        if (e && !t->err) t->err = e;
        if (nedges) *nedges += hn;
    }
    return t->err;
}

int xkrt_hip_taskdom_edges_copy(xkrt_hip_taskdom_t t, uint64_t *edges, uint64_t cap, uint64_t *n)
{
    if (!t || !n || (cap && !edges)) return -EINVAL;
    if (t->rel) return -EPERM;  // applied and consumed on the device
    uint64_t ne = 0, hn = 0;
    int rc = xkrt_hip_pool_sync(t->pool, &ne);
    if (!rc && t->hdom) rc = xkrt_hip_handle_sync(t->hdom, &hn, nullptr);
    *n = ne + hn;
    if (rc) return rc;
    if (ne + hn > cap) return -ENOSPC;
    if (hn) {
        t->he32.resize(2 * hn);
        uint64_t got = 0;
        rc = xkrt_hip_handle_edges_copy(t->hdom, t->he32.data(), hn, &got);
        if (rc) return rc;
        hn = got;
        // handle-domain ids -> program ids; an inserted empty write keeps its index
        auto prog = [&](uint32_t id) -> uint64_t {
            return id & XKRT_HIP_VIRTUAL_ID ? (XKRT_HIP_TASKDOM_VIRTUAL | (id & ~XKRT_HIP_VIRTUAL_ID)) : t->hids[id];
        };
        for (uint64_t i = 0; i < hn; ++i) {
            if (!(t->he32[2 * i] & XKRT_HIP_VIRTUAL_ID) && t->he32[2 * i] >= t->hids.size()) return -EIO;
            if (!(t->he32[2 * i + 1] & XKRT_HIP_VIRTUAL_ID) && t->he32[2 * i + 1] >= t->hids.size()) return -EIO;
            edges[2 * (ne + i)] = prog(t->he32[2 * i]);
            edges[2 * (ne + i) + 1] = prog(t->he32[2 * i + 1]);
        }
    }
    t->e32.resize(2 * ne + 2);
    t->edom.resize(ne + 1);
    uint64_t got = 0;
    rc = xkrt_hip_pool_edges_copy(t->pool, t->e32.data(), t->edom.data(), ne, &got);
    if (rc) return rc;
    for (uint64_t i = 0; i < 2 * got; ++i) edges[i] = t->e32[i];  // program ids already
    if (got < ne && hn)  // edges of failed domains were dropped: close the gap
        memmove(edges + 2 * got, edges + 2 * ne, 16 * hn);
    *n = got + hn;
    t->edges_out += got + hn;
    return 0;
}

int xkrt_hip_taskdom_ready(xkrt_hip_taskdom_t t, uint32_t *tasks, uint64_t cap, uint64_t *n)
{
    if (!t || !n) return -EINVAL;
    if (!t->rel) return -EPERM;
    return xkrt_release::ready(t->rel, tasks, cap, n);
}

int xkrt_hip_taskdom_complete(xkrt_hip_taskdom_t t, const uint32_t *tasks, uint64_t n)
{
    if (!t) return -EINVAL;
    if (!t->rel) return -EPERM;
    return xkrt_release::complete(t->rel, tasks, n);
}

int xkrt_hip_taskdom_complete_wait(xkrt_hip_taskdom_t t, const uint32_t *tasks, uint64_t n)
{
    if (!t) return -EINVAL;
    if (!t->rel) return -EPERM;
    return xkrt_release::complete_wait(t->rel, tasks, n);
}

int xkrt_hip_taskdom_stats(xkrt_hip_taskdom_t t, xkrt_hip_taskdom_stats_t *out)
{
    if (!t || !out) return -EINVAL;
    memset(out, 0, sizeof(*out));
    out->accesses = t->next_id;
    out->segments = t->nseg;
    out->blas_domains = t->blas.size();
    out->launches = t->launches;
    out->edges = t->edges_out;
    if (t->rel) {
        xkrt_release::Counts c{};
        const int rc = xkrt_release::counts(t->rel, &c);
        out->succ_nodes = c.nodes;
        out->released = c.released;
        out->completed = c.completed;
        return rc;
    }
    return 0;
}

} // extern "C"
