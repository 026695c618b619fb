/*
 * ORACLE -- test infrastructure only.  Never linked into, loaded by or called
 * from the product path; only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg use it, as the checker.
 *
 * A CPU restatement of the reference's KHP-tree dependency resolver and of
 * its BLAS coherency controller, over the fragmentation the KHP tree
 * produces, stored flat instead of as a tree:
 *
 *   - slabs: disjoint dim-0 (byte-offset) intervals, sorted.  Every endpoint
 *     of every inserted rect that falls strictly inside a slab splits it
 *     (khp-tree.hpp:1233-1303, case 3 at k = 0: the slab node and all its
 *     k = 1 descendants are shrunk and their sides re-inserted), and rect
 *     pieces that fall in no slab become new slabs (cases 1/2/4).
 *   - pieces: inside each slab, disjoint dim-1 (column) intervals, sorted;
 *     split at the dim-1 endpoints of the rects inserted over that slab
 *     (case 3 / case 4 at k = 1), new pieces for uncovered parts of the rect.
 *
 * `on_insert` is invoked exactly on the nodes the tree passes to it during an
 * insert: re-inserted sides of split slabs and pieces, new nodes, and exact
 * matches (khp-tree.hpp:961-996 insert_fixup / insert_finalize).  This keeps
 * the reference's multi-rect marking behaviour (dependency-tree.hpp:229-256:
 * a node is marked as a whole if it intersects any rect of the access).
 *
 * Parity pins (tests): the fragment lists and edge lists are compared with
 * oracle/_ref (the reference engine, built from its sources), with a
 * pointwise brute force on tiny spaces, and with closed-form streams.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* flat slab/piece store                                               */
/* ------------------------------------------------------------------ */

typedef struct {
    size_t esz;                                /* bytes per piece; first two u64 are y0, y1 */
    void (*copy)(void *dst, const void *src);  /* deep copy (state inheritance on split) */
    void (*init)(void *p);                     /* state of a never-seen region */
    void (*fini)(void *p);
} ptype_t;

typedef struct {
    uint64_t x0, x1;
    uint32_t np, cap;
    unsigned char *p;
} slab_t;

typedef struct {
    const ptype_t *pt;
    slab_t *s;
    uint32_t ns, cap;
} flat_t;

#define PC(sl, i, pt) ((void *)((sl)->p + (size_t)(i) * (pt)->esz))
#define PY0(pc) (((uint64_t *)(pc))[0])
#define PY1(pc) (((uint64_t *)(pc))[1])

typedef void (*insert_cb)(void *piece, uint64_t x0, uint64_t x1, void *ctx);

static void *xmalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p) abort();
    return p;
}

static void *xrealloc(void *q, size_t n)
{
    void *p = realloc(q, n ? n : 1);
    if (!p) abort();
    return p;
}

static void slab_reserve(slab_t *sl, const ptype_t *pt, uint32_t need)
{
    if (need <= sl->cap) return;
    uint32_t c = sl->cap ? sl->cap : 4;
    while (c < need) c *= 2;
    sl->p = xrealloc(sl->p, (size_t)c * pt->esz);
    sl->cap = c;
}

/* insert a hole at piece index i (the caller fills it) */
static void *slab_insert_at(slab_t *sl, const ptype_t *pt, uint32_t i)
{
    slab_reserve(sl, pt, sl->np + 1);
    memmove(sl->p + (size_t)(i + 1) * pt->esz, sl->p + (size_t)i * pt->esz, (size_t)(sl->np - i) * pt->esz);
    sl->np++;
    return PC(sl, i, pt);
}

static void flat_reserve(flat_t *f, uint32_t need)
{
    if (need <= f->cap) return;
    uint32_t c = f->cap ? f->cap : 16;
    while (c < need) c *= 2;
    f->s = xrealloc(f->s, (size_t)c * sizeof(slab_t));
    f->cap = c;
}

static slab_t *flat_insert_at(flat_t *f, uint32_t i)
{
    flat_reserve(f, f->ns + 1);
    memmove(f->s + i + 1, f->s + i, (size_t)(f->ns - i) * sizeof(slab_t));
    f->ns++;
    memset(f->s + i, 0, sizeof(slab_t));
    return f->s + i;
}

static void flat_free(flat_t *f)
{
    for (uint32_t i = 0; i < f->ns; ++i) {
        for (uint32_t j = 0; j < f->s[i].np; ++j) f->pt->fini(PC(&f->s[i], j, f->pt));
        free(f->s[i].p);
    }
    free(f->s);
    f->s = NULL;
    f->ns = f->cap = 0;
}

/* first slab whose x1 > x */
static uint32_t flat_lower(const flat_t *f, uint64_t x)
{
    uint32_t lo = 0, hi = f->ns;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (f->s[mid].x1 <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* first piece whose y1 > y */
static uint32_t slab_lower(const slab_t *sl, const ptype_t *pt, uint64_t y)
{
    uint32_t lo = 0, hi = sl->np;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (PY1(PC(sl, mid, pt)) <= y) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* deep-copy slab src's pieces into a new slab dst with dim-0 [x0, x1) */
static void slab_clone(slab_t *dst, const slab_t *src, const ptype_t *pt, uint64_t x0, uint64_t x1)
{
    dst->x0 = x0; dst->x1 = x1;
    dst->np = 0; dst->cap = 0; dst->p = NULL;
    slab_reserve(dst, pt, src->np);
    for (uint32_t j = 0; j < src->np; ++j) pt->copy(PC(dst, j, pt), PC(src, j, pt));
    dst->np = src->np;
}

/*
 * Insert rect [x0,x1) x [y0,y1) (khp-tree.hpp:1180-1369).  cb is on_insert.
 */
static void flat_insert(flat_t *f, uint64_t x0, uint64_t x1, uint64_t y0, uint64_t y1, insert_cb cb, void *ctx)
{
    const ptype_t *pt = f->pt;
    if (x0 >= x1 || y0 >= y1) return;

    /* (1) dim-0 splits of slabs straddling x0 or x1: the parts outside
     * [x0, x1) are the re-inserted sides (on_insert on each of their pieces);
     * the part inside keeps the original nodes. */
    uint32_t i = flat_lower(f, x0);
    if (i < f->ns && f->s[i].x0 < x0 && x0 < f->s[i].x1) {
        slab_t *sl = f->s + i;
        slab_t side;
        slab_clone(&side, sl, pt, sl->x0, x0);
        f->s[i].x0 = x0;
        slab_t *ns = flat_insert_at(f, i);
        *ns = side;
        for (uint32_t j = 0; j < ns->np; ++j) cb(PC(ns, j, pt), ns->x0, ns->x1, ctx);
        i++;
    }
    uint32_t k = flat_lower(f, x1 - 1);  /* slab containing x1-1, if any */
    if (k < f->ns && f->s[k].x0 < x1 && x1 < f->s[k].x1) {
        slab_t *sl = f->s + k;
        slab_t side;
        slab_clone(&side, sl, pt, x1, sl->x1);
        f->s[k].x1 = x1;
        slab_t *ns = flat_insert_at(f, k + 1);
        *ns = side;
        for (uint32_t j = 0; j < ns->np; ++j) cb(PC(ns, j, pt), ns->x0, ns->x1, ctx);
    }

    /* (2)+(3): walk [x0, x1): new slabs for gaps, dim-1 refinement inside */
    uint64_t x = x0;
    i = flat_lower(f, x0);
    while (x < x1) {
        if (i < f->ns && f->s[i].x0 <= x) {
            slab_t *sl = f->s + i;  /* inside [x0, x1) after (1) */
            /* (3) dim-1: split pieces straddling y0 / y1, fill gaps */
            uint32_t j = slab_lower(sl, pt, y0);
            if (j < sl->np && PY0(PC(sl, j, pt)) < y0 && y0 < PY1(PC(sl, j, pt))) {
                void *orig = PC(sl, j, pt);
                void *side = xmalloc(pt->esz);
                pt->copy(side, orig);
                PY1(side) = y0;
                PY0(orig) = y0;
                void *slot = slab_insert_at(sl, pt, j);
                memcpy(slot, side, pt->esz);
                free(side);
                cb(PC(sl, j, pt), sl->x0, sl->x1, ctx);
                j++;
            }
            uint32_t l = slab_lower(sl, pt, y1 - 1);
            if (l < sl->np && PY0(PC(sl, l, pt)) < y1 && y1 < PY1(PC(sl, l, pt))) {
                void *orig = PC(sl, l, pt);
                void *side = xmalloc(pt->esz);
                pt->copy(side, orig);
                PY0(side) = y1;
                PY1(orig) = y1;
                void *slot = slab_insert_at(sl, pt, l + 1);
                memcpy(slot, side, pt->esz);
                free(side);
                cb(PC(sl, l + 1, pt), sl->x0, sl->x1, ctx);
            }
            uint64_t y = y0;
            j = slab_lower(sl, pt, y0);
            while (y < y1) {
                if (j < sl->np && PY0(PC(sl, j, pt)) <= y) {
                    cb(PC(sl, j, pt), sl->x0, sl->x1, ctx);
                    y = PY1(PC(sl, j, pt));
                    j++;
                } else {
                    uint64_t ye = (j < sl->np && PY0(PC(sl, j, pt)) < y1) ? PY0(PC(sl, j, pt)) : y1;
                    void *np_ = slab_insert_at(sl, pt, j);
                    pt->init(np_);
                    PY0(np_) = y; PY1(np_) = ye;
                    cb(np_, sl->x0, sl->x1, ctx);
                    y = ye;
                    j++;
                }
            }
            x = sl->x1;
            i++;
        } else {
            uint64_t xe = (i < f->ns && f->s[i].x0 < x1) ? f->s[i].x0 : x1;
            slab_t *ns = flat_insert_at(f, i);
            ns->x0 = x; ns->x1 = xe;
            slab_reserve(ns, pt, 1);
            void *p = PC(ns, 0, pt);
            pt->init(p);
            PY0(p) = y0; PY1(p) = y1;
            ns->np = 1;
            cb(p, x, xe, ctx);
            x = xe;
            i++;
        }
    }
}

typedef void (*visit_cb)(void *piece, uint64_t x0, uint64_t x1, void *ctx);

static void flat_intersect(flat_t *f, uint64_t x0, uint64_t x1, uint64_t y0, uint64_t y1, visit_cb cb, void *ctx)
{
    const ptype_t *pt = f->pt;
    if (x0 >= x1 || y0 >= y1) return;
    for (uint32_t i = flat_lower(f, x0); i < f->ns && f->s[i].x0 < x1; ++i) {
        slab_t *sl = f->s + i;
        for (uint32_t j = slab_lower(sl, pt, y0); j < sl->np && PY0(PC(sl, j, pt)) < y1; ++j)
            cb(PC(sl, j, pt), sl->x0, sl->x1, ctx);
    }
}

static uint64_t flat_count(flat_t *f)
{
    uint64_t n = 0;
    for (uint32_t i = 0; i < f->ns; ++i) n += f->s[i].np;
    return n;
}

/* ------------------------------------------------------------------ */
/* dependency resolver                                                 */
/* ------------------------------------------------------------------ */

enum { M_R = 1, M_W = 2 };

typedef struct {
    uint64_t y0, y1;
    int64_t lw;       /* last_write access index, -1 = none (dependency-tree.hpp:105) */
    uint32_t nr, cr;  /* last_reads (dependency-tree.hpp:102) */
    uint32_t *r;
} dpiece_t;

static void dp_copy(void *dst, const void *src)
{
    const dpiece_t *s = src;
    dpiece_t *d = dst;
    *d = *s;
    d->cr = s->nr;
    d->r = s->nr ? xmalloc(sizeof(uint32_t) * s->nr) : NULL;
    if (s->nr) memcpy(d->r, s->r, sizeof(uint32_t) * s->nr);
}

static void dp_init(void *p)
{
    dpiece_t *d = p;
    memset(d, 0, sizeof(*d));
    d->lw = -1;
}

static void dp_fini(void *p) { free(((dpiece_t *)p)->r); }

static const ptype_t DPT = {sizeof(dpiece_t), dp_copy, dp_init, dp_fini};

typedef struct {
    uint32_t *v;
    uint32_t n, c;
} u32vec_t;

static void vpush(u32vec_t *v, uint32_t x)
{
    if (v->n == v->c) {
        v->c = v->c ? 2 * v->c : 8;
        v->v = xrealloc(v->v, sizeof(uint32_t) * v->c);
    }
    v->v[v->n++] = x;
}

typedef struct {
    const uint32_t *task;
    const uint8_t *mode;
    const uint8_t *nmark;
    const uint64_t *marks;  /* the rects on_insert tests a node against, mstride u64 per access */
    uint32_t mstride;
    u32vec_t *succ;  /* per predecessor access: successor access indices */
    uint32_t cur;    /* access being linked / put */
    uint64_t cnt;
} dctx_t;

static int rect_hits(const uint64_t *r, uint64_t x0, uint64_t x1, uint64_t y0, uint64_t y1)
{
    return r[0] < x1 && x0 < r[1] && r[2] < y1 && y0 < r[3];
}

/* task.hpp:494-512: skip if pred's last successor is an access of succ's task */
static void precedes(dctx_t *c, uint32_t pred, uint32_t succ)
{
    u32vec_t *v = c->succ + pred;
    if (v->n && c->task[v->v[v->n - 1]] == c->task[succ]) return;
    vpush(v, succ);
}

/* dependency-tree.hpp:311-355 (SEARCH_TYPE_RESOLVE) */
static void d_on_intersect(void *pc, uint64_t x0, uint64_t x1, void *ctx)
{
    (void)x0; (void)x1;
    dctx_t *c = ctx;
    dpiece_t *p = pc;
    const uint8_t m = c->mode[c->cur];
    if ((m & M_W) && p->nr) {
        for (uint32_t i = 0; i < p->nr; ++i) precedes(c, p->r[i], c->cur);
    } else if (p->lw >= 0) {
        precedes(c, (uint32_t)p->lw, c->cur);
    }
}

/* dependency-tree.hpp:229-256 */
static void d_on_insert(void *pc, uint64_t x0, uint64_t x1, void *ctx)
{
    dctx_t *c = ctx;
    dpiece_t *p = pc;
    const uint32_t a = c->cur;
    const uint8_t m = c->mode[a];
    for (int r = 0; r < c->nmark[a]; ++r) {
        if (rect_hits(c->marks + (uint64_t)c->mstride * a + 4 * r, x0, x1, p->y0, p->y1)) {
            if (m & M_W) {
                p->nr = 0;
                p->lw = a;
            } else if (m == M_R) {
                if (p->nr == p->cr) {
                    p->cr = p->cr ? 2 * p->cr : 4;
                    p->r = xrealloc(p->r, sizeof(uint32_t) * p->cr);
                }
                p->r[p->nr++] = a;
            }
            break;
        }
    }
}

static void d_count(void *pc, uint64_t x0, uint64_t x1, void *ctx)
{
    (void)pc; (void)x0; (void)x1;
    ((dctx_t *)ctx)->cnt++;
}

/*
 * Resolve one dependency domain (src/task/task.cc:285-299: LINK every access
 * of a task in slot order, then PUT every access).  Accesses are in spawn
 * order, grouped by task.  rects: 8 u64 per access, nrect 1 or 2.  Edges are
 * (pred, succ) access indices, grouped by predecessor then successor order.
 * frag (nullable): per access, the fragments intersecting its rects after its
 * task's PUT pass.  Returns the edge count (may exceed edge_cap).
 */
static int64_t resolve_impl(uint64_t nacc, const uint32_t *task, const uint8_t *mode, const uint8_t *nrect,
                            const uint64_t *rects, uint32_t rstride, const uint8_t *nmark, const uint64_t *marks,
                            uint32_t mstride, uint64_t edge_cap, uint32_t *edges, uint64_t *frag)
{
    flat_t f = {&DPT, NULL, 0, 0};
    dctx_t c = {task, mode, nmark, marks, mstride, calloc(nacc ? nacc : 1, sizeof(u32vec_t)), 0, 0};
    uint64_t i = 0;
    while (i < nacc) {
        uint64_t j = i;
        while (j < nacc && task[j] == task[i]) ++j;
        for (uint64_t a = i; a < j; ++a) {
            c.cur = (uint32_t)a;
            for (int r = 0; r < nrect[a]; ++r) {
                const uint64_t *q = rects + (uint64_t)rstride * a + 4 * r;
                flat_intersect(&f, q[0], q[1], q[2], q[3], d_on_intersect, &c);
            }
        }
        for (uint64_t a = i; a < j; ++a) {
            c.cur = (uint32_t)a;
            for (int r = 0; r < nrect[a]; ++r) {
                const uint64_t *q = rects + (uint64_t)rstride * a + 4 * r;
                flat_insert(&f, q[0], q[1], q[2], q[3], d_on_insert, &c);
            }
        }
        if (frag)
            for (uint64_t a = i; a < j; ++a) {
                c.cnt = 0;
                c.cur = (uint32_t)a;
                /* count distinct fragments hit by any rect of the access */
                for (int r = 0; r < nrect[a]; ++r) {
                    const uint64_t *q = rects + 8 * a + 4 * r;
                    flat_intersect(&f, q[0], q[1], q[2], q[3], d_count, &c);
                }
                if (nrect[a] == 2) {
                    /* a fragment hit by both rects was counted twice */
                    const uint64_t *q0 = rects + 8 * a, *q1 = rects + 8 * a + 4;
                    uint64_t ix0 = q0[0] > q1[0] ? q0[0] : q1[0], ix1 = q0[1] < q1[1] ? q0[1] : q1[1];
                    (void)ix0; (void)ix1;
                    for (uint32_t s = flat_lower(&f, q0[0]); s < f.ns && f.s[s].x0 < q0[1]; ++s)
                        for (uint32_t t = slab_lower(f.s + s, &DPT, q0[2]);
                             t < f.s[s].np && PY0(PC(f.s + s, t, &DPT)) < q0[3]; ++t) {
                            dpiece_t *p = PC(f.s + s, t, &DPT);
                            if (rect_hits(q1, f.s[s].x0, f.s[s].x1, p->y0, p->y1)) c.cnt--;
                        }
                }
                frag[a] = c.cnt;
            }
        i = j;
    }
    uint64_t ne = 0;
    for (uint64_t p = 0; p < nacc; ++p) {
        for (uint32_t k = 0; k < c.succ[p].n; ++k) {
            if (ne < edge_cap) {
                edges[2 * ne] = (uint32_t)p;
                edges[2 * ne + 1] = c.succ[p].v[k];
            }
            ne++;
        }
        free(c.succ[p].v);
    }
    free(c.succ);
    flat_free(&f);
    return (int64_t)ne;
}

int64_t xko_resolve(uint64_t nacc, const uint32_t *task, const uint8_t *mode, const uint8_t *nrect,
                    const uint64_t *rects, uint64_t edge_cap, uint32_t *edges, uint64_t *frag)
{
    return resolve_impl(nacc, task, mode, nrect, rects, 8, nrect, rects, 8, edge_cap, edges, frag);
}

/*
 * Same, for streams with up to 3 rects per access (12 u64 each) that the tree
 * inserts and intersects, and separately the rects (marks, 12 u64 each) that
 * on_insert tests a node against (dependency-tree.hpp:241-256 iterates
 * access->rects()).  For a tile both are its rects; for a segment seen by a
 * BLAS domain the inserted rects are interval_to_rects in the domain's frame
 * (dependency-tree.hpp:358-364) and the marks its own rects built with
 * ld = SIZE_MAX (access.hpp:495-513).
 */
int64_t xko_resolve_marks(uint64_t nacc, const uint32_t *task, const uint8_t *mode, const uint8_t *nrect,
                          const uint64_t *rects, const uint8_t *nmark, const uint64_t *marks, uint64_t edge_cap,
                          uint32_t *edges)
{
    return resolve_impl(nacc, task, mode, nrect, rects, 12, nmark, marks, 12, edge_cap, edges, NULL);
}

/* Fragment list of a rect stream under plain inserts (no state): for tests. */
static void nop_insert(void *pc, uint64_t x0, uint64_t x1, void *ctx) { (void)pc; (void)x0; (void)x1; (void)ctx; }

typedef struct {
    uint64_t *out;
    uint64_t cap, n;
} lctx_t;

static void list_cb(void *pc, uint64_t x0, uint64_t x1, void *ctx)
{
    lctx_t *l = ctx;
    if (l->n < l->cap) {
        uint64_t *o = l->out + 4 * l->n;
        o[0] = x0; o[1] = x1; o[2] = PY0(pc); o[3] = PY1(pc);
    }
    l->n++;
}

/* ------------------------------------------------------------------ */
/* coherency controller (memory-tree.hpp), one (ld, sizeof_type) tree   */
/* ------------------------------------------------------------------ */

#define NDEV 4
#define VIEWS_MAX 8 /* memory-tree.hpp:86 MEMORY_REPLICATE_ALLOCATION_VIEWS_MAX */

typedef struct {
    uint8_t nalloc;
    uint8_t coh;    /* per-view coherency bits */
    uint8_t fetch;  /* per-view fetching bits */
    int32_t chunk[VIEWS_MAX];
} replica_t;

typedef struct {
    uint64_t y0, y1;
    uint16_t coh, fetch;  /* per-device bits (memory-tree.hpp:256-260) */
    replica_t rep[NDEV];
} cpiece_t;

static void cp_copy(void *dst, const void *src) { memcpy(dst, src, sizeof(cpiece_t)); }
static void cp_init(void *p) { memset(p, 0, sizeof(cpiece_t)); }
static void cp_fini(void *p) { (void)p; }
static const ptype_t CPT = {sizeof(cpiece_t), cp_copy, cp_init, cp_fini};

typedef struct {
    cpiece_t *p;
    uint64_t x0, x1;
    int dst_view;
    int must_fetch;
    int src_dev;
    int64_t src_chunk;
} partite_t;

typedef struct {
    partite_t *v;
    uint32_t n, c;
} pvec_t;

static void part_cb(void *pc, uint64_t x0, uint64_t x1, void *ctx)
{
    pvec_t *v = ctx;
    if (v->n == v->c) {
        v->c = v->c ? 2 * v->c : 16;
        v->v = xrealloc(v->v, sizeof(partite_t) * v->c);
    }
    partite_t *p = v->v + v->n++;
    memset(p, 0, sizeof(*p));
    p->p = pc;
    p->x0 = x0; p->x1 = x1;
    p->dst_view = -1;
    p->src_chunk = -1;
}

/*
 * Coherency fetch planning for one (ld, sizeof_type) tree with ndev devices
 * (0 = host), synchronous completion: every fetch of an access completes
 * (fetch_callback, memory-tree.hpp:955-1061) before the next access.  Device
 * memory is unlimited (no eviction).  Per access:
 *   parts  : 6 int64 per partite (x0, x1, y0, y1, coh_before, fetch_before)
 *   fetches: 8 int64 per fetch  (x0, x1, y0, y1, src_dev, src_chunk, dst_dev, dst_chunk)
 * chunks are numbered in allocation order from 0; -1 = host memory.
 * Returns 0, or -1 when an access would add a 9th allocation view to a block
 * (memory-tree.hpp:1362-1365 LOGGER_FATAL); *fail_at = its index and the
 * offsets stop there.
 */
typedef struct {
    uint64_t x0, x1, y0, y1;
    uint64_t bytes[2];
} octx_t;

/* SEARCH_OWNERS (memory-tree.hpp:2130-2139): bytes of the query inside the
 * fragment, added to every device whose coherency bit the fragment holds */
static void own_cb(void *pc, uint64_t x0, uint64_t x1, void *ctx)
{
    octx_t *o = ctx;
    const cpiece_t *p = pc;
    const uint64_t ax0 = x0 > o->x0 ? x0 : o->x0, ax1 = x1 < o->x1 ? x1 : o->x1;
    const uint64_t ay0 = p->y0 > o->y0 ? p->y0 : o->y0, ay1 = p->y1 < o->y1 ? p->y1 : o->y1;
    if (ax0 >= ax1 || ay0 >= ay1) return;
    for (int d = 0; d < 2; ++d)
        if (p->coh & (1u << d)) o->bytes[d] += (ax1 - ax0) * (ay1 - ay0);
}

/* who_owns (memory-tree.hpp:1903-1933): the devices owning the most bytes of
 * the query rects, ties included, 0 when nobody owns a byte */
static uint32_t who_owns(flat_t *f, const uint64_t *q, int nr)
{
    octx_t o = {0, 0, 0, 0, {0, 0}};
    for (int r = 0; r < nr; ++r) {
        o.x0 = q[4 * r]; o.x1 = q[4 * r + 1]; o.y0 = q[4 * r + 2]; o.y1 = q[4 * r + 3];
        flat_intersect(f, o.x0, o.x1, o.y0, o.y1, own_cb, &o);
    }
    uint32_t owners = 0;
    uint64_t best = 0;
    for (int d = 0; d < 2; ++d) {
        if (best < o.bytes[d]) {
            best = o.bytes[d];
            owners = 1u << d;
        } else if (best && best == o.bytes[d]) {
            owners |= 1u << d;
        }
    }
    return owners;
}

/* queries answered on the state after the stream (or at the abort):
 * nq rects of 8 u64 (two rects, 4 u64 each) with qnrect[q] rects each */
int64_t xko_coherency(uint64_t nacc, const uint8_t *dev, const uint8_t *mode, const uint8_t *nrect,
                      const uint64_t *rects, const uint64_t *chunk_bytes, uint64_t part_cap, int64_t *part_off,
                      int64_t *parts, uint64_t fetch_cap, int64_t *fetch_off, int64_t *fetches,
                      int64_t *fail_at, uint64_t nq, const uint8_t *qnrect, const uint64_t *qrects, uint32_t *owners)
{
    flat_t f = {&CPT, NULL, 0, 0};
    pvec_t pv = {NULL, 0, 0};
    uint64_t np_ = 0, nf = 0;
    int64_t nchunks = 0;
    uint64_t *chunk_size = NULL;
    uint64_t chunk_cap = 0;
    int64_t rc = 0;
    *fail_at = -1;
    uint64_t a;
    for (a = 0; a < nacc; ++a) {
        part_off[a] = (int64_t)np_;
        fetch_off[a] = (int64_t)nf;
        const int d = dev[a];
        const uint8_t m = mode[a];
        const uint16_t devbit = (uint16_t)(1u << d);
        if (d == 0 && f.ns == 0) continue;  /* fetch_list_to_host: empty tree -> nothing (1191-1192) */

        /* (1) insert, (2) partition */
        for (int r = 0; r < nrect[a]; ++r) {
            const uint64_t *q = rects + 8 * a + 4 * r;
            flat_insert(&f, q[0], q[1], q[2], q[3], nop_insert, NULL);
        }
        pv.n = 0;
        for (int r = 0; r < nrect[a]; ++r) {
            const uint64_t *q = rects + 8 * a + 4 * r;
            flat_intersect(&f, q[0], q[1], q[2], q[3], part_cb, &pv);
        }
        for (uint32_t i = 0; i < pv.n; ++i) {
            partite_t *p = pv.v + i;
            if (np_ < part_cap) {
                int64_t *o = parts + 6 * np_;
                o[0] = (int64_t)p->x0; o[1] = (int64_t)p->x1; o[2] = (int64_t)p->p->y0; o[3] = (int64_t)p->p->y1;
                o[4] = p->p->coh; o[5] = p->p->fetch;
            }
            np_++;
        }

        int64_t chunk = -1;
        if (d == 0) {
            /* fetch_list_to_host_setup_partition (1110-1181) */
            for (uint32_t i = 0; i < pv.n; ++i) {
                partite_t *p = pv.v + i;
                cpiece_t *b = p->p;
                if (b->coh == 0 || (b->coh & devbit) || (b->fetch & devbit)) { p->must_fetch = 0; continue; }
                b->fetch |= devbit;
                int src = __builtin_ffs(b->coh) - 1;
                replica_t *sr = b->rep + src;
                int sv = 0;
                for (int v = 1; v < sr->nalloc; ++v)
                    if (chunk_size[sr->chunk[sv]] < chunk_size[sr->chunk[v]]) sv = v;
                p->must_fetch = 1;
                p->src_dev = src;
                p->src_chunk = sr->chunk[sv];
            }
        } else {
            /* (3) fetch_access_find_allocation (1427-1450) */
            int found = 0;
            if (pv.n) {
                replica_t *r0 = pv.v[0].p->rep + d;
                for (int j = 0; j < r0->nalloc && !found; ++j) {
                    int32_t c = r0->chunk[j];
                    int ok = 1;
                    for (uint32_t i = 1; i < pv.n && ok; ++i) {
                        replica_t *ri = pv.v[i].p->rep + d;
                        int hit = -1;
                        for (int k = 0; k < ri->nalloc; ++k) if (ri->chunk[k] == c) { hit = k; break; }
                        if (hit < 0) ok = 0; else pv.v[i].dst_view = hit;
                    }
                    if (ok) { found = 1; chunk = c; pv.v[0].dst_view = j; }
                }
            }
            if (!found) {
                /* fetch_access_allocate + fetch_access_create_allocation_views (1293-1372) */
                if ((uint64_t)nchunks == chunk_cap) {
                    chunk_cap = chunk_cap ? 2 * chunk_cap : 1024;
                    chunk_size = xrealloc(chunk_size, sizeof(uint64_t) * chunk_cap);
                }
                chunk = nchunks++;
                chunk_size[chunk] = chunk_bytes[a];
                for (uint32_t i = 0; i < pv.n; ++i) {
                    replica_t *ri = pv.v[i].p->rep + d;
                    if (ri->nalloc >= VIEWS_MAX) { rc = -1; *fail_at = (int64_t)a; goto done; }
                    ri->chunk[ri->nalloc] = (int32_t)chunk;
                    pv.v[i].dst_view = ri->nalloc++;
                }
            }
            /* (5) fetch_access_setup_copies (1471-1631), read modes only */
            if (m & M_R) {
                for (uint32_t i = 0; i < pv.n; ++i) {
                    partite_t *p = pv.v + i;
                    cpiece_t *b = p->p;
                    replica_t *dr = b->rep + d;
                    const uint8_t ab = (uint8_t)(1u << p->dst_view);
                    if (dr->coh & ab) { p->must_fetch = 0; continue; }
                    if (dr->fetch & ab) { p->must_fetch = 0; continue; } /* awaiting: not reached synchronously */
                    if (b->coh & ~1u) {
                        /* RouterAffinity::get_source: dst itself when valid there */
                        int src = (b->coh & devbit) ? d : __builtin_ffs(b->coh & ~1u) - 1;
                        replica_t *sr = b->rep + src;
                        int sv = __builtin_ffs(sr->coh) - 1;
                        p->must_fetch = 1;
                        p->src_dev = src;
                        p->src_chunk = sr->chunk[sv];
                    } else if (b->fetch & ~1u) {
                        p->must_fetch = 0; /* D2D forward: needs a concurrent fetch, not reached synchronously */
                    } else {
                        p->must_fetch = 1;
                        p->src_dev = 0;
                        p->src_chunk = -1;
                    }
                    dr->fetch |= ab;
                    b->fetch |= devbit;
                }
            }
            /* (6) fetch_access_set_coherent (1633-1667) */
            if (m & M_W) {
                for (uint32_t i = 0; i < pv.n; ++i) {
                    partite_t *p = pv.v + i;
                    cpiece_t *b = p->p;
                    for (int dd = 0; dd < NDEV; ++dd) b->rep[dd].coh = 0;
                    b->coh = 0;
                    b->rep[d].coh = (uint8_t)(1u << p->dst_view);
                    b->coh = devbit;
                }
            }
            if (!(m & M_R))
                for (uint32_t i = 0; i < pv.n; ++i) pv.v[i].must_fetch = 0;
        }

        /* (7) fetch list, then synchronous completion (SEARCH_FETCHED, 2067-2127) */
        for (uint32_t i = 0; i < pv.n; ++i) {
            partite_t *p = pv.v + i;
            if (!p->must_fetch) continue;
            if (nf < fetch_cap) {
                int64_t *o = fetches + 8 * nf;
                o[0] = (int64_t)p->x0; o[1] = (int64_t)p->x1; o[2] = (int64_t)p->p->y0; o[3] = (int64_t)p->p->y1;
                o[4] = p->src_dev; o[5] = p->src_chunk; o[6] = d; o[7] = d == 0 ? -1 : chunk;
            }
            nf++;
        }
        for (uint32_t i = 0; i < pv.n; ++i) {
            partite_t *p = pv.v + i;
            if (!p->must_fetch) continue;
            cpiece_t *b = p->p;
            replica_t *rr = b->rep + d;
            if (d != 0) {
                for (int v = 0; v < rr->nalloc; ++v)
                    if (rr->chunk[v] == chunk) {
                        rr->coh |= (uint8_t)(1u << v);
                        rr->fetch &= (uint8_t)~(1u << v);
                        break;
                    }
            }
            b->coh |= devbit;
            if (rr->fetch == 0) b->fetch &= (uint16_t)~devbit;
        }
    }
done:
    if (rc == 0) {
        part_off[nacc] = (int64_t)np_;
        fetch_off[nacc] = (int64_t)nf;
    }
    for (uint64_t q = 0; q < nq; ++q) owners[q] = who_owns(&f, qrects + 8 * q, qnrect[q]);
    /* on failure part_off[a] / fetch_off[a] already hold the start of the
     * failing access: its partites are not part of the plan */
    free(pv.v);
    free(chunk_size);
    flat_free(&f);
    return rc;
}

/*
 * Plain partition (no state): for each access, insert its rects (skipping
 * host accesses while the tree is empty when dev != NULL) and list the
 * fragments intersecting them, 4 u64 each.  Mirrors oracle/_ref's
 * khpref_partition for the geometry check.
 */
int64_t xko_partition(uint64_t nacc, const uint8_t *dev, const uint8_t *nrect, const uint64_t *rects,
                      uint64_t part_cap, uint64_t *part_off, uint64_t *parts)
{
    flat_t f = {&CPT, NULL, 0, 0};
    lctx_t l = {parts, part_cap, 0};
    for (uint64_t a = 0; a < nacc; ++a) {
        part_off[a] = l.n;
        if (dev && dev[a] == 0 && f.ns == 0) continue;
        for (int r = 0; r < nrect[a]; ++r) {
            const uint64_t *q = rects + 8 * a + 4 * r;
            flat_insert(&f, q[0], q[1], q[2], q[3], nop_insert, NULL);
        }
        for (int r = 0; r < nrect[a]; ++r) {
            const uint64_t *q = rects + 8 * a + 4 * r;
            flat_intersect(&f, q[0], q[1], q[2], q[3], list_cb, &l);
        }
    }
    part_off[nacc] = l.n;
    uint64_t total = flat_count(&f);
    (void)total;
    flat_free(&f);
    return (int64_t)l.n;
}

/* ------------------------------------------------------------------ */
/* coherency with fetches in flight (asynchronous completion)           */
/* ------------------------------------------------------------------ */

/*
 * The same tree, but a fetch completes only when an event says so
 * (fetch_callback, memory-tree.hpp:955-1061), so later accesses see the
 * fetching bits that setup_copies set (1471-1631):
 *   - a GPU read of a view being fetched awaits that fetch: no copy, the
 *     access is registered on the view (1516-1527);
 *   - a GPU read of data the GPU does not hold coherent, while some view of
 *     it is being fetched there, registers a D2D forward on the first
 *     fetching view (1567-1600, USE_D2D_FORWARDING = 1 at line 85); the
 *     forward copy is issued when that view's fetch completes;
 *   - a host read skips data the host is already fetching (1126-1130).
 * Completion of a fetch of chunk c to device d over rect R (SEARCH_FETCHED,
 * 2067-2127): every piece inside R gets, on the view holding c, coherency set
 * and fetching cleared, then device coherency set and device fetching cleared
 * once none of its views is fetching; the accesses awaiting on that view are
 * released and its forwards issued.  Registrations live on the pieces and
 * are copied when a piece splits; one fires once, at the first completion
 * that reaches a piece holding it (the reference keeps it on the shrunk
 * original node, which the same completion reaches).
 *
 * Events: ev_type[k] = 0 plans the next access; ev_type[k] = 1 completes the
 * in-flight fetch of rank ev_arg[k] mod (fetches in flight), in id order
 * (nothing when none is in flight).  Fetch ids count every fetch issued, plan
 * fetches and forwards, from 0.  Registrations fired by one completion are
 * handled in registration order, and the forwards among them take fetch ids in
 * that order.
 *
 * Output: log rows of 12 int64 [event, kind, access, x0, x1, y0, y1, a, b, c, d, id]:
 *   kind 0 partite of a planned access: a = coherency before, b = fetching before
 *   kind 1 fetch issued by a plan:      a = src dev, b = src chunk, c = dst dev, d = dst chunk, id = fetch id
 *   kind 2 access awaiting a fetch:     a = chunk awaited, id = registration
 *   kind 3 forward registered:          a = src chunk (fetching view), b = dst chunk, id = registration
 *   kind 4 awaiting access released:    a = chunk, id = registration
 *   kind 5 forward issued:              a = 1, b = src chunk, c = 1, d = dst chunk, id = fetch id
 * Returns the number of log rows (may exceed log_cap).  At a ninth
 * allocation view *fail_at = that access and the log stops before it.
 */
typedef struct {
    uint64_t y0, y1;
    uint16_t coh, fetch;
    replica_t rep[NDEV];
    uint32_t *reg;       /* registrations held by this piece */
    uint32_t nreg, creg;
} apiece_t;

static void ap_copy(void *dst, const void *src)
{
    const apiece_t *s = src;
    apiece_t *d = dst;
    *d = *s;
    d->creg = s->nreg;
    d->reg = s->nreg ? xmalloc(sizeof(uint32_t) * s->nreg) : NULL;
    if (s->nreg) memcpy(d->reg, s->reg, sizeof(uint32_t) * s->nreg);
}
static void ap_init(void *p) { memset(p, 0, sizeof(apiece_t)); }
static void ap_fini(void *p) { free(((apiece_t *)p)->reg); }
static const ptype_t APT = {sizeof(apiece_t), ap_copy, ap_init, ap_fini};

static void ap_addreg(apiece_t *b, uint32_t r)
{
    if (b->nreg == b->creg) {
        b->creg = b->creg ? 2 * b->creg : 4;
        b->reg = xrealloc(b->reg, sizeof(uint32_t) * b->creg);
    }
    b->reg[b->nreg++] = r;
}

typedef struct {
    apiece_t *p;
    uint64_t x0, x1;
    int dst_view, action, src_dev;
    int64_t src_chunk;
} apart_t;

typedef struct {
    apart_t *v;
    uint32_t n, c;
} apvec_t;

static void apart_cb(void *pc, uint64_t x0, uint64_t x1, void *ctx)
{
    apvec_t *v = ctx;
    if (v->n == v->c) {
        v->c = v->c ? 2 * v->c : 16;
        v->v = xrealloc(v->v, sizeof(apart_t) * v->c);
    }
    apart_t *p = v->v + v->n++;
    memset(p, 0, sizeof(*p));
    p->p = pc;
    p->x0 = x0; p->x1 = x1;
    p->dst_view = -1;
    p->src_chunk = -1;
}

typedef struct {
    uint64_t x0, x1, y0, y1;
    int dst_dev;
    int64_t dst_chunk;
    int done;
} afetch_t;

typedef struct {
    uint32_t acc;
    int kind;  /* 2 await, 3 forward */
    uint64_t x0, x1, y0, y1;
    int64_t chunk, dst_chunk;
    int fired;
} areg_t;

typedef struct {
    int64_t *rows;
    uint64_t cap, n;
} alog_t;

static void alog(alog_t *l, int64_t ev, int64_t kind, int64_t acc, uint64_t x0, uint64_t x1, uint64_t y0, uint64_t y1,
                 int64_t a, int64_t b, int64_t c, int64_t d, int64_t id)
{
    if (l->n < l->cap) {
        int64_t *o = l->rows + 12 * l->n;
        o[0] = ev; o[1] = kind; o[2] = acc;
        o[3] = (int64_t)x0; o[4] = (int64_t)x1; o[5] = (int64_t)y0; o[6] = (int64_t)y1;
        o[7] = a; o[8] = b; o[9] = c; o[10] = d; o[11] = id;
    }
    l->n++;
}

typedef struct {
    const afetch_t *f;
    areg_t *regs;
    u32vec_t hit;  /* registrations fired by this completion */
} acomp_t;

/* SEARCH_FETCHED on one piece inside the completed fetch's rect (memory-tree.hpp:2067-2127) */
static void acomplete_cb(void *pc, uint64_t x0, uint64_t x1, void *ctx)
{
    (void)x0; (void)x1;
    acomp_t *c = ctx;
    apiece_t *b = pc;
    const int d = c->f->dst_dev;
    const uint16_t devbit = (uint16_t)(1u << d);
    replica_t *rr = b->rep + d;
    if (d != 0) {
        for (int v = 0; v < rr->nalloc; ++v)
            if (rr->chunk[v] == c->f->dst_chunk) {
                rr->coh |= (uint8_t)(1u << v);
                rr->fetch &= (uint8_t)~(1u << v);
                break;
            }
        /* the view's awaiting accesses and forwards move to this completion */
        uint32_t k = 0;
        for (uint32_t i = 0; i < b->nreg; ++i) {
            areg_t *r = c->regs + b->reg[i];
            if (r->chunk == c->f->dst_chunk) {
                if (!r->fired) {
                    r->fired = 1;
                    vpush(&c->hit, b->reg[i]);
                }
            } else {
                b->reg[k++] = b->reg[i];
            }
        }
        b->nreg = k;
    }
    b->coh |= devbit;
    if (rr->fetch == 0) b->fetch &= (uint16_t)~devbit;
}

static int u32cmp(const void *a, const void *b)
{
    const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

int64_t xko_coherency_events(uint64_t nacc, const uint8_t *dev, const uint8_t *mode, const uint8_t *nrect,
                             const uint64_t *rects, const uint64_t *chunk_bytes, uint64_t nev, const uint8_t *ev_type,
                             const uint64_t *ev_arg, uint64_t log_cap, int64_t *log_rows, int64_t *fail_at)
{
    flat_t f = {&APT, NULL, 0, 0};
    apvec_t pv = {NULL, 0, 0};
    alog_t lg = {log_rows, log_cap, 0};
    afetch_t *fe = NULL;
    uint64_t nfe = 0, cfe = 0;
    areg_t *rg = NULL;
    uint64_t nrg = 0, crg = 0;
    int64_t nchunks = 0;
    uint64_t *chunk_size = NULL, chunk_cap = 0;
    int64_t rc = 0;
    uint64_t a = 0;
    *fail_at = -1;
#define NEW_FETCH(X0, X1, Y0, Y1, DD, DC)                                                    \
    do {                                                                                      \
        if (nfe == cfe) { cfe = cfe ? 2 * cfe : 256; fe = xrealloc(fe, sizeof(afetch_t) * cfe); } \
        fe[nfe].x0 = (X0); fe[nfe].x1 = (X1); fe[nfe].y0 = (Y0); fe[nfe].y1 = (Y1);            \
        fe[nfe].dst_dev = (DD); fe[nfe].dst_chunk = (DC); fe[nfe].done = 0; nfe++;            \
    } while (0)
#define NEW_REG(ACC, KIND, P, CH, DCH)                                                      \
    do {                                                                                      \
        if (nrg == crg) { crg = crg ? 2 * crg : 256; rg = xrealloc(rg, sizeof(areg_t) * crg); } \
        rg[nrg].acc = (uint32_t)(ACC); rg[nrg].kind = (KIND);                                 \
        rg[nrg].x0 = (P)->x0; rg[nrg].x1 = (P)->x1; rg[nrg].y0 = (P)->p->y0; rg[nrg].y1 = (P)->p->y1; \
        rg[nrg].chunk = (CH); rg[nrg].dst_chunk = (DCH); rg[nrg].fired = 0;                   \
        ap_addreg((P)->p, (uint32_t)nrg); nrg++;                                              \
    } while (0)

    for (uint64_t ev = 0; ev < nev; ++ev) {
        if (ev_type[ev] == 1) {
            /* complete the in-flight fetch of rank ev_arg mod in-flight count */
            uint64_t nin = 0;
            for (uint64_t k = 0; k < nfe; ++k) nin += !fe[k].done;
            if (!nin) continue;
            uint64_t want = ev_arg[ev] % nin, id = 0;
            for (id = 0; id < nfe; ++id)
                if (!fe[id].done && want-- == 0) break;
            fe[id].done = 1;
            acomp_t c = {fe + id, rg, {NULL, 0, 0}};
            const afetch_t cf = fe[id];
            c.f = &cf;
            flat_intersect(&f, cf.x0, cf.x1, cf.y0, cf.y1, acomplete_cb, &c);
            qsort(c.hit.v, c.hit.n, sizeof(uint32_t), u32cmp);
            for (uint32_t i = 0; i < c.hit.n; ++i) {
                const areg_t *r = rg + c.hit.v[i];
                if (r->kind == 2) {
                    alog(&lg, (int64_t)ev, 4, r->acc, r->x0, r->x1, r->y0, r->y1, r->chunk, 0, 0, 0, c.hit.v[i]);
                } else {
                    /* the forward copy: from the view just fetched to the forward's view */
                    alog(&lg, (int64_t)ev, 5, r->acc, r->x0, r->x1, r->y0, r->y1, 1, r->chunk, 1, r->dst_chunk,
                         (int64_t)nfe);
                    NEW_FETCH(r->x0, r->x1, r->y0, r->y1, 1, r->dst_chunk);
                }
            }
            free(c.hit.v);
            continue;
        }
        if (a >= nacc) continue;
        const int d = dev[a];
        const uint8_t m = mode[a];
        const uint16_t devbit = (uint16_t)(1u << d);
        if (d == 0 && f.ns == 0) { a++; continue; }  /* fetch_list_to_host: empty tree (1191-1192) */
        for (int r = 0; r < nrect[a]; ++r) {
            const uint64_t *q = rects + 8 * a + 4 * r;
            flat_insert(&f, q[0], q[1], q[2], q[3], nop_insert, NULL);
        }
        pv.n = 0;
        for (int r = 0; r < nrect[a]; ++r) {
            const uint64_t *q = rects + 8 * a + 4 * r;
            flat_intersect(&f, q[0], q[1], q[2], q[3], apart_cb, &pv);
        }
        for (uint32_t i = 0; i < pv.n; ++i) {
            const apart_t *p = pv.v + i;
            alog(&lg, (int64_t)ev, 0, (int64_t)a, p->x0, p->x1, p->p->y0, p->p->y1, p->p->coh, p->p->fetch, 0, 0, 0);
        }
        int64_t chunk = -1;
        if (d == 0) {
            /* fetch_list_to_host_setup_partition (1110-1181) */
            for (uint32_t i = 0; i < pv.n; ++i) {
                apart_t *p = pv.v + i;
                apiece_t *b = p->p;
                if (b->coh == 0 || (b->coh & devbit) || (b->fetch & devbit)) continue;
                b->fetch |= devbit;
                const int src = __builtin_ffs(b->coh) - 1;
                replica_t *sr = b->rep + src;
                int sv = 0;
                for (int v = 1; v < sr->nalloc; ++v)
                    if (chunk_size[sr->chunk[sv]] < chunk_size[sr->chunk[v]]) sv = v;
                p->action = 1;
                p->src_dev = src;
                p->src_chunk = sr->chunk[sv];
            }
        } else {
            /* fetch_access_find_allocation (1427-1450) */
            int found = 0;
            if (pv.n) {
                replica_t *r0 = pv.v[0].p->rep + d;
                for (int j = 0; j < r0->nalloc && !found; ++j) {
                    const int32_t c = r0->chunk[j];
                    int ok = 1;
                    for (uint32_t i = 1; i < pv.n && ok; ++i) {
                        replica_t *ri = pv.v[i].p->rep + d;
                        int hit = -1;
                        for (int k = 0; k < ri->nalloc; ++k)
                            if (ri->chunk[k] == c) { hit = k; break; }
                        if (hit < 0) ok = 0; else pv.v[i].dst_view = hit;
                    }
                    if (ok) { found = 1; chunk = c; pv.v[0].dst_view = j; }
                }
            }
            if (!found) {
                /* fetch_access_allocate + create_allocation_views (1293-1372) */
                if ((uint64_t)nchunks == chunk_cap) {
                    chunk_cap = chunk_cap ? 2 * chunk_cap : 1024;
                    chunk_size = xrealloc(chunk_size, sizeof(uint64_t) * chunk_cap);
                }
                chunk = nchunks++;
                chunk_size[chunk] = chunk_bytes[a];
                for (uint32_t i = 0; i < pv.n; ++i) {
                    replica_t *ri = pv.v[i].p->rep + d;
                    if (ri->nalloc >= VIEWS_MAX) { rc = -1; *fail_at = (int64_t)a; goto done; }
                    ri->chunk[ri->nalloc] = (int32_t)chunk;
                    pv.v[i].dst_view = ri->nalloc++;
                }
            }
            /* fetch_access_setup_copies (1471-1631) */
            if (m & M_R) {
                for (uint32_t i = 0; i < pv.n; ++i) {
                    apart_t *p = pv.v + i;
                    apiece_t *b = p->p;
                    replica_t *dr = b->rep + d;
                    const uint8_t ab = (uint8_t)(1u << p->dst_view);
                    if (dr->coh & ab) continue;
                    if (dr->fetch & ab) {  /* awaiting (1516-1527) */
                        alog(&lg, (int64_t)ev, 2, (int64_t)a, p->x0, p->x1, b->y0, b->y1, chunk, 0, 0, 0, (int64_t)nrg);
                        NEW_REG(a, 2, p, chunk, -1);
                        continue;
                    }
                    if (b->coh & ~1u) {
                        /* RouterAffinity::get_source: dst itself when valid there */
                        const int src = (b->coh & devbit) ? d : __builtin_ffs(b->coh & ~1u) - 1;
                        replica_t *sr = b->rep + src;
                        const int sv = __builtin_ffs(sr->coh) - 1;
                        p->action = 1;
                        p->src_dev = src;
                        p->src_chunk = sr->chunk[sv];
                    } else if (b->fetch & ~1u) {
                        /* D2D forward from the first fetching view of the fetching device (1567-1600) */
                        const int fd = (b->fetch & devbit) ? d : __builtin_ffs(b->fetch & ~1u) - 1;
                        replica_t *fr = b->rep + fd;
                        const int fv = __builtin_ffs(fr->fetch) - 1;
                        alog(&lg, (int64_t)ev, 3, (int64_t)a, p->x0, p->x1, b->y0, b->y1, fr->chunk[fv], chunk, 0, 0,
                             (int64_t)nrg);
                        NEW_REG(a, 3, p, fr->chunk[fv], chunk);
                    } else {
                        p->action = 1;
                        p->src_dev = 0;
                        p->src_chunk = -1;
                    }
                    dr->fetch |= ab;
                    b->fetch |= devbit;
                }
            }
            /* fetch_access_set_coherent (1633-1667) */
            if (m & M_W) {
                for (uint32_t i = 0; i < pv.n; ++i) {
                    apiece_t *b = pv.v[i].p;
                    for (int dd = 0; dd < NDEV; ++dd) b->rep[dd].coh = 0;
                    b->rep[d].coh = (uint8_t)(1u << pv.v[i].dst_view);
                    b->coh = devbit;
                }
            }
        }
        /* fetch_list_from_partition (1067-1108): the copies are issued; they complete by events */
        for (uint32_t i = 0; i < pv.n; ++i) {
            const apart_t *p = pv.v + i;
            if (p->action != 1) continue;
            alog(&lg, (int64_t)ev, 1, (int64_t)a, p->x0, p->x1, p->p->y0, p->p->y1, p->src_dev, p->src_chunk, d,
                 d == 0 ? -1 : chunk, (int64_t)nfe);
            NEW_FETCH(p->x0, p->x1, p->p->y0, p->p->y1, d, d == 0 ? -1 : chunk);
        }
        a++;
    }
done:
#undef NEW_FETCH
#undef NEW_REG
    free(pv.v);
    free(fe);
    free(rg);
    free(chunk_size);
    flat_free(&f);
    (void)rc;  /* an abort is reported through *fail_at; the log holds what came before it */
    return (int64_t)lg.n;
}

/* ------------------------------------------------------------------------
 * The runtime's successor release on one host core, the CPU side of the
 * device release engine's measurement: every edge applied as
 * __task_precedes does (include/xkrt/task/task.hpp:453-486: the successor's
 * wait counter + 1, the successor on the predecessor access's list), every
 * task committed (task.hpp:535-551: the guard of 1 dropped), then the ready
 * tasks run in FIFO order as empty tasks, each completion walking its
 * accesses' lists (src/task.cc:160-247; detached accesses, mode 8, skipped).
 * acc_task: task of each access (tasks 0 .. ntask - 1); edges: (pred, succ)
 * access pairs.  order (nullable): completion order.  Returns the number of
 * tasks that ran.
 * ------------------------------------------------------------------------ */
int64_t xko_release_run(uint64_t ntask, uint64_t nacc, const uint32_t *acc_task, const uint8_t *acc_mode,
                        uint64_t ne, const uint32_t *edges, uint32_t *order)
{
    uint32_t *wc = (uint32_t *)malloc(4 * (ntask + 1));
    uint64_t *first = (uint64_t *)calloc(ntask + 2, 8);  /* accesses of task t: [first[t], first[t + 1]) */
    uint64_t *head = (uint64_t *)malloc(8 * (nacc + 1));
    uint64_t *next = (uint64_t *)malloc(8 * (ne + 1));
    uint32_t *succ = (uint32_t *)malloc(4 * (ne + 1));
    uint32_t *queue = (uint32_t *)malloc(4 * (ntask + 1));
    if (!wc || !first || !head || !next || !succ || !queue) return -1;
    for (uint64_t t = 0; t < ntask; ++t) wc[t] = 1;
    for (uint64_t a = 0; a < nacc; ++a) {
        head[a] = UINT64_MAX;
        first[acc_task[a] + 1] = a + 1;
    }
    for (uint64_t t = 1; t <= ntask; ++t)
        if (first[t] < first[t - 1]) first[t] = first[t - 1];
    for (uint64_t e = 0; e < ne; ++e) {
        const uint32_t p = edges[2 * e], s = edges[2 * e + 1];
        wc[acc_task[s]]++;
        succ[e] = acc_task[s];
        next[e] = head[p];
        head[p] = e;
    }
    uint64_t qh = 0, qt = 0;
    for (uint64_t t = 0; t < ntask; ++t)
        if (--wc[t] == 0) queue[qt++] = (uint32_t)t;
    while (qh < qt) {
        const uint32_t t = queue[qh++];
        if (order) order[qh - 1] = t;
        for (uint64_t a = first[t]; a < first[t + 1]; ++a) {
            if (acc_mode[a] & 8) continue;
            for (uint64_t e = head[a]; e != UINT64_MAX; e = next[e])
                if (--wc[succ[e]] == 0) queue[qt++] = succ[e];
        }
    }
    free(wc); free(first); free(head); free(next); free(succ); free(queue);
    return (int64_t)qh;
}
