/*
 * xkrt_hip.h -- C ABI of the MI355X (gfx950) dependency resolver and
 * coherency fetch planner.
 *
 * Plain C types only: pointers, sizes, POD structs.  Every entry point
 * returns 0 on success or a negative errno:
 *   -EINVAL     bad argument (null handle, zero ld, non-contiguous tasks, ...)
 *   -ENODEV     no HIP device / the HIP runtime failed
 *   -ENOMEM     device memory budget or allocation exhausted
 *   -ENOSPC     a per-bin capacity (fragments, readers, records) was exceeded:
 *               the domain failed (reset or destroy it); the other domains of
 *               its pool are not affected
 *   -EOVERFLOW  the pool's edge / candidate buffers overflowed; every domain of
 *               the pool is poisoned
 *   -ERANGE     a coordinate does not fit the domain's 32-bit frame, or a
 *               device record leaves its domain's grid (that domain fails)
 *   -ECANCELED  a batch for a domain that failed earlier (reset it first)
 *   -ENOTSUP    an access the device path does not reproduce bit-exactly
 *               (a wrapped 2-rect tile that the reference would mark beyond
 *               its rects, see xkrt_hip_dep_submit)
 *
 * Reference interfaces replaced (rpereira-dev/xkaapi):
 *   DependencyDomain::link / put           include/xkrt/memory/access/dependency-domain.hpp:48-80
 *   KBLASDependencyTree<2>::link / put      include/xkrt/memory/access/blas/dependency-tree.hpp:367-414
 *   task_get_dependency_domain_blas_matrix  src/task/task.cc:153-201  (one domain per (ld, sizeof_type))
 *   __access_precedes (edge dedup)          include/xkrt/task/task.hpp:494-512
 *   KBLASMemoryTree<2>::fetch               include/xkrt/memory/access/blas/memory-tree.hpp:1806-1874
 *   task_get_memory_controller              src/task/task.cc:50-150   (one tree per (ld, sizeof_type))
 */
#ifndef XKRT_HIP_H
#define XKRT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* access modes, identical to xkrt_access_mode_t (include/xkrt/memory/access/mode.h:56-62) */
enum {
    XKRT_HIP_MODE_R  = 1,
    XKRT_HIP_MODE_W  = 2,
    XKRT_HIP_MODE_RW = 3,
    XKRT_HIP_MODE_V  = 4,
    XKRT_HIP_MODE_D  = 8
};

/* access kinds */
enum {
    XKRT_HIP_KIND_TILE    = 0,  /* column-major tile: addr, m rows, n columns */
    XKRT_HIP_KIND_SEGMENT = 1,  /* byte segment [addr, addr + m) seen by a BLAS domain */
    XKRT_HIP_KIND_HANDLE  = 2   /* task domains only: an opaque handle (addr), see xkrt_hip_taskdom_submit */
};

/* One access.  A tile is the region of the reference's BLAS access
 * constructor (include/xkrt/memory/access/access.hpp:539-574) with the
 * offsets folded into addr; ld and sizeof_type are the domain's.  A segment
 * is a 1-D access (access.hpp:479-514) that the reference also links against
 * every BLAS domain (src/task/task.cc:240-257) through the three rects of
 * interval_to_rects (access.hpp:77-139). */
typedef struct xkrt_hip_access {
    uint64_t addr;   /* begin address in bytes */
    uint64_t m;      /* tile: rows (elements); segment: size in bytes */
    uint64_t n;      /* tile: columns; segment: ignored */
    uint32_t task;   /* spawn-order task number; a task's accesses are contiguous, slot = order inside it */
    uint8_t  mode;   /* XKRT_HIP_MODE_* bits */
    uint8_t  kind;   /* XKRT_HIP_KIND_* */
    uint8_t  pad[2];
} xkrt_hip_access_t;

/* Device-frame rect record (32 B; arrays of them are what the kernels read).
 * Coordinates are in the domain frame: x = byte offset in the column,
 * y = column - the domain's y base (see xkrt_hip_dep_frame).  tm packs the
 * launch-local task number (bits 8..31) and the mode (bits 0..7).  acc is the
 * domain access id the edges refer to; dom the domain's index in its pool
 * (xkrt_hip_dep_index). */
typedef struct xkrt_hip_rect {
    uint32_t x0, x1, y0, y1;
    uint32_t acc, tm, dom, pad;
} xkrt_hip_rect_t;

typedef struct xkrt_hip_dep_config {
    uint32_t bin_x_bytes;  /* bin extent in dim 0 (bytes); 0 = auto */
    uint32_t bin_y_cols;   /* bin extent in dim 1 (columns); 0 = auto */
    uint32_t kernel;       /* LDS capacity class: 0 = small, 1 = medium, 2 = large */
    uint32_t pad;
    uint64_t max_rects;    /* max rect records per submit (sizes scratch) */
    uint64_t max_edges;    /* edge buffer capacity before a drain is required */
    uint64_t budget_bytes; /* device memory budget of the domain; 0 = unlimited */
    uint64_t y_hint_lo;    /* expected column range [lo, hi) (pre-sizes the bin grid); lo == hi: none */
    uint64_t y_hint_hi;
} xkrt_hip_dep_config_t;

typedef struct xkrt_hip_dep_stats {
    uint64_t accesses;      /* accesses resolved so far */
    uint64_t rects;         /* rect records resolved so far */
    uint64_t edges;         /* edges produced so far */
    uint64_t candidates;    /* edge candidates before the cross-bin dedup */
    uint64_t launches;      /* resolve launches */
    uint64_t bins;          /* bins in the grid */
    uint64_t max_frags;     /* largest fragment count seen in one bin */
    uint64_t max_reads;     /* largest reader-pool use seen in one bin */
    uint64_t frag_visits;   /* fragment records scanned by the resolve kernel */
    uint64_t device_bytes;  /* device memory held by the domain */
    uint64_t err;           /* sticky device error bits (0 = healthy).  The domain's own: 2 bin record
                               list, 4 bin fragments, 8 bin readers, 128 record outside the grid; its
                               pool's: 16 candidates, 32 dedup table, 64 edges */
} xkrt_hip_dep_stats_t;

typedef struct xkrt_hip_dep_domain *xkrt_hip_dep_domain_t;

/* Selects the HIP device used by later calls from this thread. */
int xkrt_hip_init(int device);

/* Creates one dependency domain keyed by (ld, sizeof_type), the seam of
 * task_get_dependency_domain_blas_matrix (src/task/task.cc:153-201).  cfg may
 * be NULL (defaults).  stream is a hipStream_t (NULL = a stream of its own). */
int xkrt_hip_dep_domain_create(uint64_t ld, uint64_t sizeof_type, const xkrt_hip_dep_config_t *cfg,
                               void *stream, xkrt_hip_dep_domain_t *out);
int xkrt_hip_dep_domain_destroy(xkrt_hip_dep_domain_t d);

/* Creates a segment dependency domain, the seam of the task domain's
 * IntervalDependencyTree (src/task/task.cc:240-259; the K = 1 tree of
 * interval/dependency-tree.hpp:187-363).  It takes byte segments only
 * (XKRT_HIP_KIND_SEGMENT; a tile is refused with -EINVAL) and gives the edges
 * of the reference's interval tree: each segment is cut into the rects of a
 * frame of col_bytes-byte columns (at most 3) and LINKed and PUT exactly, the
 * same per-point rule as the one-dimensional tree.  col_bytes and cfg's bin
 * extents set how the address space is tiled into bins (bin_x_bytes <=
 * col_bytes); the frame spans 2^32 columns from the first segment seen.
 * Submit, sync, edges and destroy as for xkrt_hip_dep_domain_create's
 * domains.  A task whose accesses mix segments and tiles gives its segments
 * to this domain and, LINK-only, to every BLAS domain (xkrt_hip_dep_submit). */
int xkrt_hip_seg_domain_create(uint64_t col_bytes, const xkrt_hip_dep_config_t *cfg, void *stream,
                               xkrt_hip_dep_domain_t *out);

/* The replay of task_get_dependency_domain_blas_matrix (src/task/task.cc:
 * 177-198): a BLAS domain created while segment accesses are uncompleted gets
 * them PUT, with no LINK.  That PUT marks no node unless the segment starts in
 * the first ld * sizeof_type bytes of memory (see xkrt_hip_dep_submit), so
 * the call changes no edge; it records the dim-0 extent of the inserted rects
 * for the wrapped-tile check.  -ENOTSUP, and nothing recorded, when a segment
 * starts below ld * sizeof_type or the domain has nodes at column 0; -EINVAL
 * for a tile, an empty segment or a segment domain.  Host only: nothing is
 * launched. */
int xkrt_hip_dep_replay_segments(xkrt_hip_dep_domain_t d, const xkrt_hip_access_t *acc, uint64_t n);

/* Enqueues the resolution of n accesses in spawn order: every access of a
 * task is LINKed (slot order) before any is PUT, as task_dependency_resolve
 * does (src/task/task.cc:285-299).  Access ids are assigned consecutively
 * from *first_id (out).  Asynchronous: edges become visible after
 * xkrt_hip_dep_sync.  A wrapped tile (2 rects) is refused with -ENOTSUP, and
 * nothing of the batch is enqueued, when an earlier rect of the domain spans
 * its wrap gap in dim 0: there the reference marks whole split fragments
 * beyond the tile (dependency-tree.hpp:229-256).
 * A segment links like any access but its PUT changes no state: the
 * reference tests the tree's nodes against the segment's own rects, built
 * with ld = SIZE_MAX (access.hpp:495-513), and those meet only nodes that
 * start at column 0 (interval overlap is strict, so the empty rect
 * [0, SIZE_MAX) x [1, 1) meets a node spanning columns 0 and 1).  Such nodes
 * exist once an access touched the first ld * sizeof_type bytes of memory;
 * from then on, and for a segment in those bytes, the batch is refused with
 * -ENOTSUP (the reference would mark whole nodes there). */
int xkrt_hip_dep_submit(xkrt_hip_dep_domain_t d, const xkrt_hip_access_t *acc, uint64_t n, uint64_t *first_id);

/* Same, from device-resident rect records (no host conversion, no copy):
 * nrect records, as xkrt_hip_dep_build_records writes them; [ylo, yhi)
 * bounds every record's y (domain frame) and sizes d's grid; naccess =
 * accesses the records belong to (for accounting).  One launch.  Records
 * name their domain (dom): in a pool they may belong to several domains.
 * The device checks every record against its domain's grid: a record that
 * leaves it (or an empty one) fails its domain with -ERANGE at the next
 * sync, before any of that domain's state changes in the launch; a record
 * naming no domain slot poisons the pool. */
int xkrt_hip_dep_submit_device(xkrt_hip_dep_domain_t d, const xkrt_hip_rect_t *rec, uint64_t nrect,
                               uint64_t naccess, uint32_t ylo, uint32_t yhi);

/* Converts accesses to device-frame rect records on the host (the work
 * xkrt_hip_dep_submit does before its copy).  out holds cap records; the
 * record count goes to *nrect; task numbers become launch-local from 0;
 * access ids start at first_id.  *ylo / *yhi receive the y bounds. */
int xkrt_hip_dep_build_records(xkrt_hip_dep_domain_t d, const xkrt_hip_access_t *acc, uint64_t n,
                               uint32_t first_id, uint64_t cap, xkrt_hip_rect_t *out, uint64_t *nrect,
                               uint32_t *ylo, uint32_t *yhi);

/* y base of the domain frame (fixed at the first submit or from the hint). */
int xkrt_hip_dep_frame(xkrt_hip_dep_domain_t d, uint64_t *ybase);

/* Waits for every enqueued launch; *nedges = edges not yet consumed (of the
 * whole pool).  Returns the pool's error, else the domain's own (a domain
 * that failed in a shared pool fails alone). */
int xkrt_hip_dep_sync(xkrt_hip_dep_domain_t d, uint64_t *nedges);

/* Device pointer to the unconsumed edges, 2 x u32 (pred id, succ id) each,
 * and their count (synchronises first). */
int xkrt_hip_dep_edges_device(xkrt_hip_dep_domain_t d, const uint32_t **edges, uint64_t *n);

/* Copies up to cap unconsumed edges to host memory (synchronises first),
 * then consumes them all.  *n = number of edges that were unconsumed. */
int xkrt_hip_dep_edges_copy(xkrt_hip_dep_domain_t d, uint32_t *host_edges, uint64_t cap, uint64_t *n);

/* Drops every unconsumed edge (asynchronous). */
int xkrt_hip_dep_edges_consume(xkrt_hip_dep_domain_t d);

int xkrt_hip_dep_stats(xkrt_hip_dep_domain_t d, xkrt_hip_dep_stats_t *out);

/* Stream the domain enqueues on (a hipStream_t). */
void *xkrt_hip_dep_stream(xkrt_hip_dep_domain_t d);

/* Forgets every access of the domain (all points back to "never accessed"),
 * drops unconsumed edges (of a private domain) and clears its failed state
 * (a private domain: the pool's too); the bin grid and the frame stay.  Asynchronous.  The reference equivalent is deleting the
 * domain in runtime_t::reset (src/memory/invalidate.cc:68-131). */
int xkrt_hip_dep_reset(xkrt_hip_dep_domain_t d);

/* Retires the accesses with id < horizon (ids as first_id and the edges
 * report them): the caller states that their tasks have completed.  The
 * reference links a completed predecessor to nothing (__task_precedes,
 * include/xkrt/task/task.hpp:452-487), and every access before horizon is
 * completed, so each later edge from one of them would be dropped there:
 * the domain forgets them instead (a point last written by one becomes
 * unwritten, their reader entries go) and the remaining ids drop by
 * horizon -- later edges are exactly the reference's.  This is how a domain
 * outlives 2^32 accesses: ids stay 32-bit, and the id an edge or first_id
 * reports plus xkrt_hip_dep_id_base is the access's id over the domain's
 * life.  After it, first_id continues from the old value - horizon, and
 * caller ids (pools) must be given minus horizon.  -EBUSY while records
 * are staged or edges unconsumed (they carry the old ids) or a serving
 * engine drives the pool; -EINVAL past the ids handed out; -ECANCELED on a
 * failed domain.  Synchronises the stream, then enqueues one kernel. */
int xkrt_hip_dep_retire(xkrt_hip_dep_domain_t d, uint64_t horizon);

/* Sum of the horizons retired so far (0 for a new domain; reset keeps it). */
int xkrt_hip_dep_id_base(xkrt_hip_dep_domain_t d, uint64_t *id_base);

/* Per-kernel HIP-event timing of every later launch (count, resolve over all
 * tiers, dedup, final). */
int xkrt_hip_dep_timing(xkrt_hip_dep_domain_t d, int enable);

/* Accumulated milliseconds per kernel over *launches timed launches, as 6
 * doubles: count, 0, 0, resolve, dedup, final (slots 1 and 2 are kept for the
 * layout and read 0); synchronises; clear != 0 restarts the accumulation. */
int xkrt_hip_dep_kernel_times(xkrt_hip_dep_domain_t d, double *ms, uint64_t *launches, int clear);

/* Device-to-device streaming copy of `bytes` (a multiple of 16) on `stream`:
 * the reference kernel whose HBM byte counters calibrate the resolver's. */
int xkrt_hip_copy(void *dst, const void *src, uint64_t bytes, void *stream);

/* Per timed launch, milliseconds from its first kernel's start to its last
 * kernel's end (up to cap values; *n = launches recorded). */
int xkrt_hip_dep_launch_times(xkrt_hip_dep_domain_t d, float *ms, uint64_t cap, uint64_t *n, int clear);

/* ------------------------------------------------------------------------
 * Pools: many dependency domains (clients) resident on one device and
 * resolved together.  A pool holds the bin state of up to 256 domains and
 * admits a domain only while its device-memory budget allows (the serving
 * loop's admission control).  Batches of any domains are staged into the
 * pool's open launch and resolved by one launch -- the same kernels as one
 * domain, over the union of the batches' bins -- optionally replayed from a
 * captured hipGraph.  Each domain keeps its own access ids; edges carry
 * their domain's index.  A domain from xkrt_hip_dep_domain_create is a
 * private pool with one domain.
 * No counterpart in the reference, which resolves each task inline in the
 * spawning thread (src/task/task.cc:285-299).
 * ------------------------------------------------------------------------ */

typedef struct xkrt_hip_pool *xkrt_hip_pool_t;

typedef struct xkrt_hip_pool_config {
    uint32_t kernel;        /* LDS capacity class of every bin: 0, 1, 2 (see xkrt_hip_dep_config_t) */
    uint32_t graph;         /* 1: launches of staged batches replay a captured hipGraph */
    uint64_t max_rects;     /* rect records per launch (sizes the staging and scratch buffers) */
    uint64_t max_edges;     /* edge buffer capacity before a drain is required */
    uint64_t budget_bytes;  /* device memory budget: bins of every domain + scratch; 0 = unlimited */
    uint32_t direct_dedup;  /* 1: each bin's candidates go straight into the launch's dedup table (one
                               kernel less per launch: for small, latency-bound launches); 0: they are
                               staged and deduplicated by a kernel of their own (large launches) */
    uint32_t small_launch_rects; /* launches of at most this many rects read their records in place from
                               pinned host memory (no upload) and start at the second LDS tier (one
                               resolve kernel less); 0 = never */
    uint32_t host_edges;    /* 1: the edge buffer lives in pinned host memory that the device writes
                               (a launch's edges are on the host when it completes, with no copy) */
    uint32_t pad;
} xkrt_hip_pool_config_t;

typedef struct xkrt_hip_pool_stats {
    uint64_t domains;        /* domains admitted */
    uint64_t bins_capacity;  /* bins allocated on the device */
    uint64_t bins_used;      /* bins held by domains */
    uint64_t bytes_per_bin;
    uint64_t device_bytes;   /* bins + scratch */
    uint64_t launches;
    uint64_t graph;          /* 1 when a captured launch is ready */
    uint64_t poisoned;
} xkrt_hip_pool_stats_t;

int xkrt_hip_pool_create(const xkrt_hip_pool_config_t *cfg, void *stream, xkrt_hip_pool_t *p);
int xkrt_hip_pool_destroy(xkrt_hip_pool_t p);

/* Admits a domain for (ld, sizeof_type) over columns [y_lo, y_hi) with the
 * given bin extents (powers of two; 0 = 1024 B x 128 columns).  -ENOMEM: the
 * budget cannot hold its bins now (retry after other domains leave).  Leave
 * with xkrt_hip_dep_domain_destroy.  y_lo == y_hi: no range; the first batch
 * frames the domain and its grid grows with later batches (within the
 * budget), as a single domain's does. */
int xkrt_hip_pool_domain_create(xkrt_hip_pool_t p, uint64_t ld, uint64_t sizeof_type, uint32_t bin_x_bytes,
                                uint32_t bin_y_cols, uint64_t y_lo, uint64_t y_hi, xkrt_hip_dep_domain_t *d);

/* Admits a segment domain (see xkrt_hip_seg_domain_create) over columns
 * [y_lo, y_hi) of col_bytes bytes, i.e. addresses [y_lo, y_hi) * col_bytes. */
int xkrt_hip_pool_seg_domain_create(xkrt_hip_pool_t p, uint64_t col_bytes, uint32_t bin_x_bytes, uint32_t bin_y_cols,
                                    uint64_t y_lo, uint64_t y_hi, xkrt_hip_dep_domain_t *d);

/* Appends a batch of d's accesses (whole tasks, spawn order) to the open
 * launch; ids as xkrt_hip_dep_submit.  -ENOSPC: the open launch is full;
 * -ECANCELED: d failed earlier (xkrt_hip_dep_sync(d) gives its error). */
int xkrt_hip_pool_stage(xkrt_hip_pool_t p, xkrt_hip_dep_domain_t d, const xkrt_hip_access_t *acc, uint64_t n,
                        uint64_t *first_id);

/* Stages nd batches, batch k of domain d[k] (acc[k], n[k] accesses), as one
 * unit: either all of them are appended to the open launch (first ids in
 * first_id[k], nullable) or none is, and the first refusal's code is returned
 * (-ENOSPC: they do not fit the open launch together; launch it and stage
 * again).  A task whose accesses go to several domains is staged this way.
 * ids (nullable; ids[k] nullable): the access ids the edges of batch k name,
 * instead of the domain's consecutive ids -- strictly increasing within the
 * domain across its batches (-EINVAL otherwise), e.g. program-wide ids shared
 * by several domains. */
int xkrt_hip_pool_stage_group(xkrt_hip_pool_t p, uint32_t nd, const xkrt_hip_dep_domain_t *d,
                              const xkrt_hip_access_t *const *acc, const uint32_t *const *ids, const uint64_t *n,
                              uint64_t *first_id);

/* Resolves every staged batch in one launch (asynchronous).  *ticket (may
 * be NULL) = the launch's sequence number, 1 for the pool's first launch. */
int xkrt_hip_pool_launch(xkrt_hip_pool_t p, uint64_t *ticket);
/* *done = every launch up to this ticket has completed (non-blocking). */
int xkrt_hip_pool_poll(xkrt_hip_pool_t p, uint64_t *done);
/* Blocks until the launch with this ticket has completed. */
int xkrt_hip_pool_wait(xkrt_hip_pool_t p, uint64_t ticket);
/* Staged but not launched: records, accesses, batches. */
int xkrt_hip_pool_open(xkrt_hip_pool_t p, uint64_t *nrect, uint64_t *naccess, uint64_t *nbatches);
/* Waits for every launch; returns the pool's error only (a domain that
 * failed alone is reported by xkrt_hip_dep_sync on it). */
int xkrt_hip_pool_sync(xkrt_hip_pool_t p, uint64_t *nedges);
/* Copies and consumes every unconsumed edge: edges = (pred id, succ id)
 * pairs, doms = the domain index of each.  Edges of failed domains are
 * dropped (*n counts the rest). */
int xkrt_hip_pool_edges_copy(xkrt_hip_pool_t p, uint32_t *edges, uint32_t *doms, uint64_t cap, uint64_t *n);
/* For consumers on the device: the edge buffer ((pred, succ) u32 pairs), the
 * domain of each edge and the device word holding the number of unconsumed
 * edges (read it on the pool's stream, after a launch). */
int xkrt_hip_pool_edges_device(xkrt_hip_pool_t p, const uint32_t **edges, const uint32_t **doms,
                               const uint64_t **count);
/* Consumes every edge on the device, in stream order (no host copy). */
int xkrt_hip_pool_edges_consume(xkrt_hip_pool_t p);
int xkrt_hip_pool_stats(xkrt_hip_pool_t p, xkrt_hip_pool_stats_t *out);
int xkrt_hip_pool_timing(xkrt_hip_pool_t p, int enable);
/* Per timed launch, milliseconds from its first kernel's start to its last kernel's end. */
int xkrt_hip_pool_launch_times(xkrt_hip_pool_t p, float *ms, uint64_t cap, uint64_t *n, int clear);
void *xkrt_hip_pool_stream(xkrt_hip_pool_t p);
/* The domain's index in its pool (the dom field of records and edges). */
int xkrt_hip_dep_index(xkrt_hip_dep_domain_t d, uint32_t *idx);

/* ------------------------------------------------------------------------
 * Serving engine: a host thread that drives one pool for many clients.
 * Clients submit batches (whole tasks, spawn order) from any thread; the
 * engine admits a client's domain against the pool's budget at its first
 * batch (clients wait while the budget is full), stages every arrived batch
 * of admitted clients into one launch whenever fewer than `depth` launches
 * are in flight (continuous batching; max_batches = 1 gives the plain
 * per-batch form), retires launches as they complete, and releases a closed
 * client's bins once its last batch is resolved; a second host thread copies
 * each retired launch's edges out per client and stamps its batches ready
 * (submit times are stamped at submit).
 * While an engine runs, the pool's own stage / launch / edge calls return
 * -EBUSY.  No counterpart in the reference (its resolver runs inline in the
 * spawning thread, src/task/task.cc:285-299).
 * ------------------------------------------------------------------------ */

typedef struct xkrt_hip_server *xkrt_hip_server_t;

typedef struct xkrt_hip_server_config {
    uint32_t depth;          /* launches in flight, 1..16 (0 = 2) */
    uint32_t max_batches;    /* batches per launch; 0 = every arrived batch that fits */
    uint64_t max_rects;      /* soft cap of rects per launch; 0 = the pool's max_rects */
    uint64_t reserve_bytes;  /* bin memory allocated up front (within the budget), so
                                admission does not allocate while serving */
    uint32_t prearm;         /* 1: the next launch is queued on the device while the current one
                                runs, held by a gate kernel that the engine opens by writing pinned
                                memory once the batches are staged (no launch calls on the critical
                                path); needs depth <= 2 and a pool with small_launch_rects > 0 (with
                                depth 2 the gate of the launch after next is armed once the
                                oldest launch has freed its staging buffer) */
    uint32_t pad;
} xkrt_hip_server_config_t;

typedef struct xkrt_hip_server_stats {
    uint64_t launches;
    uint64_t batches;        /* batches resolved */
    uint64_t admitted;       /* client domains admitted */
    uint64_t left;           /* clients that left */
    uint64_t waited;         /* admission attempts refused for lack of budget */
    uint64_t harvests;       /* edge copies */
    uint64_t consumes;       /* edge buffer resets */
    uint64_t gated;          /* launches started by opening a gate (prearm) */
    int64_t error;           /* first pool-level error (negative errno), 0 if none */
} xkrt_hip_server_stats_t;

int xkrt_hip_server_create(xkrt_hip_pool_t p, const xkrt_hip_server_config_t *cfg, xkrt_hip_server_t *s);
/* Stops the engine (after the launches in flight), releases every client's
 * domain; the pool stays usable. */
int xkrt_hip_server_destroy(xkrt_hip_server_t s);
/* Registers a client: a domain for (ld, sizeof_type) over columns
 * [y_lo, y_hi) with the given bin extents, admitted at its first batch. */
int xkrt_hip_server_client(xkrt_hip_server_t s, uint64_t ld, uint64_t sizeof_type, uint32_t bin_x_bytes,
                           uint32_t bin_y_cols, uint64_t y_lo, uint64_t y_hi, uint32_t *client);
/* Queues a copy of n accesses of the client; *request = its number (0, 1, ...
 * over all clients).  The client's access ids continue from its previous
 * batch, from 0. */
int xkrt_hip_server_submit(xkrt_hip_server_t s, uint32_t client, const xkrt_hip_access_t *acc, uint64_t n,
                           uint64_t *request);
/* The client submits no more batches; it leaves after the last one. */
int xkrt_hip_server_close(xkrt_hip_server_t s, uint32_t client);
/* Waits until every submitted batch is resolved (or refused) and every
 * closed client has left; -ETIMEDOUT after timeout_us (0 = no limit). */
int xkrt_hip_server_drain(xkrt_hip_server_t s, uint64_t timeout_us);
/* Per request first .. first + n - 1: microseconds since the engine started
 * at submit and at ready (the launch that resolved it completed and its
 * edges are in its client's host buffer), and status (1 pending, 0 resolved,
 * negative errno: refused; a client whose domain fails on the device gets
 * its error, e.g. -ENOSPC, on that batch and every later one, alone). */
int xkrt_hip_server_times(xkrt_hip_server_t s, uint64_t first, uint64_t n, double *submit_us, double *ready_us,
                          int32_t *status);
/* The client's edges so far: (pred id, succ id) pairs in its own access ids;
 * *n = pairs available (up to cap copied). */
int xkrt_hip_server_edges(xkrt_hip_server_t s, uint32_t client, uint32_t *edges, uint64_t cap, uint64_t *n);
int xkrt_hip_server_stats(xkrt_hip_server_t s, xkrt_hip_server_stats_t *out);

/* ------------------------------------------------------------------------
 * Task domains: the dependency resolution of one task domain over every kind
 * of access it holds, the seam of task_dependency_resolve (src/task/task.cc:
 * 204-299).  A task domain owns a pool with one segment domain, created at
 * the first segment (the IntervalDependencyTree of task.cc:240-259), and one
 * BLAS domain per (ld, sizeof_type), created at the first tile of that pair
 * (task_get_dependency_domain_blas_matrix, task.cc:153-201).  A segment goes
 * to the segment domain and, LINK-only, to every BLAS domain existing when it
 * is linked; a BLAS domain created while segments are in flight gets them
 * replayed (xkrt_hip_dep_replay_segments).  Every batch is staged into all
 * its domains as one unit and resolved by one launch.  Access ids are
 * program-wide (64-bit, consecutive in submit order); each predecessor's
 * edges come from one domain (a segment's from the segment domain, a tile's
 * from its pair's), so the domains' edges, mapped to program ids, are the
 * reference's edges with no cross-domain duplicate.
 * Handle accesses (XKRT_HIP_KIND_HANDLE) go to a handle domain created at the
 * first one (the DependencyMap of task.cc:222-234, see xkrt_hip_handle_*):
 * their edges join the others' in program ids, an empty write the map inserts
 * as XKRT_HIP_TASKDOM_VIRTUAL | k; with release on, the device runs them itself
 * (an empty task, completed as soon as it is released).  Limits:
 * 2^18 handle accesses per batch (-EINVAL), 2^22 over the domain's life
 * (-ENOSPC).
 * Host memory: 32 bytes per segment (the replay), 4 bytes per handle access
 * (its program id).
 * ------------------------------------------------------------------------ */

typedef struct xkrt_hip_taskdom *xkrt_hip_taskdom_t;

/* One access of a task domain (48 B). */
typedef struct xkrt_hip_task_access {
    uint64_t addr;         /* begin address in bytes */
    uint64_t m;            /* tile: rows (elements); segment: size in bytes */
    uint64_t n;            /* tile: columns; segment: ignored */
    uint64_t ld;           /* tile: leading dimension (elements); segment: ignored */
    uint32_t sizeof_type;  /* tile: element size (bytes); segment: ignored */
    uint32_t task;         /* spawn-order task number; a task's accesses are contiguous */
    uint8_t  mode;         /* XKRT_HIP_MODE_* bits */
    uint8_t  kind;         /* XKRT_HIP_KIND_TILE, _SEGMENT or _HANDLE */
    uint8_t  concurrency;  /* handle: XKRT_HIP_CONC_* (0 = sequential); otherwise ignored */
    uint8_t  pad[5];
} xkrt_hip_task_access_t;

/* Program id bit of an empty write inserted by the task domain's handle map
 * (its k-th one: XKRT_HIP_TASKDOM_VIRTUAL | k; an empty task with one VW
 * access, see xkrt_hip_handle_sync). */
#define XKRT_HIP_TASKDOM_VIRTUAL (1ull << 63)

typedef struct xkrt_hip_taskdom_config {
    uint32_t kernel;          /* LDS capacity class of every domain (see xkrt_hip_dep_config_t) */
    uint32_t release;         /* 1: edges are applied on the device and tasks released there (see
                                 xkrt_hip_taskdom_ready); 0: edges are copied out */
    uint64_t max_rects;       /* rect records per launch (0 = 2^20) */
    uint64_t max_edges;       /* edge buffer capacity before edges must be copied out (0 = 2^22) */
    uint64_t budget_bytes;    /* device memory budget of all its domains; 0 = unlimited */
    uint64_t seg_col_bytes;   /* segment domain: column width of its frame (0 = 65536) */
    uint32_t seg_bin_x_bytes; /* segment domain bins (0 = min(column width, 4096) B x 1 column) */
    uint32_t seg_bin_y_cols;
    uint32_t blas_bin_x_bytes;/* BLAS domain bins (0 = 1024 B x 128 columns) */
    uint32_t blas_bin_y_cols;
    uint64_t max_tasks;       /* release: tasks of the domain's lifetime (0 = 2^20) */
    uint64_t max_accesses;    /* release: accesses (0 = 4 x max_tasks) */
    uint64_t max_succ;        /* release: successor-list entries, one per edge resolved (0 = 2 x max_accesses) */
    uint64_t max_virtual;     /* release: empty writes the handle map may insert (0 = max_tasks) */
} xkrt_hip_taskdom_config_t;

typedef struct xkrt_hip_taskdom_stats {
    uint64_t accesses;        /* accesses resolved or enqueued */
    uint64_t segments;        /* of which segments */
    uint64_t blas_domains;    /* BLAS domains created */
    uint64_t launches;
    uint64_t edges;           /* edges copied out so far */
    uint64_t succ_nodes;      /* release: successor-list entries taken, one per edge resolved */
    uint64_t released;        /* release: tasks released */
    uint64_t completed;       /* release: tasks completed */
} xkrt_hip_taskdom_stats_t;

int xkrt_hip_taskdom_create(const xkrt_hip_taskdom_config_t *cfg, void *stream, xkrt_hip_taskdom_t *out);
int xkrt_hip_taskdom_destroy(xkrt_hip_taskdom_t t);

/* Enqueues a batch of whole tasks in spawn order (task numbers non-decreasing
 * across batches); ids from *first_id (out, nullable).  Launches when the
 * open launch is full.  Either the whole batch is enqueued or nothing of it:
 * -EINVAL (bad access, task order, more than 255 BLAS domains), -ENOTSUP (a
 * case a domain refuses, see xkrt_hip_dep_submit and
 * xkrt_hip_dep_replay_segments), -ENOMEM (budget). */
int xkrt_hip_taskdom_submit(xkrt_hip_taskdom_t t, const xkrt_hip_task_access_t *acc, uint64_t n,
                            uint64_t *first_id);
/* Launches what is open and waits for every launch; *nedges = edges waiting
 * to be copied out.  A domain's device error is returned here (and later). */
int xkrt_hip_taskdom_sync(xkrt_hip_taskdom_t t, uint64_t *nedges);
/* Copies out and consumes every waiting edge: (pred, succ) program access ids,
 * 2 u64 each (an inserted empty write: XKRT_HIP_TASKDOM_VIRTUAL | k).  -ENOSPC, nothing consumed, *n = edges waiting, when cap is
 * smaller.  Call after xkrt_hip_taskdom_sync.  -EPERM with release on. */
int xkrt_hip_taskdom_edges_copy(xkrt_hip_taskdom_t t, uint64_t *edges, uint64_t cap, uint64_t *n);

/* Release on the device (config release = 1; tasks numbered 0, 1, ... in
 * submit order).  What the runtime does with an edge, a commit and a
 * completion (__task_precedes, __task_commit, include/xkrt/task/task.hpp:
 * 453-551; __task_complete, src/task.cc:160-247), done by kernels on the
 * pool's stream: after each launch its edges are applied -- an edge from a
 * task not yet completed adds one to the successor's wait counter and puts
 * it on the predecessor access's list -- and its tasks' commit guards are
 * dropped; a completion walks the task's lists (detached accesses, mode D,
 * excepted) and decrements.  A counter reaching 0 releases its task into a
 * queue in pinned host memory.
 * xkrt_hip_taskdom_ready copies up to cap released tasks not read yet, in
 * release order, without blocking (tasks whose launch has not been started
 * are not there: xkrt_hip_taskdom_sync starts it).
 * xkrt_hip_taskdom_complete enqueues the completion of released tasks; a task
 * not released, or completed twice, fails the domain (-EINVAL at the next
 * sync).  Capacities: max_tasks, max_accesses, max_succ (-ENOSPC).
 * xkrt_hip_taskdom_complete_wait is the synchronous form, as __task_complete
 * is on the host: it returns once the device has released the successors of
 * these tasks (and of every completion enqueued before), so the next
 * xkrt_hip_taskdom_ready sees them; it polls a word in pinned memory that the
 * completion kernel writes, without synchronising the stream (-ETIMEDOUT
 * after 10 s). */
int xkrt_hip_taskdom_ready(xkrt_hip_taskdom_t t, uint32_t *tasks, uint64_t cap, uint64_t *n);
int xkrt_hip_taskdom_complete(xkrt_hip_taskdom_t t, const uint32_t *tasks, uint64_t n);
int xkrt_hip_taskdom_complete_wait(xkrt_hip_taskdom_t t, const uint32_t *tasks, uint64_t n);
int xkrt_hip_taskdom_stats(xkrt_hip_taskdom_t t, xkrt_hip_taskdom_stats_t *out);

/* ------------------------------------------------------------------------
 * Handle dependency domains: OpenMP-style dependencies on opaque handles
 * (pointers), the seam of the task domain's DependencyMap
 * (include/xkrt/memory/access/handle/dependency-map.hpp:52-330; created at
 * src/task/task.cc:222-234).  Per handle, the map keeps the concurrent writes
 * since the last sequential write, the sequential reads since, and that
 * write; a concurrent write after reads, or a read after concurrent writes,
 * goes through an empty sequential write the map inserts (a task of its own,
 * insert_empty_write, 94-135).  Edges are as the reference makes them when no
 * task has completed (every edge materialised), deduplicated as
 * __access_precedes does (include/xkrt/task/task.hpp:494-512).
 * Each submit is one launch; handles are independent and resolved in
 * parallel, each handle's accesses in spawn order.
 * ------------------------------------------------------------------------ */

typedef struct xkrt_hip_handle_domain *xkrt_hip_handle_domain_t;

#define XKRT_HIP_CONC_SEQUENTIAL  1u  /* ACCESS_CONCURRENCY_* (include/xkrt/memory/access/concurrency.h:41-46) */
#define XKRT_HIP_CONC_COMMUTATIVE 2u
#define XKRT_HIP_CONC_CONCURRENT  4u
#define XKRT_HIP_VIRTUAL_ID 0x80000000u  /* id bit of an inserted empty write */

typedef struct xkrt_hip_handle_access {  /* 16 B */
    uint64_t handle;       /* the handle (any value but UINT64_MAX) */
    uint32_t task;         /* spawn-order task number, < 2^31; a task's accesses are contiguous */
    uint8_t  mode;         /* XKRT_HIP_MODE_* bits */
    uint8_t  concurrency;  /* XKRT_HIP_CONC_* */
    uint8_t  pad[2];
} xkrt_hip_handle_access_t;

typedef struct xkrt_hip_handle_config {
    uint64_t max_handles;   /* distinct handles of the domain's life (0 = 2^16) */
    uint64_t max_accesses;  /* accesses of its life, < 2^31 (0 = 2^22) */
    uint64_t max_virtual;   /* empty writes it may insert (0 = 2^20) */
    uint64_t max_batch;     /* accesses per submit (0 = 2^18) */
    uint64_t max_edges;     /* edge buffer before a copy is needed (0 = 2^22), in 64 stripes of max_edges / 64:
                               a handle's edges go to stripe (its table slot mod 64) */
} xkrt_hip_handle_config_t;

int xkrt_hip_handle_domain_create(const xkrt_hip_handle_config_t *cfg, void *stream, xkrt_hip_handle_domain_t *out);
int xkrt_hip_handle_domain_destroy(xkrt_hip_handle_domain_t h);
/* Resolves a batch of whole tasks (asynchronous); ids consecutive from
 * *first_id (out, nullable).  -EINVAL: bad access or task order; -ENOSPC:
 * max_accesses reached. */
int xkrt_hip_handle_submit(xkrt_hip_handle_domain_t h, const xkrt_hip_handle_access_t *acc, uint64_t n,
                           uint64_t *first_id);
/* Same, with the access ids the edges name given (ids[i] of acc[i], strictly
 * increasing across the domain's submits, below max_accesses): e.g. a task
 * domain's program ids. */
int xkrt_hip_handle_submit_ids(xkrt_hip_handle_domain_t h, const xkrt_hip_handle_access_t *acc, const uint32_t *ids,
                               uint64_t n);
/* For consumers on the device (the domain's stream order): the edge buffer,
 * *nstripes stripes of *stripe_cap (pred, succ) u32 pairs, the u64 count of
 * each stripe, and the u32 device word counting the inserted empty writes.
 * xkrt_hip_handle_edges_consume clears the counts in stream order. */
int xkrt_hip_handle_edges_device(xkrt_hip_handle_domain_t h, const uint32_t **edges, const uint64_t **stripe_counts,
                                 uint32_t *nstripes, uint64_t *stripe_cap, const uint32_t **nvirtual);
int xkrt_hip_handle_edges_consume(xkrt_hip_handle_domain_t h);
/* Waits for the domain's launches: *nedges waiting edges, *nvirtual empty
 * writes inserted so far (ids XKRT_HIP_VIRTUAL_ID | k, k < *nvirtual; each
 * one is an empty task with one VW access).  -ENOSPC when a capacity
 * overflowed (the domain's edges are then incomplete). */
int xkrt_hip_handle_sync(xkrt_hip_handle_domain_t h, uint64_t *nedges, uint64_t *nvirtual);
/* Copies out and consumes the waiting edges, (pred id, succ id) u32 pairs;
 * -ENOSPC and nothing consumed when cap is smaller. */
int xkrt_hip_handle_edges_copy(xkrt_hip_handle_domain_t h, uint32_t *edges, uint64_t cap, uint64_t *n);

/* ------------------------------------------------------------------------
 * Coherency fetch planner: one memory tree per (ld, sizeof_type), a host
 * (device 0) and one GPU (device 1), no eviction.  By default completion is
 * synchronous (the fetches of an access complete before the next access);
 * with async_completion the fetches stay in flight until
 * xkrt_hip_mcc_complete, and later accesses see them being fetched.
 *
 * Replaces KBLASMemoryTree<2>::fetch (memory-tree.hpp:1806-1874) with its
 * host path (fetch_list_to_host, 1110-1215), allocation-view reuse and
 * creation (1293-1450), copy setup (1471-1631), set coherent (1633-1667),
 * fetch list (1067-1108) and completion (955-1061, 2067-2127); and
 * who_owns (1903-1933).  One tree per (ld, sizeof_type) as
 * task_get_memory_controller creates them (src/task/task.cc:50-150).
 * ------------------------------------------------------------------------ */

typedef struct xkrt_hip_mcc_config {
    uint64_t y_lo, y_hi;   /* column range of the tree's frame: columns addr / (ld * sizeof_type) in [y_lo, y_hi) */
    uint64_t budget_bytes; /* device memory budget of the tree state; 0 = unlimited */
    uint32_t async_completion; /* 1: fetches complete when xkrt_hip_mcc_complete says so (fetch_callback,
                                  memory-tree.hpp:955-1061); 0: at the end of their access */
    uint32_t pad;
} xkrt_hip_mcc_config_t;

typedef struct xkrt_hip_mcc_access {
    uint64_t addr;   /* begin address in bytes (a multiple of sizeof_type) */
    uint64_t m;      /* rows (elements); the tile must not wrap a column */
    uint64_t n;      /* columns */
    uint8_t  dev;    /* 0 = host, 1 = the GPU */
    uint8_t  mode;   /* XKRT_HIP_MODE_* bits */
    uint8_t  pad[6];
} xkrt_hip_mcc_access_t;

typedef struct xkrt_hip_mcc_stats {
    uint64_t accesses;      /* accesses planned so far */
    uint64_t gpu_accesses;  /* of which on the GPU */
    uint64_t chunks;        /* GPU chunks allocated */
    uint64_t last_parts;    /* valid-region rows of the last plan call */
    uint64_t last_fetches;  /* fetch rows of the last plan call */
    uint64_t device_bytes;  /* device memory of the tree state */
    uint64_t poisoned;      /* 1 after an abort (ninth view) or a capacity error */
    float    last_ms;       /* GPU time of the last plan call, first kernel to last */
    uint32_t pad;
    uint64_t last_fetch_base;  /* id of the first fetch row of the last plan call (async: ids of row r = base + r) */
    uint64_t in_flight;        /* async: fetches issued and not completed */
    uint64_t registrations;    /* async: accesses awaiting and forwards registered, not yet fired */
} xkrt_hip_mcc_stats_t;

typedef struct xkrt_hip_mcc *xkrt_hip_mcc_t;

int xkrt_hip_mcc_create(uint64_t ld, uint64_t sizeof_type, const xkrt_hip_mcc_config_t *cfg, void *stream,
                        xkrt_hip_mcc_t *h);
int xkrt_hip_mcc_destroy(xkrt_hip_mcc_t h);

/* Plans n accesses in order (synchronous).  *nparts / *nfetches = rows of
 * this call.  *fail_at = index in this call of the access at which the
 * reference aborts with a ninth allocation view (memory-tree.hpp:1362-1365),
 * or -1; the rows stop before it and the tree refuses later calls
 * (-ECANCELED).  -ENOTSUP: a tile wrapping a column (two rects). */
int xkrt_hip_mcc_plan(xkrt_hip_mcc_t h, const xkrt_hip_mcc_access_t *acc, uint64_t n, uint64_t *nparts,
                      uint64_t *nfetches, int64_t *fail_at);

/* Valid-region split of the last plan call: part_off[n + 1] row offsets per
 * access; parts = 6 int64 per row (x0, x1 bytes; y0, y1 columns;
 * coherency bits before the access, 1 = host, 2 = GPU; fetching bits before
 * it, always 0 with synchronous completion). */
int xkrt_hip_mcc_parts_copy(xkrt_hip_mcc_t h, uint64_t *part_off, int64_t *parts, uint64_t cap);

/* Fetch list of the last plan call: fetch_off[n + 1]; 8 int64 per row
 * (x0, x1, y0, y1, src device, src chunk, dst device, dst chunk; chunk -1 =
 * host memory; chunks are numbered in allocation order from 0). */
int xkrt_hip_mcc_fetches_copy(xkrt_hip_mcc_t h, uint64_t *fetch_off, int64_t *fetches, uint64_t cap);

/* Asynchronous completion (trees created with async_completion).  Every fetch
 * row has an id (fetch rows of a plan call: last_fetch_base + row; forwards:
 * see below).  A GPU read of a view already being fetched does not fetch again:
 * the access awaits that fetch (memory-tree.hpp:1516-1527).  A GPU read of data
 * the GPU holds nowhere coherent while some view of it is being fetched there
 * registers a D2D forward from that view instead (1567-1600).  Both are
 * registrations, listed per access of the last plan call by
 * xkrt_hip_mcc_awaits_copy: awaits_off[n + 1] row offsets; rows of 8 int64
 * (x0, x1 bytes; y0, y1 columns; kind 2 = await, 3 = forward; the chunk whose
 * fetch fires it; the forward's destination chunk or -1; registration id). */
int xkrt_hip_mcc_awaits_copy(xkrt_hip_mcc_t h, uint64_t *awaits_off, int64_t *rows, uint64_t cap);

/* Completes n fetches in flight, in this order (fetch_callback,
 * memory-tree.hpp:955-1061; SEARCH_FETCHED 2067-2127): over each fetch's rect
 * the destination view becomes coherent and stops fetching, the device becomes
 * coherent, and stops fetching once none of its views is.  The registrations
 * on the fetched view fire: awaiting accesses are released and forwards are
 * issued as new fetches (ids continuing the tree's count, in the order of the
 * completions and, within one, of the registrations).  -EINVAL: an id not in
 * flight (or twice).  *nreleased / *nforwards = rows of this call. */
int xkrt_hip_mcc_complete(xkrt_hip_mcc_t h, const uint64_t *fetch_ids, uint64_t n, uint64_t *nreleased,
                          uint64_t *nforwards);

/* What the last xkrt_hip_mcc_complete call fired, per completion k:
 * rel_off[n + 1] / fwd_off[n + 1] row offsets.  released: 7 int64 per row
 * (access index in the tree's order, x0, x1, y0, y1, chunk, registration id);
 * forwards: 9 int64 per row (fetch id, x0, x1, y0, y1, src device, src chunk,
 * dst device, dst chunk). */
int xkrt_hip_mcc_completed_copy(xkrt_hip_mcc_t h, uint64_t *rel_off, int64_t *released, uint64_t rel_cap,
                                uint64_t *fwd_off, int64_t *forwards, uint64_t fwd_cap);

/* Fetch execution (trees with async_completion): the copies behind the fetch
 * rows, as the reference's HIP driver issues them per fetch with
 * hipMemcpyParam2DAsync (src/driver/driver_hip.cc:661-751, queue.cc:165-284).
 * xkrt_hip_mcc_exec_bind gives the tree its memory: host_base = host address
 * of byte 0 of column y_lo of the matrix (ld * sizeof_type bytes per column),
 * pinned memory the device can map (hipHostMalloc); arena = device memory for
 * the GPU chunks (arena_bytes), each chunk placed there on first use and
 * holding its allocating access's tile column-major (pitch = its rows times
 * sizeof_type). */
int xkrt_hip_mcc_exec_bind(xkrt_hip_mcc_t h, void *host_base, void *arena, uint64_t arena_bytes);

/* Enqueues the copies of n fetches in flight on `stream` (a hipStream_t), one
 * kernel for the batch: host to GPU, GPU to GPU (another view, forwards), GPU
 * to host, each the 2D copy of its fetch's rect.  It completes nothing: once
 * the stream has run them, pass the same ids to xkrt_hip_mcc_complete.
 * -ENOMEM: the arena cannot place a chunk; -EINVAL: an id not in flight. */
int xkrt_hip_mcc_exec(xkrt_hip_mcc_t h, const uint64_t *fetch_ids, uint64_t n, void *stream);

/* The GPU chunk each access of the last plan call reads or writes through
 * (its allocation view, memory-tree.hpp:1293-1450), -1 for host accesses. */
int xkrt_hip_mcc_access_chunks(xkrt_hip_mcc_t h, int64_t *chunks);

/* Device address of a chunk (placed in the arena on first use), its pitch in
 * bytes, and the rect it holds (x0, x1 bytes; y0, y1 columns). */
int xkrt_hip_mcc_chunk_ptr(xkrt_hip_mcc_t h, uint32_t chunk, void **ptr, uint64_t *pitch, int64_t *rect);

/* who_owns for n rects on the current state: device bitfield owning the
 * most bytes (ties: several bits; 0 when nobody owns a byte). */
int xkrt_hip_mcc_who_owns(xkrt_hip_mcc_t h, const xkrt_hip_mcc_access_t *acc, uint64_t n, uint32_t *owners);

/* The same answers as xkrt_hip_mcc_who_owns, with no allocation: the queries,
 * their rects and the answers are staged in a static device array of the
 * library (1.6 MiB per GPU, held until the process ends) instead of device
 * memory allocated per call.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (tight scheduling loops that
 * place tasks by owner, code paths that forbid hipMalloc).  At most
 * XKRT_HIP_WHO_OWNS_STATIC_MAX queries per call (-E2BIG beyond).  Every call
 * on one GPU shares the array: calls that overlap on one GPU (from several
 * threads) race and must be serialised by the caller.  Synchronous. */
#define XKRT_HIP_WHO_OWNS_STATIC_MAX 32768
int xkrt_hip_mcc_who_owns_static(xkrt_hip_mcc_t h, const xkrt_hip_mcc_access_t *acc, uint64_t n, uint32_t *owners);

int xkrt_hip_mcc_stats(xkrt_hip_mcc_t h, xkrt_hip_mcc_stats_t *out);

#ifdef __cplusplus
}
#endif

#endif /* XKRT_HIP_H */
