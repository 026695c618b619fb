// MI355X (gfx950, wave64) dependency resolver for 2D (ld, sizeof_type)
// region accesses: RAW / WAR / WAW predecessor edges with the reference's
// semantics (include/xkrt/memory/access/blas/dependency-tree.hpp:229-355,
// include/xkrt/task/task.hpp:494-512, src/task/task.cc:285-299).
//
// Data layout in HBM, per domain:
//   * the domain frame is cut into a dense grid of bins, BX bytes (dim 0) by
//     BY columns (dim 1), both powers of two;
//   * each bin holds a writer painter map -- fragments tiling the bin's
//     written area, each with its last writer (access id) -- and a log of the
//     exact readers clipped to the bin (AoS, 5 u32 per fragment / entry:
//     x0, x1, y0, y1, lw or acc; fcap fragments and rcap entries per bin),
//     and two 64-bit single-bin masks (fragment / entry i: its access lies
//     in this bin alone; kept by the register replay, cleared elsewhere);
//   * rect records of a launch are an array of 32 B structs, two 16-byte
//     words each {x0, x1, y0, y1}, {acc, task<<8|mode, domain, 0}: one record
//     is two dwordx4 loads by one lane (the record an incidence refers to is
//     read whole, never a single field);
//   * incidences (record x bin) are 32 B: {x0, x1, y0, y1}, {acc, tm, record, one}
//     (one: the record's access is this record alone and lies in this bin).
//
// One launch resolves one batch of records in these kernels:
//   k_count   records -> per-bin record counts; each incidence goes into its
//             bin's bucket, each bin the launch touches into the touched list
//   k_resolve_reg  one wave64 per touched bin of at most 64 records, fragments
//             and log entries: the bin's whole state in registers (one
//             fragment, log entry and record per lane), replayed -- LINK of
//             every access of a task, then PUT -- with readlane and ballots.
//             Candidates (pred, succ access, succ task) are staged; an edge
//             whose predecessor's access lies in this bin alone, found for a
//             task with one record here, is unique and staged as an edge.  A
//             bin that does not fit goes on to the first LDS tier's list.
//   k_resolve one kernel per LDS tier, one wave64 per bin, on the bins passed
//             on: a bin that does not fit a tier (with headroom) goes on to the
//             next one's list.  The bin's fragments, reader log and incidences
//             move to LDS, its records are put in spawn order (rank sort up to
//             64, bitonic on 64-bit keys above) and replayed; W paints over the
//             writer map, an exact R appends to the reader log; the bin is
//             written back compacted.  Candidates are deduplicated per bin in
//             LDS, then either staged in a candidate buffer (throughput form)
//             or inserted into the launch's table at once (direct form: one
//             kernel less per launch)
//   k_dedup   (staged form) candidates -> open-addressing table keyed
//             (pred, succ task), keeping the lowest succ access id
//             (= lowest slot, since LINK runs in slot order); edges known
//             unique skip the table.  Each block lists its new keys and edges
//             in its own part of eslot
//   k_final   edge positions by a prefix over the blocks' counts; table entries
//             and unique edges -> (pred id, succ id) edges; table entries reset;
//             the launch's edge count and error bits to pinned host memory
//
// Because every fragment of a bin carries a uniform state, the edges found
// per bin are exactly the per-point rule; the union over bins followed by
// the (pred, task) dedup is the reference's edge set.

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <cstddef>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>
#include <vector>

#include "../../include/xkrt_hip.h"

#define NONE 0xFFFFFFFFu
#define CSTRIPES 16u   // candidate buffer stripes, each with its own counter (64 B apart)
#define CSTRIPES_C 8u  // stripes [0, 8) hold candidates for the dedup table, [8, 16) edges known
                       // unique where they were found (k_resolve_reg), which bypass the table
#define DIRECT_SLOT 0x80000000u  // eslot entry of such an edge: this bit | its candidate index
#define BLKSLOTS 4096u // per-block statistics slots per tier (4 tiers)
#define EMPTY64 0xFFFFFFFFFFFFFFFFull

// Error bits.  The first group is per domain (a word per domain slot, sticky
// until the domain is reset or its slot is reused): a domain that fails stops
// touching state, the other domains of its pool go on.  The second group is
// per pool: the launch's shared scratch overflowed, so every later kernel of
// the pool stops.
enum : uint32_t {
    ERR_LIST = 2u,      // a bin got more records than the kernel's LDS list holds
    ERR_FRAG = 4u,      // a bin's fragment list overflowed
    ERR_READS = 8u,     // a bin's reader pool overflowed
    ERR_RANGE = 128u,   // a record outside its domain's grid (or naming no domain)
    ERR_CAND = 16u,     // candidate buffer overflow (staged form)
    ERR_HASH = 32u,     // dedup table full
    ERR_EDGES = 64u,    // edge buffer overflow
};
constexpr uint32_t ERR_DOMAIN_BITS = ERR_LIST | ERR_FRAG | ERR_READS | ERR_RANGE;

struct DevCounters {
    unsigned int nrec;       // rect records of this launch (written by k_count)
    unsigned int ntouched;   // bins that got records this launch (listed by k_count; zeroed by k_final)
    unsigned int total;      // incidences allocated this launch
    unsigned int ntier[4];   // bins on each LDS tier's list this launch (escalated from the tier below)
    unsigned int ncand;      // candidates this launch
    unsigned int err;        // sticky ERR_* bits
    unsigned int maxfrag;    // stats: max fragments in a bin
    unsigned int maxrs;      // stats: max reader-pool use in a bin
    unsigned long long edge_total;   // unconsumed edges
    unsigned long long edge_base;    // edge_total when the launch started
    unsigned long long cand_total;   // stats
    unsigned long long frag_visits;  // stats
    unsigned long long nlaunch;      // launches completed (k_final counts them)
};

constexpr uint32_t MAX_DOMS = 256;

// What the host learns of a launch without a stream sync: k_final writes it to
// pinned host memory, slot = launch number % NSNAP.
struct LaunchSnap {
    unsigned long long launch;      // launch number (1, 2, ...)
    unsigned long long edge_total;  // unconsumed edges after it
    unsigned int err;               // sticky pool error bits after it
    unsigned int pad[3];
    unsigned int derr[MAX_DOMS / 4];  // per domain slot, 8 error bits (slot d: word d / 4, byte d % 4)
};
constexpr uint32_t NSNAP = 64;

// one dependency domain's bin grid inside the pool's bin arrays
struct DomDesc {
    uint32_t bin0;            // first pool bin of the grid
    uint32_t nbx, nby;        // bin grid
    uint32_t bxs, bys;        // log2 of the bin extents
    uint32_t ybin0;           // frame y of bin row 0
    uint32_t xmax;            // LDs: dim-0 extent of the frame
    uint32_t pad;
};

struct KArgs {
    const uint4 *rec;         // 2 x uint4 per rect record: {x0, x1, y0, y1}, {acc, tm, domain, 0}
    uint32_t nrec_arg;        // records of this launch, or NONE: read from *nrec_ptr
    const uint32_t *nrec_ptr; // record count uploaded with the records (a header 32 B before them)
    const DomDesc *dom;       // domain table
    const uint32_t *bin_dom;  // domain of each pool bin
    uint32_t *bin_nf, *bin_nrs;
    uint4 *bin_one;           // per bin, 64-bit masks {fragments lo, hi, log entries lo, hi}: the
                              // access of fragment / entry i lies in this bin alone (0: unknown)
    uint32_t *frag;           // nbins * fcap fragments, 5 words each (x0, x1, y0, y1, lw)
    uint32_t *rs;             // nbins * rcap reader-log entries, 5 words each (x0, x1, y0, y1, acc)
    uint32_t fcap, rcap;
    uint32_t *bin_cnt;        // incidences per bin this launch (zeroed by k_resolve)
    uint32_t *touched;        // bins with records this launch, in first-touch order
    uint4 *binrec;  // per bin a bucket of lmax incidences, 2 x uint4 each: {x0, x1, y0, y1}, {acc, tm, record, one}
    uint4 *active;  // per tier, nbins entries {slot, incidence start, record count, nf | ng << 16}
    uint32_t lmax;            // bucket capacity: records per bin and launch of the highest tier
    uint32_t nbins;
    uint32_t max_tier;                              // tiers 0..max_tier; storage sized for max_tier
    uint32_t first_tier;                            // the LDS tier of the bins k_resolve_reg passes on
    uint32_t krpw;                                  // k_count: records per wave (16 or 64)
    uint32_t *cand;           // staged form: CSTRIPES stripes of cand_cap candidates (3 u32 each)
    uint32_t cand_cap;        // per stripe
    uint32_t *cstr;           // CSTRIPES candidate counters, 16 u32 apart
    uint4 *blk;               // per (tier, block): max fragments, max log entries, fragment visits (lo, hi)
    unsigned long long *hkeys;
    uint32_t *hvals;
    uint32_t hmask;
    uint32_t *eslot;
    uint32_t *dcnt;           // staged form: new keys per k_dedup block this launch
    uint32_t *edges;          // (pred id, succ id) per edge
    uint32_t *edom;           // domain of each edge
    unsigned long long edge_cap;
    DevCounters *ctr;
    uint32_t *dom_err;        // MAX_DOMS words: per-domain error bits
    LaunchSnap *snap;         // pinned host ring, NSNAP slots
};

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const uint32_t l = lane_id();
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(v, d, 64);
        if (l >= d) v += y;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_bcast(uint32_t v, int src) { return __shfl(v, src, 64); }

__device__ __forceinline__ uint64_t mix64(uint64_t k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

// ---------------------------------------------------------------- binning

// The bins of record (q0, q1): [bx0, bx1] x [by0, by1] of its domain's grid.
// False when the record names no domain slot, is empty, or leaves its domain's
// grid (records may come from the caller's device memory): its domain is marked
// ERR_RANGE and the record is dropped, so its domain touches no state in this
// launch (k_resolve skips the bins of failed domains).
__device__ __forceinline__ bool rec_bins(const KArgs &a, const uint4 &q0, const uint4 &q1, DomDesc &g,
                                         uint32_t &bx0, uint32_t &bx1, uint32_t &by0, uint32_t &by1)
{
    if (q1.z >= MAX_DOMS) {
        atomicOr(&a.ctr->err, ERR_RANGE);  // no domain to charge: the pool's launch is unsound
        return false;
    }
    g = a.dom[q1.z];
    const uint32_t derr = a.dom_err[q1.z];  // read beside the descriptor: one round trip
    const bool ok = g.nbx != 0 && q0.x < q0.y && q0.y <= g.xmax && q0.z < q0.w && q0.z >= g.ybin0 &&
                    ((q0.w - 1 - g.ybin0) >> g.bys) < g.nby;
    if (!ok) {
        atomicOr(&a.dom_err[q1.z], ERR_RANGE);
        return false;
    }
    if (derr) return false;  // a failed domain: its records are not counted
    bx0 = q0.x >> g.bxs;
    bx1 = (q0.y - 1) >> g.bxs;
    by0 = (q0.z - g.ybin0) >> g.bys;
    by1 = (q0.w - 1 - g.ybin0) >> g.bys;
    return true;
}

// The incidences (record x bin) of 64 consecutive records, spread over the
// wave's lanes: record j (lane j) has n_j = w_j * h_j bins, inclusive prefix
// incl_j; incidence k belongs to the lane j with incl_{j-1} <= k < incl_j and
// is bin (bx0_j + off % w_j, by0_j + off / w_j), off = k - incl_{j-1}.  So a
// record covering many bins does not stall its wave.
// k_count's records per wave (KArgs::krpw): 16 spreads a small launch over more waves; a large
// launch takes 64, a lane's worth, since 16 leave most lanes idle when records cover one bin
constexpr uint32_t KRPW_SMALL = 16, KRPW_LARGE = 64;
constexpr uint64_t KRPW_LARGE_FROM = 1u << 17;  // records of a launch from which it takes 64

struct WaveRecs {
    uint32_t r, valid, bx0, w, by0, n, incl, bin0, nbx;
    uint32_t one;  // the record's access is this one record, in one bin
    uint4 q0;      // the record, read once (a small launch reads it from host memory)
    uint32_t acc, tm;
};

// Lane j takes record r = (wave's first) + j; this wave's records are those below
// nrec, the launch's below nall.  Every record is read once: its neighbours' access
// ids (an access's records are adjacent: a wrapped tile has two, a segment up to
// three) come from the adjacent lanes, and only the wave's two outer neighbours are
// extra loads, issued beside the records' own.
__device__ __forceinline__ WaveRecs wave_recs(const KArgs &a, uint32_t r, uint32_t nrec, uint32_t nall)
{
    WaveRecs q;
    q.r = r;
    q.valid = r < nrec;
    q.bx0 = q.w = q.by0 = q.n = q.bin0 = q.nbx = q.one = 0;
    uint4 q0 = make_uint4(0, 0, 0, 0), q1 = make_uint4(0, 0, 0, 0);
    uint32_t pa = NONE, na = NONE;
    const bool head = q.valid && lane_id() == 0, tail = q.valid && r + 1 == nrec;
    if (q.valid) {
        q0 = a.rec[2ull * r];
        q1 = a.rec[2ull * r + 1];
    }
    if (head && r > 0) pa = a.rec[2ull * (r - 1) + 1].x;
    if (tail && r + 1 < nall) na = a.rec[2ull * (r + 1) + 1].x;
    const uint32_t up = __shfl(q1.x, (int)((lane_id() + 63) & 63), 64);
    const uint32_t dn = __shfl(q1.x, (int)((lane_id() + 1) & 63), 64);
    if (!head) pa = up;
    if (!tail) na = dn;
    q.q0 = q0;
    q.acc = q1.x;
    q.tm = q1.y;
    if (q.valid) {
        DomDesc g;
        uint32_t bx0, bx1, by0, by1;
        if (rec_bins(a, q0, q1, g, bx0, bx1, by0, by1)) {
            q.bx0 = bx0; q.w = bx1 - bx0 + 1; q.by0 = by0; q.n = q.w * (by1 - by0 + 1);
            q.bin0 = g.bin0; q.nbx = g.nbx;
            q.one = q.n == 1 && pa != q1.x && na != q1.x ? 1u : 0u;
        }
    }
    q.incl = wave_incl_scan(q.n);
    return q;
}

// owner lane of incidence k and the bin it names (all lanes call it together)
__device__ __forceinline__ uint32_t wave_rec_bin(const WaveRecs &q, uint32_t k, uint32_t &owner)
{
    uint32_t lo = 0;
#pragma unroll
    for (uint32_t step = 32; step >= 1; step >>= 1) {
        const uint32_t v = __shfl(q.incl, (int)(lo + step - 1), 64);
        if (v <= k) lo += step;
    }
    owner = lo > 63 ? 63 : lo;
    const uint32_t incl = __shfl(q.incl, (int)owner, 64), n = __shfl(q.n, (int)owner, 64);
    const uint32_t bx0 = __shfl(q.bx0, (int)owner, 64), w = __shfl(q.w, (int)owner, 64);
    const uint32_t by0 = __shfl(q.by0, (int)owner, 64), bin0 = __shfl(q.bin0, (int)owner, 64);
    const uint32_t nbx = __shfl(q.nbx, (int)owner, 64);
    const uint32_t off = k - (incl - n);
    return bin0 + (by0 + off / max(w, 1u)) * nbx + bx0 + off % max(w, 1u);
}

// k_count's LDS: each wave's first-touched bins, listed with one update of the
// touched-list counter per block (or per TBUF bins of a wave), then the four
// waves' counts and the block's list position.  One update per 64 incidences
// kept the waves of a large launch queued on that one address.
constexpr uint32_t TBUF = 256;
constexpr uint32_t COUNT_LDS = 4 * TBUF * 4 + 8 * 4;  // 256 threads = 4 waves

// this wave's buffered first-touched bins to the touched list
__device__ __forceinline__ void flush_touched(const KArgs &a, uint32_t *tb, uint32_t n)
{
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(&a.ctr->ntouched, n);
    base = wave_bcast(base, 0);
    for (uint32_t i = lane_id(); i < n; i += 64) a.touched[base + i] = tb[i];
    __builtin_amdgcn_wave_barrier();  // the buffer is read before any lane refills it
}

__global__ void __launch_bounds__(256) k_count(KArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    uint32_t *tb = reinterpret_cast<uint32_t *>(lds_raw) + (threadIdx.x >> 6) * TBUF;
    uint32_t ntb = 0;  // wave-uniform: entries of tb in use
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid == 0) {
        a.ctr->total = 0;
        for (int t = 0; t < 4; ++t) {
            a.ctr->ntier[t] = 0;
        }
        a.ctr->ncand = 0;
        a.ctr->edge_base = a.ctr->edge_total;
        for (uint32_t q = 0; q < CSTRIPES; ++q) a.cstr[16 * q] = 0;
    }
    const uint32_t nrec = a.nrec_arg != NONE ? a.nrec_arg : *a.nrec_ptr;
    if (tid == 0) a.ctr->nrec = nrec;  // for the later kernels
    if (a.ctr->err) return;  // a poisoned pool touches no state
    const uint32_t lane = lane_id();
    // krpw records per wave: their incidences (several per record in general) are
    // spread over the wave's 64 lanes
    const uint32_t wave = tid >> 6, nwaves = (gridDim.x * blockDim.x) >> 6, krpw = a.krpw;
    for (uint32_t r0 = wave * krpw; r0 < nrec; r0 += nwaves * krpw) {
        const WaveRecs q = wave_recs(a, r0 + lane, lane < krpw ? min(nrec, r0 + krpw) : 0u, nrec);
        const uint32_t total = __shfl(q.incl, 63, 64);
        for (uint32_t k0 = 0; k0 < total; k0 += 64) {
            const uint32_t k = k0 + lane;
            uint32_t owner;
            const uint32_t bin = wave_rec_bin(q, min(k, total - 1), owner);
            const uint32_t one = __shfl(q.one, (int)owner, 64);
            // the owner's record, from its lane
            const uint4 q0 = make_uint4(__shfl(q.q0.x, (int)owner, 64), __shfl(q.q0.y, (int)owner, 64),
                                        __shfl(q.q0.z, (int)owner, 64), __shfl(q.q0.w, (int)owner, 64));
            const uint32_t acc = __shfl(q.acc, (int)owner, 64), tm = __shfl(q.tm, (int)owner, 64);
            bool first = false;
            if (k < total) {
                // the incidence goes straight into its bin's bucket; a bin with more than lmax
                // incidences fits no tier and fails in k_resolve
                const uint32_t r = r0 + owner;
                const uint32_t pos = atomicAdd(&a.bin_cnt[bin], 1u);
                first = pos == 0;
                if (pos < a.lmax) {
                    uint4 *o = a.binrec + 2ull * ((size_t)bin * a.lmax + pos);
                    o[0] = q0;
                    o[1] = make_uint4(acc, tm, r, one);
                }
            }
            // the bins this launch touches, listed once each
            const uint64_t m = __ballot(first);
            if (m) {
                const uint32_t nm = (uint32_t)__popcll(m);
                if (ntb + nm > TBUF) {
                    flush_touched(a, tb, ntb);
                    ntb = 0;
                }
                if (first) tb[ntb + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = bin;
                ntb += nm;
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    // the block's buffered bins: one counter update, each wave copies its own
    uint32_t *bc = reinterpret_cast<uint32_t *>(lds_raw) + 4 * TBUF;
    const uint32_t w = threadIdx.x >> 6;
    if (lane == 0) bc[w] = ntb;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t tot = bc[0] + bc[1] + bc[2] + bc[3];
        bc[4] = tot ? atomicAdd(&a.ctr->ntouched, tot) : 0u;
    }
    __syncthreads();
    uint32_t base = bc[4];
    for (uint32_t u = 0; u < w; ++u) base += bc[u];
    for (uint32_t i = lane; i < ntb; i += 64) a.touched[base + i] = tb[i];
}

// ---------------------------------------------------------------- resolve
//
// Bin state: a painter map of the writers -- fragments tiling the bin, each
// with a uniform last writer `lw` (access id, NONE = never written) -- and a
// log of the exact readers (rect clipped to the bin, access id).  Access ids
// grow in spawn order, so a logged reader r is among the readers-since-the-
// last-writer of point x iff acc(r) > lw(x) and x is in r.

// Per-tier LDS sizes besides (fragments, log entries, records): the (pred,
// task) dedup set, the candidate staging buffer (3 u32 each) and the readers
// of one fragment kept for the exact coverage test; the small tier keeps
// them small so that more bins are in flight per CU.
template <int FCAP, int GCAP, int LCAP, bool DIR>
struct BinLds {
    static constexpr uint32_t DSZ = LCAP >= 1024 ? 1024u : (LCAP >= 512 ? 512u : (LCAP >= 256 ? 256u : 64u));
    static constexpr uint32_t CCAP = LCAP >= 256 ? 256u : 64u;
    static constexpr uint32_t KCAP = LCAP >= 256 ? 128u : 64u;
    uint32_t fx0[2][FCAP], fx1[2][FCAP], fy0[2][FCAP], fy1[2][FCAP], flw[2][FCAP];
    uint32_t gx0[GCAP], gx1[GCAP], gy0[GCAP], gy1[GCAP], gacc[GCAP];
    unsigned long long L[LCAP];  // sort keys: record index << 32 | staging position
    // the bin's records in spawn order, clipped to the bin
    uint32_t rx0[LCAP], rx1[LCAP], ry0[LCAP], ry1[LCAP], racc[LCAP], rtm[LCAP];
    unsigned long long D[DSZ];  // (pred, task) keys of the current run / group
    uint32_t Dv[DSZ];           // lowest successor access id found for the key
    uint32_t Du[DSZ];           // slots of D in use
    uint32_t C[3 * CCAP];
    uint32_t kx0[KCAP], kx1[KCAP], ky0[KCAP], ky1[KCAP];
    uint32_t cx[KCAP + 1], cy[KCAP + 1];
    uint32_t ovl[FCAP];         // a writer's overlapping fragments
    uint32_t okc[FCAP];         // per overlapping fragment: qualifying readers | 1 << 31 if one covers f n rect
    uint32_t novl;
    uint32_t ccount;
    uint32_t dcount;            // entries of Du
    uint32_t tdom;              // domain << 24: or-ed into task numbers of candidates
};

struct BinState {
    uint32_t nf;      // fragments in the current bank
    uint32_t fb;      // current fragment bank
    uint32_t ng;      // reader-log entries
    uint32_t err;     // local error bits
    uint32_t slot;    // the bin
    uint32_t ng0;     // log entries in HBM when the bin started
    uint32_t gload;   // 1 once entries [0, ng0) are in LDS (they are loaded on first use)
    uint32_t gdirty;  // LDS log entries from this index on differ from HBM
    uint32_t fdirty;  // 1 when the fragments must be written back
};

// The reader log is read only by a writer's LINK and by pruning; bins whose
// records only read append to it without loading it.
template <int FCAP, int GCAP, int LCAP, bool DIR>
__device__ void ensure_log(const KArgs &a, BinLds<FCAP, GCAP, LCAP, DIR> &s, BinState &st)
{
    if (st.gload) return;
    const uint32_t *gr = a.rs + (size_t)st.slot * a.rcap * 5;
    for (uint32_t i = lane_id(); i < st.ng0; i += 64) {
        const uint32_t *q = gr + 5 * i;
        const uint32_t g0 = q[0], g1 = q[1], g2 = q[2], g3 = q[3], g4 = q[4];
        s.gx0[i] = g0; s.gx1[i] = g1; s.gy0[i] = g2; s.gy1[i] = g3; s.gacc[i] = g4;
    }
    st.gload = 1;
    __syncthreads();
}

// Inserts candidate (pred p, successor access acc, task with domain bits)
// into the launch's dedup table, keeping the lowest acc per key.  Returns the
// table slot when the key is new, NONE otherwise.
__device__ __forceinline__ uint32_t hash_insert(const KArgs &a, uint32_t p, uint32_t acc, uint32_t task)
{
    const unsigned long long key = ((unsigned long long)p << 32) | task;
    const uint32_t mask = a.hmask;
    uint32_t h = (uint32_t)mix64(key) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const unsigned long long old = atomicCAS(&a.hkeys[h], EMPTY64, key);
        if (old == EMPTY64 || old == key) {
            atomicMin(&a.hvals[h], acc);
            return old == EMPTY64 ? h : NONE;
        }
        h = (h + 1) & mask;
    }
    atomicOr(&a.ctr->err, ERR_HASH);
    return NONE;
}

// the new key in table slot h takes edge position pos of this launch
__device__ __forceinline__ void edge_at(const KArgs &a, unsigned long long pos, uint32_t h)
{
    const unsigned long long ebase = a.ctr->edge_base;
    if (pos >= a.edge_cap || pos - ebase > a.hmask) atomicOr(&a.ctr->err, ERR_EDGES);
    else a.eslot[pos - ebase] = h;
}

// the staged candidates of this block into the candidate buffer (staged
// form) or the dedup table (direct form: one edge counter update per 64)
template <bool DIR, class S>
__device__ void flush_cands_to(const KArgs &a, S &s)
{
    __syncthreads();
    const uint32_t n = s.ccount;
    const uint32_t lane = lane_id();
    const uint32_t stripe = blockIdx.x % CSTRIPES_C;
    uint32_t base = 0;
    if (n && lane == 0) base = atomicAdd(&a.cstr[16 * stripe], n);
    if (!DIR && n) {  // staged form: the candidates go to the buffer for k_dedup
        base = wave_bcast(base, 0);
        if ((uint64_t)base + n > a.cand_cap) {
            if (lane == 0) atomicOr(&a.ctr->err, ERR_CAND);
        } else {
            uint32_t *dst = a.cand + 3ull * ((size_t)stripe * a.cand_cap + base);
            for (uint32_t i = lane; i < 3 * n; i += 64) dst[i] = s.C[i];
        }
    }
    for (uint32_t i0 = 0; DIR && i0 < n; i0 += 64) {
        const uint32_t i = i0 + lane;
        const uint32_t h = i < n ? hash_insert(a, s.C[3 * i], s.C[3 * i + 1], s.C[3 * i + 2]) : NONE;
        const uint64_t nm = __ballot(h != NONE);
        if (nm == 0) continue;
        uint32_t blo = 0, bhi = 0;
        if (lane == 0) {
            const unsigned long long b0 = atomicAdd(&a.ctr->edge_total, (unsigned long long)__popcll(nm));
            blo = (uint32_t)b0;
            bhi = (uint32_t)(b0 >> 32);
        }
        blo = __shfl(blo, 0, 64);
        bhi = __shfl(bhi, 0, 64);
        if (h != NONE)
            edge_at(a, (((unsigned long long)bhi << 32) | blo) + (unsigned long long)__popcll(nm & ((1ull << lane) - 1ull)),
                    h);
    }
    __syncthreads();
    if (lane == 0) s.ccount = 0;
    __syncthreads();
}

template <int FCAP, int GCAP, int LCAP, bool DIR>
__device__ __forceinline__ void flush_cands(const KArgs &a, BinLds<FCAP, GCAP, LCAP, DIR> &s)
{
    flush_cands_to<DIR>(a, s);
}

// room for m more candidates in the staging buffer.  It is flushed only when
// full: every flush takes a position from a stripe counter that many blocks
// share, and a flush per bin put that round trip on every bin's path.
template <int FCAP, int GCAP, int LCAP, bool DIR>
__device__ __forceinline__ void room(const KArgs &a, BinLds<FCAP, GCAP, LCAP, DIR> &s, uint32_t m)
{
    __syncthreads();
    if (s.ccount + m > BinLds<FCAP, GCAP, LCAP, DIR>::CCAP) flush_cands(a, s);
}

// Record predecessor p of access acc (task `task`): the run's dedup set keeps
// the lowest successor access per (pred, task) -- LINK runs in slot order in
// the reference, so the lowest access id is the slot that owns the edge.  When
// the set is full the candidate goes straight to the global candidate buffer
// (the cross-bin dedup keeps the minimum too).
template <int FCAP, int GCAP, int LCAP, bool DIR>
__device__ __forceinline__ void emit(const KArgs &a, BinLds<FCAP, GCAP, LCAP, DIR> &s, uint32_t p, uint32_t acc,
                                     uint32_t task)
{
    task |= s.tdom;  // (pred, task) keys of different domains never meet
    const unsigned long long key = ((unsigned long long)p << 32) | task;
    constexpr uint32_t DSZ = BinLds<FCAP, GCAP, LCAP, DIR>::DSZ;
    uint32_t h = (uint32_t)(mix64(key) & (DSZ - 1));
    for (uint32_t probe = 0; probe < 32; ++probe) {
        const unsigned long long old = atomicCAS(&s.D[h], EMPTY64, key);
        if (old == EMPTY64 || old == key) {
            atomicMin(&s.Dv[h], acc);
            if (old == EMPTY64) s.Du[atomicAdd(&s.dcount, 1u)] = h;
            return;
        }
        h = (h + 1) & (DSZ - 1);
    }
    // the bin's set is crowded: this candidate goes on alone
    const uint32_t stripe = blockIdx.x % CSTRIPES_C;
    const uint32_t pos = atomicAdd(&a.cstr[16 * stripe], 1u);
    if (DIR) {
        const uint32_t slot = hash_insert(a, p, acc, task);
        if (slot != NONE) edge_at(a, atomicAdd(&a.ctr->edge_total, 1ull), slot);
    } else if (pos < a.cand_cap) {
        uint32_t *dst = a.cand + 3ull * ((size_t)stripe * a.cand_cap + pos);
        dst[0] = p;
        dst[1] = acc;
        dst[2] = task;
    } else {
        atomicOr(&a.ctr->err, ERR_CAND);
    }
}

// move the run's (pred, task, lowest successor) entries to the staging buffer
template <int FCAP, int GCAP, int LCAP, bool DIR>
__device__ void drain_dedup(const KArgs &a, BinLds<FCAP, GCAP, LCAP, DIR> &s)
{
    __syncthreads();
    const uint32_t n = s.dcount;
    for (uint32_t c = 0; c < n; c += 64) {
        room(a, s, min(64u, n - c));
        const uint32_t u = c + lane_id();
        if (u < n) {
            const uint32_t h = s.Du[u];
            const unsigned long long key = s.D[h];
            const uint32_t pos = s.ccount + u - c;
            s.C[3 * pos] = (uint32_t)(key >> 32);
            s.C[3 * pos + 1] = s.Dv[h];
            s.C[3 * pos + 2] = (uint32_t)key;
            s.D[h] = EMPTY64;
            s.Dv[h] = NONE;
        }
        __syncthreads();
        if (lane_id() == 0) s.ccount += min(64u, n - c);
        __syncthreads();
    }
    if (lane_id() == 0) s.dcount = 0;
    __syncthreads();
}

__device__ __forceinline__ bool overlaps(uint32_t ax0, uint32_t ax1, uint32_t ay0, uint32_t ay1, uint32_t bx0,
                                         uint32_t bx1, uint32_t by0, uint32_t by1)
{
    return ax0 < bx1 && bx0 < ax1 && ay0 < by1 && by0 < ay1;
}

// LINK (dependency-tree.hpp:311-355).  Per point x of the rect: a writer links
// to the readers since the last writer of x when there are some, else to the
// last writer; any other access links to the last writer.
template <int FCAP, int GCAP, int LCAP, bool DIR>
__device__ void do_link(const KArgs &a, BinLds<FCAP, GCAP, LCAP, DIR> &s, BinState &st, uint32_t X0, uint32_t X1,
                        uint32_t Y0, uint32_t Y1, uint32_t mode, uint32_t acc, uint32_t task)
{
    const uint32_t fb = st.fb;
    const uint32_t lane = lane_id();
    if (!(mode & 2u)) {
        for (uint32_t c = 0; c < st.nf; c += 64) {
            const uint32_t f = c + lane;
            if (f < st.nf && s.flw[fb][f] != NONE &&
                overlaps(s.fx0[fb][f], s.fx1[fb][f], s.fy0[fb][f], s.fy1[fb][f], X0, X1, Y0, Y1))
                emit(a, s, s.flw[fb][f], acc, task);
        }
        __syncthreads();
        return;
    }
    // writer.  (1) the fragments f overlapping the rect, compacted; (2) one lane
    // per logged reader g: g is a predecessor iff it touches some Q = f n rect
    // with acc(g) > lw(f) -- emitted once -- and it counts towards Q; (3) lw(f)
    // is a predecessor iff some point of Q has no such reader: none touches Q,
    // or (not one covers Q and) the exact coverage test finds an uncovered point.
    ensure_log(a, s, st);
    if (lane == 0) s.novl = 0;
    __syncthreads();
    for (uint32_t c = 0; c < st.nf; c += 64) {
        const uint32_t f = c + lane;
        const bool ov = f < st.nf && overlaps(s.fx0[fb][f], s.fx1[fb][f], s.fy0[fb][f], s.fy1[fb][f], X0, X1, Y0, Y1);
        const uint64_t m = __ballot(ov);
        if (ov) {
            const uint32_t k = s.novl + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            s.ovl[k] = f;
            s.okc[k] = 0;
        }
        __syncthreads();
        if (lane == 0) s.novl += (uint32_t)__popcll(m);
        __syncthreads();
    }
    const uint32_t no = s.novl;
    // one lane per (logged reader, overlapping fragment) pair, reader-major: a lane per reader
    // walking every overlapping fragment left most lanes idle on bins with few readers and
    // many fragments, each step a dependent LDS read
    const uint32_t npairs = st.ng * no;
    for (uint32_t t0 = 0; t0 < npairs; t0 += 64) {
        const uint32_t t = t0 + lane;
        bool pred = false;
        uint32_t ga = 0;
        if (t < npairs) {
            const uint32_t g = t / no, k = t - g * no;
            const uint32_t gx0 = s.gx0[g], gx1 = s.gx1[g], gy0 = s.gy0[g], gy1 = s.gy1[g];
            ga = s.gacc[g];
            const uint32_t f = s.ovl[k], lw = s.flw[fb][f];
            if (lw == NONE || ga > lw) {
                const uint32_t qx0 = max(s.fx0[fb][f], X0), qx1 = min(s.fx1[fb][f], X1);
                const uint32_t qy0 = max(s.fy0[fb][f], Y0), qy1 = min(s.fy1[fb][f], Y1);
                if (overlaps(gx0, gx1, gy0, gy1, qx0, qx1, qy0, qy1)) {
                    pred = true;
                    const bool full = gx0 <= qx0 && qx1 <= gx1 && gy0 <= qy0 && qy1 <= gy1;
                    atomicAdd(&s.okc[k], 1u);
                    if (full) atomicOr(&s.okc[k], 1u << 31);
                }
            }
        }
        // a reader is emitted once per pass over its pairs; its repeats across lanes and
        // passes collapse in the (pred, task) set
        const uint64_t pm = __ballot(pred);
        const uint32_t prev = __shfl(ga, (int)(lane + 63u) & 63, 64);  // the lane below (all lanes take part)
        if (pred && (lane == 0 || !((pm >> (lane - 1)) & 1ull) || prev != ga)) emit(a, s, ga, acc, task);
    }
    __syncthreads();
    for (uint32_t c = 0; c < no; c += 64) {
        const uint32_t k = c + lane;
        uint32_t lw = NONE, kc = 0;
        if (k < no) {
            lw = s.flw[fb][s.ovl[k]];
            kc = s.okc[k];
        }
        if (k < no && lw != NONE && kc == 0) emit(a, s, lw, acc, task);
        // partially covered fragments: exact test, one fragment at a time
        uint64_t m = __ballot(k < no && lw != NONE && kc != 0 && !(kc >> 31));
        while (m) {
            const uint32_t ff = s.ovl[c + (uint32_t)__builtin_ctzll(m)];
            m &= m - 1;
            const uint32_t rx0 = max(s.fx0[fb][ff], X0), rx1 = min(s.fx1[fb][ff], X1);
            const uint32_t ry0 = max(s.fy0[fb][ff], Y0), ry1 = min(s.fy1[fb][ff], Y1);
            const uint32_t rlw = s.flw[fb][ff];
            uint32_t nk = 0;
            for (uint32_t g0 = 0; g0 < st.ng; g0 += 64) {
                const uint32_t g = g0 + lane;
                const bool in = g < st.ng && s.gacc[g] > rlw &&
                                overlaps(s.gx0[g], s.gx1[g], s.gy0[g], s.gy1[g], rx0, rx1, ry0, ry1);
                const uint64_t km = __ballot(in);
                if (in) {
                    const uint32_t kk = nk + (uint32_t)__popcll(km & ((1ull << lane) - 1ull));
                    if (kk < BinLds<FCAP, GCAP, LCAP, DIR>::KCAP) {
                        s.kx0[kk] = max(s.gx0[g], rx0); s.kx1[kk] = min(s.gx1[g], rx1);
                        s.ky0[kk] = max(s.gy0[g], ry0); s.ky1[kk] = min(s.gy1[g], ry1);
                    }
                }
                nk += (uint32_t)__popcll(km);
            }
            __syncthreads();  // kx/ky rows are read by other lanes below
            if (nk > BinLds<FCAP, GCAP, LCAP, DIR>::KCAP) {
                st.err |= ERR_READS;
                return;
            }
            // If Q minus the readers is not empty it holds a point (cx, cy) with
            // cx in {qx0} U {reader x1} and cy in {qy0} U {reader y1}: the lowest
            // uncovered point of the leftmost uncovered column.
            const uint32_t nc = nk + 1;
            if (lane == 0) { s.cx[0] = rx0; s.cy[0] = ry0; }
            for (uint32_t q = lane; q < nk; q += 64) {
                s.cx[q + 1] = s.kx1[q] < rx1 ? s.kx1[q] : rx0;
                s.cy[q + 1] = s.ky1[q] < ry1 ? s.ky1[q] : ry0;
            }
            __syncthreads();
            bool unc = false;
            for (uint32_t p = lane; p < nc * nc && !unc; p += 64) {
                const uint32_t px = s.cx[p % nc], py = s.cy[p / nc];
                bool cov = false;
                for (uint32_t q = 0; q < nk && !cov; ++q)
                    cov = s.kx0[q] <= px && px < s.kx1[q] && s.ky0[q] <= py && py < s.ky1[q];
                unc = !cov;
            }
            if (__ballot(unc) != 0 && lane == 0) emit(a, s, rlw, acc, task);
            __syncthreads();
        }
    }
    __syncthreads();
}

template <int FCAP, int GCAP, int LCAP, bool DIR>
__device__ void prune_log(const KArgs &a, BinLds<FCAP, GCAP, LCAP, DIR> &s, BinState &st);

// A run of whole task groups none of whose records writes: LINK of a
// non-writer reads only the writer map, and PUT of a reader only appends to
// the log, so every record of the run links in parallel (one per lane) and
// the exact readers are then logged in spawn order.
template <int FCAP, int GCAP, int LCAP, bool DIR>
__device__ void do_run(const KArgs &a, BinLds<FCAP, GCAP, LCAP, DIR> &s, BinState &st, uint32_t i0, uint32_t i1)
{
    const uint32_t fb = st.fb;
    const uint32_t lane = lane_id();
    // one lane per (record, fragment) pair, record-major: a lane per record walking every
    // fragment left most lanes idle on short runs
    const uint32_t nf = st.nf, npairs = (i1 - i0) * nf;
    for (uint32_t t0 = 0; t0 < npairs; t0 += 64) {
        const uint32_t t = t0 + lane;
        if (t < npairs) {
            const uint32_t r = t / nf, f = t - r * nf, q = i0 + r;
            const uint32_t lw = s.flw[fb][f];
            if (lw != NONE &&
                overlaps(s.fx0[fb][f], s.fx1[fb][f], s.fy0[fb][f], s.fy1[fb][f], s.rx0[q], s.rx1[q], s.ry0[q], s.ry1[q]))
                emit(a, s, lw, s.racc[q], s.rtm[q] >> 8);
        }
    }
    __syncthreads();
    // log the exact readers
    for (uint32_t c = i0; c < i1; c += 64) {
        const uint32_t q = c + lane;
        const bool rd = q < i1 && (s.rtm[q] & 0xffu) == 1u;
        const uint64_t m = __ballot(rd);
        const uint32_t n = (uint32_t)__popcll(m);
        if (!n) continue;
        if (st.ng + n > (uint32_t)GCAP) {
            prune_log(a, s, st);
            if (st.ng + n > (uint32_t)GCAP) {
                st.err |= ERR_READS;
                return;
            }
        }
        if (rd) {
            const uint32_t g = st.ng + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            s.gx0[g] = s.rx0[q]; s.gx1[g] = s.rx1[q]; s.gy0[g] = s.ry0[q]; s.gy1[g] = s.ry1[q];
            s.gacc[g] = s.racc[q];
        }
        st.ng += n;
        __syncthreads();
    }
}

// PUT of a writer (dependency-tree.hpp:238-242): the rect's points get
// last_write = acc.  Fragments under the rect are cut away (up to 4 guillotine
// pieces of f minus the rect survive, keeping f's writer) and the rect,
// clipped to the bin, becomes one fragment.  Logged readers are not touched:
// their ids are now older than the rect's writer.
template <int FCAP, int GCAP, int LCAP, bool DIR>
__device__ void do_paint(BinLds<FCAP, GCAP, LCAP, DIR> &s, BinState &st, uint32_t X0, uint32_t X1, uint32_t Y0,
                         uint32_t Y1, uint32_t acc)
{
    const uint32_t fb = st.fb, nb = fb ^ 1u;
    uint32_t base = 0;
    for (uint32_t c = 0; c < st.nf; c += 64) {
        const uint32_t f = c + lane_id();
        uint32_t fx0 = 0, fx1 = 0, fy0 = 0, fy1 = 0, lw = NONE;
        bool live = f < st.nf, ov = false;
        if (live) {
            fx0 = s.fx0[fb][f]; fx1 = s.fx1[fb][f]; fy0 = s.fy0[fb][f]; fy1 = s.fy1[fb][f];
            lw = s.flw[fb][f];
            ov = overlaps(fx0, fx1, fy0, fy1, X0, X1, Y0, Y1);
        }
        // pieces of f minus the rect: top, bottom, then left and right of the middle band
        const bool pt = ov && fx0 < X0, pbm = ov && X1 < fx1, pl = ov && fy0 < Y0, pr = ov && Y1 < fy1;
        const uint32_t npc = live ? (ov ? (uint32_t)pt + pbm + pl + pr : 1u) : 0u;
        const uint32_t incl = wave_incl_scan(npc);
        const uint32_t tot = wave_bcast(incl, 63);
        if (base + tot + 1 > (uint32_t)FCAP) {
            st.err |= ERR_FRAG;
            return;
        }
        uint32_t o = base + incl - npc;
        if (live && !ov) {
            s.fx0[nb][o] = fx0; s.fx1[nb][o] = fx1; s.fy0[nb][o] = fy0; s.fy1[nb][o] = fy1; s.flw[nb][o] = lw;
        } else if (ov) {
            const uint32_t mx0 = max(fx0, X0), mx1 = min(fx1, X1);
            if (pt) { s.fx0[nb][o] = fx0; s.fx1[nb][o] = X0; s.fy0[nb][o] = fy0; s.fy1[nb][o] = fy1; s.flw[nb][o] = lw; ++o; }
            if (pbm) { s.fx0[nb][o] = X1; s.fx1[nb][o] = fx1; s.fy0[nb][o] = fy0; s.fy1[nb][o] = fy1; s.flw[nb][o] = lw; ++o; }
            if (pl) { s.fx0[nb][o] = mx0; s.fx1[nb][o] = mx1; s.fy0[nb][o] = fy0; s.fy1[nb][o] = Y0; s.flw[nb][o] = lw; ++o; }
            if (pr) { s.fx0[nb][o] = mx0; s.fx1[nb][o] = mx1; s.fy0[nb][o] = Y1; s.fy1[nb][o] = fy1; s.flw[nb][o] = lw; ++o; }
        }
        base += tot;
    }
    if (lane_id() == 0) {
        s.fx0[nb][base] = X0; s.fx1[nb][base] = X1; s.fy0[nb][base] = Y0; s.fy1[nb][base] = Y1;
        s.flw[nb][base] = acc;
    }
    st.nf = base + 1;
    st.fb = nb;
    st.fdirty = 1;
    __syncthreads();
}

// drop the logged readers that no point still sees: every fragment they
// touch has a newer writer
template <int FCAP, int GCAP, int LCAP, bool DIR>
__device__ void prune_log(const KArgs &a, BinLds<FCAP, GCAP, LCAP, DIR> &s, BinState &st)
{
    ensure_log(a, s, st);
    st.gdirty = 0;  // compaction moves entries
    const uint32_t fb = st.fb;
    uint32_t base = 0;
    for (uint32_t g0 = 0; g0 < st.ng; g0 += 64) {
        const uint32_t g = g0 + lane_id();
        uint32_t x0 = 0, x1 = 0, y0 = 0, y1 = 0, ga = 0;
        bool alive = false;
        if (g < st.ng) {
            x0 = s.gx0[g]; x1 = s.gx1[g]; y0 = s.gy0[g]; y1 = s.gy1[g]; ga = s.gacc[g];
            for (uint32_t f = 0; f < st.nf && !alive; ++f) {
                const uint32_t lw = s.flw[fb][f];
                alive = (lw == NONE || lw < ga) && overlaps(s.fx0[fb][f], s.fx1[fb][f], s.fy0[fb][f], s.fy1[fb][f],
                                                            x0, x1, y0, y1);
            }
        }
        __syncthreads();
        const uint32_t incl = wave_incl_scan(alive ? 1u : 0u);
        const uint32_t tot = wave_bcast(incl, 63);
        if (alive) {
            const uint32_t o = base + incl - 1;
            s.gx0[o] = x0; s.gx1[o] = x1; s.gy0[o] = y0; s.gy1[o] = y1; s.gacc[o] = ga;
        }
        base += tot;
        __syncthreads();
    }
    st.ng = base;
}

// PUT of an exact reader (dependency-tree.hpp:243-244): log it
template <int FCAP, int GCAP, int LCAP, bool DIR>
__device__ void do_log(const KArgs &a, BinLds<FCAP, GCAP, LCAP, DIR> &s, BinState &st, uint32_t X0, uint32_t X1, uint32_t Y0,
                       uint32_t Y1, uint32_t acc)
{
    if (st.ng >= (uint32_t)GCAP) {
        prune_log(a, s, st);
        if (st.ng >= (uint32_t)GCAP) {
            st.err |= ERR_READS;
            return;
        }
    }
    if (lane_id() == 0) {
        const uint32_t g = st.ng;
        s.gx0[g] = X0; s.gx1[g] = X1; s.gy0[g] = Y0; s.gy1[g] = Y1; s.gacc[g] = acc;
    }
    st.ng++;
    __syncthreads();
}

// A bin that outgrows this tier's LDS capacities is abandoned -- nothing is
// written back, its record list stays -- and queued for the next tier; the
// candidates it staged before are a subset of what the retry emits, and the
// cross-bin dedup absorbs the repeats.  The last tier reports the overflow.
template <int FCAP, int GCAP, int LCAP, bool DIR>
__device__ __forceinline__ void escalate(const KArgs &a, uint32_t tier, uint4 e, uint32_t err)
{
    if (lane_id() != 0) return;
    if (tier < a.max_tier) {
        // tier 1 has tier 0's fragment and log room and only more record room; in a launch
        // without a tier-1 kernel tier 0 passes its bins straight to tier 2
        const uint32_t next = tier + 1;
        const uint32_t pos = atomicAdd(&a.ctr->ntier[next], 1u);
        a.active[(size_t)next * a.nbins + pos] = e;
    } else {
        atomicOr(&a.dom_err[e.w >> 24], err);  // the bin's domain fails; the pool's other domains go on
        a.bin_cnt[e.x] = 0;
    }
}

// the first 64 records, fragments and log entries of a bin (one per lane),
// loaded ahead of the bin
struct BinPre {
    uint4 q0, q1;
    uint32_t f0, f1, f2, f3, f4;
    uint32_t g0, g1, g2, g3, g4;
    uint32_t mf, mg;  // this lane's word of the bin's single-bin marks (with the log)
};

// With `log`, the reader log of a bin of at most 64 entries comes too (the
// register replay below takes it whole); otherwise the log is loaded on first
// use (ensure_log).
__device__ __forceinline__ void bin_pre_load(const KArgs &a, uint4 e, BinPre &p, bool log)
{
    const uint32_t i = lane_id();
    const uint32_t slot = e.x, cnt = e.z, nf0 = e.w & 0x7ffu, ng0 = (e.w >> 11) & 0xfffu;
    p.q0 = make_uint4(0, 0, 0, 0);
    p.q1 = make_uint4(NONE, 0, NONE, 0);
    p.f0 = p.f1 = p.f2 = p.f3 = p.f4 = 0;
    p.g0 = p.g1 = p.g2 = p.g3 = p.g4 = 0;
    p.mf = p.mg = 0;
    if (log && (cnt || nf0)) {
        const uint32_t *w = reinterpret_cast<const uint32_t *>(a.bin_one + slot);
        p.mf = w[i >> 5];
        p.mg = w[2 + (i >> 5)];
    }
    if (i < cnt && i < a.lmax) {
        const uint4 *ir = a.binrec + 2ull * ((size_t)slot * a.lmax + i);
        p.q0 = ir[0];
        p.q1 = ir[1];
    }
    if (i < nf0 && nf0 <= a.fcap) {
        const uint32_t *q = a.frag + (size_t)slot * a.fcap * 5 + 5 * i;
        p.f0 = q[0]; p.f1 = q[1]; p.f2 = q[2]; p.f3 = q[3]; p.f4 = q[4];
    }
    if (log && i < ng0 && ng0 <= 64u && ng0 <= a.rcap) {
        const uint32_t *q = a.rs + (size_t)slot * a.rcap * 5 + 5 * i;
        p.g0 = q[0]; p.g1 = q[1]; p.g2 = q[2]; p.g3 = q[3]; p.g4 = q[4];
    }
}

// ---------------------------------------------------------------- register replay
//
// k_resolve_reg<DIR, 1>, the first resolve kernel of a launch, replays every
// touched bin of at most 256 records, 64 fragments and 64 log entries with the
// bin's whole state in registers: fragment f in lane f, log entry g in lane
// g, record r (spawn order) in lane r.  k_resolve_reg<DIR, 2> then takes the
// bins whose map outgrew 64 fragments, with fragment f in lane f % 64 of half
// f / 64 (up to 128).  Rects reach the other lanes by readlane, and ballots
// take the place of the LDS scans, counters and barriers of the general
// replay (do_link, do_run, do_paint, do_log), whose semantics it keeps.  The
// kernel is bound by instruction issue, not memory, so each step loops over
// whichever side is shorter: a run of readers tests its records (one per
// lane) against the fragments in turn, or each record against all fragments;
// a writer tests its overlapping fragments in turn against all log entries,
// or the log entries in turn against all fragments.  LDS only moves values
// between lanes -- the record sort, a paint's new map, a log append or
// compaction, the corner candidates of the coverage test -- and stages the
// candidates.  A bin that does not fit, or would outgrow the fragments or
// the 64 log entries (or meet more than 63 readers in one coverage test),
// goes on -- from the first kernel to the second when its map is what grew,
// else to the first LDS tier -- with nothing written back; the candidates it
// staged are a subset of what the next replay emits for it, and the
// cross-bin dedup absorbs the repeats.

__device__ __forceinline__ uint32_t rl(uint32_t v, uint32_t l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
__device__ __forceinline__ uint32_t popc(uint64_t m) { return (uint32_t)__popcll(m); }
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ uint32_t ctz(uint64_t m) { return (uint32_t)__builtin_ctzll(m); }
// set bits of m below this lane
__device__ __forceinline__ uint32_t popc_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ bool lane_in(uint64_t m) { return (m >> lane_id()) & 1ull; }
__device__ __forceinline__ uint64_t lanes_lt(uint32_t n) { return n >= 64u ? ~0ull : (1ull << n) - 1ull; }

constexpr uint32_t REG_RECS = 256;  // records of a bin the register replay takes (in chunks of 64)

template <int KF>
struct RegLds {
    static constexpr uint32_t CCAP = 128;
    uint32_t X[7][64 * KF];    // lane exchange: records (7 words), a new map (6), log entries (6)
    uint32_t cx[64], cy[64];   // corner candidates of a coverage test
    uint32_t C[3 * CCAP];      // staged candidates (pred, succ access, domain << 24 | succ task)
    uint32_t E[3 * CCAP];      // staged edges known unique, the same three words
    uint32_t ccount;
    uint8_t ord[REG_RECS];     // a bin of more than 64 records: bucket positions in spawn order
    uint4 win[64];             // list entries of the block's next 64 bins
};

// a bin's state: fragment lane + 64 k in half k, one log entry per lane; fo /
// go: the access of the fragment's writer / of the logged reader lies in this
// bin alone (kept for the first 64 fragments)
template <int KF>
struct RegBin {
    uint32_t fx0[KF], fx1[KF], fy0[KF], fy1[KF], flw[KF], fo[KF];  // fragment (lane + 64 k < nf)
    uint32_t gx0, gx1, gy0, gy1, gacc, go;                          // log entry (lane < ng)
};

// word v of fragment f (uniform f)
template <int KF>
__device__ __forceinline__ uint32_t rlf(const uint32_t (&v)[KF], uint32_t f)
{
    if (KF == 1 || f < 64) return rl(v[0], f & 63u);
    return rl(v[KF - 1], f - 64);
}

// The staged edges known unique into the edge stripes: k_dedup gives them
// positions without the table.
template <class S>
__device__ __forceinline__ void flush_direct(const KArgs &a, S &s, uint32_t n)
{
    __syncthreads();  // the staged entries are in LDS
    const uint32_t lane = lane_id();
    const uint32_t stripe = CSTRIPES_C + blockIdx.x % (CSTRIPES - CSTRIPES_C);
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(&a.cstr[16 * stripe], n);
    base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
    if ((uint64_t)base + n > a.cand_cap) {
        if (lane == 0) atomicOr(&a.ctr->err, ERR_CAND);
    } else {
        uint32_t *dst = a.cand + 3ull * ((size_t)stripe * a.cand_cap + base);
        for (uint32_t i = lane; i < 3 * n; i += 64) dst[i] = s.E[i];
    }
    __syncthreads();  // E is refilled after this
}

// Stages (p, acc, task) of every lane of m: the lanes of md (a subset) as edges
// known unique -- the predecessor's access lies in this bin alone and the
// successor's task has one record here, so no other bin and no other record
// can find this (pred, task) key -- the others as candidates for the dedup
// table.  cc / ec: candidates / edges staged.  Staged edges leave LDS only
// between bins (a bin that fails drops its own, see k_resolve_reg); when the
// buffer is full they go on as candidates, which the table dedups just as well.
template <bool DIR, class S>
__device__ __forceinline__ void reg_append(const KArgs &a, S &s, uint32_t &cc, uint32_t &ec, uint64_t m,
                                           uint64_t md, uint32_t p, uint32_t acc, uint32_t task)
{
    if (DIR || ec + popc(md) > S::CCAP) md = 0;  // the direct form inserts everything into the table
    const uint64_t mc = m & ~md;
    if (mc) {
        const uint32_t n = popc(mc);
        if (cc + n > S::CCAP) {
            if (lane_id() == 0) s.ccount = cc;
            flush_cands_to<DIR>(a, s);
            cc = 0;
        }
        if (lane_in(mc)) {
            const uint32_t pos = 3 * (cc + popc_below(mc));
            s.C[pos] = p;
            s.C[pos + 1] = acc;
            s.C[pos + 2] = task;
        }
        cc += n;
    }
    if (md) {
        const uint32_t n = popc(md);
        if (lane_in(md)) {
            const uint32_t pos = 3 * (ec + popc_below(md));
            s.E[pos] = p;
            s.E[pos + 1] = acc;
            s.E[pos + 2] = task;
        }
        ec += n;
    }
}

// the lanes of m holding the first occurrence of each distinct value of v
__device__ __forceinline__ uint64_t distinct_first(uint64_t m, uint32_t v)
{
    uint64_t first = 0;
    while (m) {
        const uint32_t l = ctz(m);
        const uint32_t x = rl(v, l);
        first |= 1ull << l;
        m &= ~__ballot(v == x);
    }
    return first;
}

// the same over the fragment halves: first[k] holds, per distinct value of v
// among the fragments of m, the fragment of lowest index
template <int KF>
__device__ __forceinline__ void distinct_first_f(const uint64_t (&m)[KF], const uint32_t (&v)[KF],
                                                 uint64_t (&first)[KF])
{
    uint64_t mm[KF];
#pragma unroll
    for (int k = 0; k < KF; ++k) {
        mm[k] = m[k];
        first[k] = 0;
    }
#pragma unroll
    for (int k = 0; k < KF; ++k)
        while (mm[k]) {
            const uint32_t l = ctz(mm[k]);
            const uint32_t x = rl(v[k], l);
            first[k] |= 1ull << l;
#pragma unroll
            for (int k2 = k; k2 < KF; ++k2) mm[k2] &= ~__ballot(v[k2] == x);
        }
}

// True when Q minus the log entries of lanes K (clipped to Q) holds a point.
// If the clipped areas sum below Q's area it does.  Otherwise: if it does, it
// holds one (cx, cy) with cx in {Qx0} U {entry x1 < Qx1} and cy in {Qy0} U
// {entry y1 < Qy1} -- from any uncovered point, move left to the largest such
// x at or below it, then down likewise; no entry covers the points passed.  x
// candidate j sits in lane j; the y candidates are taken in turn.  |K| <= 63.
template <class S, class B>
__device__ __forceinline__ bool reg_uncovered(S &s, const B &b, uint64_t K, uint32_t Qx0, uint32_t Qx1, uint32_t Qy0,
                                              uint32_t Qy1)
{
    const uint32_t lane = lane_id();
    const uint32_t kx0 = max(b.gx0, Qx0), kx1 = min(b.gx1, Qx1), ky0 = max(b.gy0, Qy0), ky1 = min(b.gy1, Qy1);
    {
        const uint64_t ar = lane_in(K) ? (uint64_t)(kx1 - kx0) * (ky1 - ky0) : 0ull;
        uint64_t sum = 0;
        for (uint64_t kk = K; kk; kk &= kk - 1) {
            const uint32_t l = ctz(kk);
            sum += ((uint64_t)rl((uint32_t)(ar >> 32), l) << 32) | rl((uint32_t)ar, l);
        }
        if (sum < (uint64_t)(Qx1 - Qx0) * (Qy1 - Qy0)) return true;
    }
    if (lane == 0) { s.cx[0] = Qx0; s.cy[0] = Qy0; }
    if (lane_in(K)) {
        const uint32_t r = 1 + popc_below(K);
        s.cx[r] = kx1 < Qx1 ? kx1 : Qx0;
        s.cy[r] = ky1 < Qy1 ? ky1 : Qy0;
    }
    __syncthreads();
    const uint32_t nc = popc(K) + 1;
    const uint32_t px = s.cx[lane < nc ? lane : 0u], cyv = s.cy[lane < nc ? lane : 0u];
    __syncthreads();  // read before the next test writes them
    bool unc = false;
    for (uint32_t jy = 0; jy < nc && !unc; ++jy) {
        const uint32_t py = rl(cyv, jy);
        bool cov = lane >= nc;
        for (uint64_t kk = K; kk; kk &= kk - 1) {
            const uint32_t l = ctz(kk);
            cov = cov || (rl(kx0, l) <= px && px < rl(kx1, l) && rl(ky0, l) <= py && py < rl(ky1, l));
        }
        unc = __ballot(!cov) != 0;
    }
    return unc;
}

// drops the logged readers that no point still sees (every fragment they
// touch has a newer writer) and compacts the log; gdirty drops to 0 when
// entries move
template <int KF>
__device__ __forceinline__ void reg_prune(RegLds<KF> &s, RegBin<KF> &b, uint32_t nf, uint32_t &ng, uint32_t &gdirty)
{
    const uint32_t lane = lane_id();
    uint64_t alive = 0;
    if (nf <= ng) {  // fragments in turn, log entries in lanes
        bool sees = false;
        for (uint32_t f = 0; f < nf; ++f) {
            const uint32_t lw = rlf<KF>(b.flw, f);
            sees = sees || ((lw == NONE || lw < b.gacc) &&
                            overlaps(rlf<KF>(b.fx0, f), rlf<KF>(b.fx1, f), rlf<KF>(b.fy0, f), rlf<KF>(b.fy1, f),
                                     b.gx0, b.gx1, b.gy0, b.gy1));
        }
        alive = __ballot(lane < ng && sees);
    } else {
        for (uint32_t g = 0; g < ng; ++g) {
            const uint32_t ga = rl(b.gacc, g);
            const uint32_t x0 = rl(b.gx0, g), x1 = rl(b.gx1, g), y0 = rl(b.gy0, g), y1 = rl(b.gy1, g);
            bool sees = false;
#pragma unroll
            for (int k = 0; k < KF; ++k)
                sees = sees || (lane + 64 * k < nf && (b.flw[k] == NONE || b.flw[k] < ga) &&
                                overlaps(b.fx0[k], b.fx1[k], b.fy0[k], b.fy1[k], x0, x1, y0, y1));
            if (__ballot(sees)) alive |= 1ull << g;
        }
    }
    if (alive == lanes_lt(ng)) return;
    if (lane_in(alive)) {
        const uint32_t o = popc_below(alive);
        s.X[0][o] = b.gx0; s.X[1][o] = b.gx1; s.X[2][o] = b.gy0; s.X[3][o] = b.gy1; s.X[4][o] = b.gacc;
        s.X[5][o] = b.go;
    }
    __syncthreads();
    ng = popc(alive);
    if (lane < ng) {
        b.gx0 = s.X[0][lane]; b.gx1 = s.X[1][lane]; b.gy0 = s.X[2][lane]; b.gy1 = s.X[3][lane]; b.gacc = s.X[4][lane];
        b.go = s.X[5][lane];
    }
    __syncthreads();
    gdirty = 0;
}

// PUT of a writer (do_paint): up to four guillotine pieces of every fragment
// under the rect survive with their writer, in fragment order, and the rect
// becomes the last fragment.  False when the map would exceed 64 KF fragments.
template <int KF>
__device__ __forceinline__ bool reg_paint(RegLds<KF> &s, RegBin<KF> &b, uint32_t &nf, uint32_t X0, uint32_t X1,
                                          uint32_t Y0, uint32_t Y1, uint32_t acc, uint32_t one, bool whole)
{
    const uint32_t lane = lane_id();
    if (whole) {  // the rect is the whole bin: one fragment
        if (lane == 0) { b.fx0[0] = X0; b.fx1[0] = X1; b.fy0[0] = Y0; b.fy1[0] = Y1; b.flw[0] = acc; b.fo[0] = one; }
        nf = 1;
        return true;
    }
    uint32_t base = 0;
    bool live[KF], ov[KF], pt[KF], pbm[KF], pl[KF], pr[KF];
    uint32_t o[KF];
#pragma unroll
    for (int k = 0; k < KF; ++k) {
        live[k] = lane + 64 * k < nf;
        ov[k] = live[k] && overlaps(b.fx0[k], b.fx1[k], b.fy0[k], b.fy1[k], X0, X1, Y0, Y1);
        // pieces of f minus the rect: top, bottom, then left and right of the middle band
        pt[k] = ov[k] && b.fx0[k] < X0;
        pbm[k] = ov[k] && X1 < b.fx1[k];
        pl[k] = ov[k] && b.fy0[k] < Y0;
        pr[k] = ov[k] && Y1 < b.fy1[k];
        const uint32_t npc = live[k] ? (ov[k] ? (uint32_t)pt[k] + pbm[k] + pl[k] + pr[k] : 1u) : 0u;
        const uint64_t c1 = __ballot(npc >= 1), c2 = __ballot(npc >= 2), c3 = __ballot(npc >= 3),
                       c4 = __ballot(npc >= 4);
        o[k] = base + popc_below(c1) + popc_below(c2) + popc_below(c3) + popc_below(c4);
        base += popc(c1) + popc(c2) + popc(c3) + popc(c4);
    }
    const uint32_t tot = base;
    if (tot + 1 > 64u * KF) return false;
#pragma unroll
    for (int k = 0; k < KF; ++k) {
        uint32_t q = o[k];
        auto put = [&](uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1) {
            s.X[0][q] = x0; s.X[1][q] = x1; s.X[2][q] = y0; s.X[3][q] = y1; s.X[4][q] = b.flw[k]; s.X[5][q] = b.fo[k];
            ++q;
        };
        if (live[k] && !ov[k]) {
            put(b.fx0[k], b.fx1[k], b.fy0[k], b.fy1[k]);
        } else if (ov[k]) {
            const uint32_t mx0 = max(b.fx0[k], X0), mx1 = min(b.fx1[k], X1);
            if (pt[k]) put(b.fx0[k], X0, b.fy0[k], b.fy1[k]);
            if (pbm[k]) put(X1, b.fx1[k], b.fy0[k], b.fy1[k]);
            if (pl[k]) put(mx0, mx1, b.fy0[k], Y0);
            if (pr[k]) put(mx0, mx1, Y1, b.fy1[k]);
        }
    }
    if (lane == 0) {
        s.X[0][tot] = X0; s.X[1][tot] = X1; s.X[2][tot] = Y0; s.X[3][tot] = Y1; s.X[4][tot] = acc; s.X[5][tot] = one;
    }
    __syncthreads();
    nf = tot + 1;
#pragma unroll
    for (int k = 0; k < KF; ++k) {
        const uint32_t f = lane + 64 * k;
        if (f < nf) {
            b.fx0[k] = s.X[0][f]; b.fx1[k] = s.X[1][f]; b.fy0[k] = s.X[2][f]; b.fy1[k] = s.X[3][f];
            b.flw[k] = s.X[4][f]; b.fo[k] = s.X[5][f];
        }
    }
    __syncthreads();
    return true;
}

// LINK of one record with the reader rule: the last writers of its points
// (single: the record's task has one record in this bin)
template <bool DIR, int KF>
__device__ __forceinline__ void reg_link_reader(const KArgs &a, RegLds<KF> &s, uint32_t &cc, uint32_t &ec,
                                                const RegBin<KF> &b, uint32_t nf, uint32_t X0, uint32_t X1, uint32_t Y0,
                                                uint32_t Y1, uint32_t acc, uint32_t task, bool single)
{
    uint64_t m[KF], first[KF];
#pragma unroll
    for (int k = 0; k < KF; ++k)
        m[k] = __ballot(lane_id() + 64 * k < nf && b.flw[k] != NONE &&
                        overlaps(b.fx0[k], b.fx1[k], b.fy0[k], b.fy1[k], X0, X1, Y0, Y1));
    distinct_first_f<KF>(m, b.flw, first);
#pragma unroll
    for (int k = 0; k < KF; ++k)
        reg_append<DIR>(a, s, cc, ec, first[k], single ? first[k] & __ballot(b.fo[k] != 0) : 0ull, b.flw[k], acc,
                        task);
}

// LINK of one record with the writer rule.  Per overlapping fragment f and
// Q = f n rect: every logged reader newer than lw(f) touching Q is a
// predecessor, and lw(f) is one when some point of Q has no such reader.
// False when a coverage test would meet more than 63 readers.
template <bool DIR, int KF>
__device__ __forceinline__ bool reg_link_writer(const KArgs &a, RegLds<KF> &s, uint32_t &cc, uint32_t &ec,
                                                const RegBin<KF> &b, uint32_t nf, uint32_t ng, uint32_t X0, uint32_t X1,
                                                uint32_t Y0, uint32_t Y1, uint32_t acc, uint32_t task, bool single)
{
    const uint32_t lane = lane_id();
    bool ov[KF];
    uint32_t qx0[KF], qx1[KF], qy0[KF], qy1[KF];
    uint64_t OV[KF], elw[KF], part[KF];  // per half: overlapping, writer a predecessor, readers cover in part
    uint32_t nov = 0;
#pragma unroll
    for (int k = 0; k < KF; ++k) {
        ov[k] = lane + 64 * k < nf && overlaps(b.fx0[k], b.fx1[k], b.fy0[k], b.fy1[k], X0, X1, Y0, Y1);
        qx0[k] = max(b.fx0[k], X0); qx1[k] = min(b.fx1[k], X1); qy0[k] = max(b.fy0[k], Y0); qy1[k] = min(b.fy1[k], Y1);
        OV[k] = __ballot(ov[k]);
        elw[k] = part[k] = 0;
        nov += popc(OV[k]);
    }
    uint64_t PG = 0;  // log lanes that are predecessors
    if (nov <= ng) {  // overlapping fragments in turn, log entries in lanes
#pragma unroll
        for (int k = 0; k < KF; ++k)
            for (uint64_t om = OV[k]; om; om &= om - 1) {
                const uint32_t l = ctz(om);
                const uint32_t Qx0 = rl(qx0[k], l), Qx1 = rl(qx1[k], l), Qy0 = rl(qy0[k], l), Qy1 = rl(qy1[k], l);
                const uint32_t lw = rl(b.flw[k], l);
                const bool t = lane < ng && (lw == NONE || b.gacc > lw) &&
                               overlaps(b.gx0, b.gx1, b.gy0, b.gy1, Qx0, Qx1, Qy0, Qy1);
                const uint64_t K = __ballot(t);
                PG |= K;
                if (lw == NONE) continue;
                if (!K) {
                    elw[k] |= 1ull << l;
                    continue;
                }
                if (!__ballot(t && b.gx0 <= Qx0 && Qx1 <= b.gx1 && b.gy0 <= Qy0 && Qy1 <= b.gy1)) part[k] |= 1ull << l;
            }
    } else {  // log entries in turn, fragments in lanes
        uint32_t nk[KF];
        bool full[KF];
#pragma unroll
        for (int k = 0; k < KF; ++k) {
            nk[k] = 0;
            full[k] = false;
        }
        for (uint32_t g = 0; g < ng; ++g) {
            const uint32_t ga = rl(b.gacc, g);
            const uint32_t gx0 = rl(b.gx0, g), gx1 = rl(b.gx1, g), gy0 = rl(b.gy0, g), gy1 = rl(b.gy1, g);
            bool any = false;
#pragma unroll
            for (int k = 0; k < KF; ++k) {
                const bool t = ov[k] && (b.flw[k] == NONE || ga > b.flw[k]) &&
                               overlaps(gx0, gx1, gy0, gy1, qx0[k], qx1[k], qy0[k], qy1[k]);
                nk[k] += t ? 1u : 0u;
                full[k] = full[k] || (t && gx0 <= qx0[k] && qx1[k] <= gx1 && gy0 <= qy0[k] && qy1[k] <= gy1);
                any = any || t;
            }
            if (__ballot(any)) PG |= 1ull << g;
        }
#pragma unroll
        for (int k = 0; k < KF; ++k) {
            elw[k] = __ballot(ov[k] && b.flw[k] != NONE && nk[k] == 0);
            part[k] = __ballot(ov[k] && b.flw[k] != NONE && nk[k] != 0 && !full[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < KF; ++k)
        for (; part[k]; part[k] &= part[k] - 1) {
            const uint32_t l = ctz(part[k]);
            const uint32_t Qx0 = rl(qx0[k], l), Qx1 = rl(qx1[k], l), Qy0 = rl(qy0[k], l), Qy1 = rl(qy1[k], l);
            const uint32_t lw = rl(b.flw[k], l);
            const uint64_t K =
                __ballot(lane < ng && b.gacc > lw && overlaps(b.gx0, b.gx1, b.gy0, b.gy1, Qx0, Qx1, Qy0, Qy1));
            if (popc(K) > 63u) return false;
            if (reg_uncovered(s, b, K, Qx0, Qx1, Qy0, Qy1)) elw[k] |= 1ull << l;
        }
    // log entries with one access id are the records of one access (a wrapped tile): never
    // single-bin, so the edges known unique hold each reader once
    reg_append<DIR>(a, s, cc, ec, PG, single ? PG & __ballot(b.go != 0) : 0ull, b.gacc, acc, task);
    uint64_t first[KF];
    distinct_first_f<KF>(elw, b.flw, first);
#pragma unroll
    for (int k = 0; k < KF; ++k)
        reg_append<DIR>(a, s, cc, ec, first[k], single ? first[k] & __ballot(b.fo[k] != 0) : 0ull, b.flw[k], acc,
                        task);
    return true;
}

// KF = 1: the touched list; bins whose map outgrows 64 fragments go on to the
// KF = 2 kernel (list 0) when the pool has one (class >= 1), the others to the
// first LDS tier.  KF = 2: list 0; what it cannot hold goes to the first LDS
// tier.
template <bool DIR, int KF>
__global__ void __launch_bounds__(64, (KF == 1 ? 6 : 4)) k_resolve_reg(KArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    RegLds<KF> &s = *reinterpret_cast<RegLds<KF> *>(lds_raw);
    const uint32_t lane = lane_id();
    if (a.ctr->err) return;  // a poisoned pool touches no state
    const uint32_t nlist = KF == 1 ? min(a.ctr->ntouched, a.nbins) : min(a.ctr->ntier[0], a.nbins);
    const uint32_t next = a.first_tier;  // the LDS tier that takes the bins not replayed here
    const bool wide = KF == 1 && a.first_tier >= 2;  // a KF = 2 kernel follows
    unsigned long long visits = 0;
    uint32_t maxf = 0, maxg = 0;
    uint32_t cc = 0, ec = 0;  // candidates staged in s.C, edges known unique in s.E
    // The same software pipeline over the block's bins as k_resolve: list entries 64 at
    // a time, one per lane; the next bin's records, fragments and log loaded before this
    // bin's write-back.
    const uint32_t G = gridDim.x;
    auto entry = [&](uint32_t x) {
        if (x >= nlist) return make_uint4(0, 0, 0, 0);
        if (KF != 1) return a.active[x];
        const uint32_t slot = a.touched[x];
        return make_uint4(slot, 0u, a.bin_cnt[slot], a.bin_nf[slot] | (a.bin_nrs[slot] << 11) | (a.bin_dom[slot] << 24));
    };
    // the entries of bins k in [wbase, wbase + 64), one per lane, kept in LDS (s.win)
    uint32_t wbase = 0;
    s.win[lane] = entry(blockIdx.x + lane * G);
    __syncthreads();
    auto entry_k = [&](uint32_t k) {
        if (k >= wbase + 64) {
            wbase += 64;
            const uint4 e = entry(blockIdx.x + (wbase + lane) * G);
            __syncthreads();
            s.win[lane] = e;
            __syncthreads();
        }
        const uint4 e = s.win[k - wbase];
        return make_uint4(uni(e.x), uni(e.y), uni(e.z), uni(e.w));
    };
    uint4 e_cur = entry_k(0), e_nxt = entry_k(1);
    BinPre pre;
    bin_pre_load(a, e_cur, pre, true);
    for (uint32_t bi = blockIdx.x, k = 0; bi < nlist; bi += G, ++k) {
        const uint4 e = e_cur;
        const BinPre cur = pre;
        e_cur = e_nxt;
        const uint32_t slot = e.x, cnt = e.z, nf0 = e.w & 0x7ffu, ng0 = (e.w >> 11) & 0xfffu, dm = e.w >> 24;
        if (a.dom_err[dm]) {  // a failed domain touches no state
            bin_pre_load(a, e_cur, pre, true);
            e_nxt = entry_k(k + 2);
            if (lane == 0) a.bin_cnt[slot] = 0;
            continue;
        }
        const bool fits = cnt <= REG_RECS && cnt <= a.lmax && ng0 <= 64u;
        bool ok = fits && nf0 <= 64u * KF;
        bool grew = fits && nf0 > 64u * KF;  // the map is what does not fit
        if (ec > RegLds<KF>::CCAP / 2) {  // room for this bin's edges known unique
            flush_direct(a, s, ec);
            ec = 0;
        }
        const uint32_t ec0 = ec;  // a bin that fails drops the edges it staged
        const DomDesc g = a.dom[dm];
        const uint32_t local = slot - g.bin0;
        const uint32_t bxi = local % g.nbx, byi = local / g.nbx;
        const uint32_t BX0 = bxi << g.bxs, BX1 = min((bxi + 1) << g.bxs, g.xmax);
        const uint32_t BY0 = g.ybin0 + (byi << g.bys), BY1 = BY0 + (1u << g.bys);
        const uint32_t tdom = dm << 24;

        RegBin<KF> b;
        uint32_t nf = nf0, ng = ng0, gdirty = ng0, fdirty = 0, painted = 0;
        b.fx0[0] = cur.f0; b.fx1[0] = cur.f1; b.fy0[0] = cur.f2; b.fy1[0] = cur.f3; b.flw[0] = cur.f4;
        b.fo[0] = lane < nf0 ? (cur.mf >> (lane & 31u)) & 1u : 0u;
#pragma unroll
        for (int h = 1; h < KF; ++h) {  // fragments past the first 64 (no single-bin marks)
            const uint32_t f = lane + 64 * h;
            b.fx0[h] = b.fx1[h] = b.fy0[h] = b.fy1[h] = 0;
            b.flw[h] = NONE;
            b.fo[h] = 0;
            if (ok && f < nf0) {
                const uint32_t *q = a.frag + (size_t)slot * a.fcap * 5 + 5 * f;
                b.fx0[h] = q[0]; b.fx1[h] = q[1]; b.fy0[h] = q[2]; b.fy1[h] = q[3]; b.flw[h] = q[4];
            }
        }
        b.gx0 = cur.g0; b.gx1 = cur.g1; b.gy0 = cur.g2; b.gy1 = cur.g3; b.gacc = cur.g4;
        b.go = lane < ng0 ? (cur.mg >> (lane & 31u)) & 1u : 0u;
        if (nf0 == 0) {  // a bin never written: one fragment, no writer
            b.fx0[0] = BX0; b.fx1[0] = BX1; b.fy0[0] = BY0; b.fy1[0] = BY1; b.flw[0] = NONE; b.fo[0] = 0;
            nf = 1;
            fdirty = 1;
        }
        // A bin of more than 64 records is put in spawn order in LDS (s.ord: bucket positions by
        // rank of record index) and replayed in chunks of up to 64 records that end at task
        // group boundaries; the fragments and the log stay in registers across chunks.
        const bool big = cnt > 64u;
        const uint4 *bucket = a.binrec + 2ull * (size_t)slot * a.lmax;
        if (ok && big) {
            // the record indices in LDS (position p at X[p / 64][p % 64]), then each position's
            // rank among them, one column of 64 positions at a time
            const uint32_t ncol = (cnt + 63) / 64;
            for (uint32_t j = 0; j < ncol; ++j) {
                const uint32_t p = lane + 64 * j;
                s.X[j][lane] = p < cnt ? bucket[2 * p + 1].z : NONE;
            }
            __syncthreads();
            bool sorted = true;
            for (uint32_t j = 0; j < ncol; ++j) {
                const uint32_t p = lane + 64 * j;
                sorted = sorted && (p == 0 || p >= cnt || s.X[(p - 1) >> 6][(p - 1) & 63u] < s.X[j][lane]);
            }
            for (uint32_t j = 0; j < ncol; ++j) {
                const uint32_t p = lane + 64 * j, key = s.X[j][lane];
                uint32_t rank = p;
                if (__ballot(!sorted)) {
                    rank = 0;
                    for (uint32_t q = 0; q < cnt; ++q) rank += s.X[q >> 6][q & 63u] < key ? 1u : 0u;
                }
                if (p < cnt) s.ord[rank] = (uint8_t)p;
            }
            __syncthreads();
        }
        for (uint32_t c0 = 0; ok && c0 < cnt;) {
        const uint32_t m = min(64u, cnt - c0);  // records of this chunk before its cut
        // the records, clipped to the bin, in spawn order (rank by record index, distinct
        // within a bin; k_count's buckets are often in order already)
        uint4 q0 = cur.q0, q1 = cur.q1;
        if (big) {
            q0 = make_uint4(0, 0, 0, 0);
            q1 = make_uint4(NONE, 0, NONE, 0);
            if (lane < m) {
                const uint32_t p = s.ord[c0 + lane];
                q0 = bucket[2 * p];
                q1 = bucket[2 * p + 1];
            }
        }
        uint32_t rx0 = max(q0.x, BX0), rx1 = min(q0.y, BX1), ry0 = max(q0.z, BY0), ry1 = min(q0.w, BY1);
        uint32_t racc = q1.x, rtm = q1.y, rone = q1.w;
        if (!big) {
            const uint32_t key = lane < cnt ? cur.q1.z : NONE;
            const uint32_t pkey = __shfl_up(key, 1, 64);
            if (__ballot(lane > 0 && lane < cnt && pkey > key)) {
                uint32_t rank = 0;
                for (uint32_t j = 0; j < cnt; ++j) rank += rl(key, j) < key ? 1u : 0u;
                if (lane < cnt) {
                    s.X[0][rank] = rx0; s.X[1][rank] = rx1; s.X[2][rank] = ry0; s.X[3][rank] = ry1;
                    s.X[4][rank] = racc; s.X[5][rank] = rtm; s.X[6][rank] = rone;
                }
                __syncthreads();
                if (lane < cnt) {
                    rx0 = s.X[0][lane]; rx1 = s.X[1][lane]; ry0 = s.X[2][lane]; ry1 = s.X[3][lane];
                    racc = s.X[4][lane]; rtm = s.X[5][lane]; rone = s.X[6][lane];
                }
                __syncthreads();
            }
        }
        const uint32_t rtask = rtm >> 8, rmode = rtm & 0xffu;
        const uint32_t ptask = __shfl_up(rtask, 1, 64);
        uint32_t nc = m;  // records of this chunk: it ends at a task group boundary
        if (c0 + m < cnt) {
            const uint32_t ntask = bucket[2 * s.ord[c0 + m] + 1].y >> 8;  // the next chunk's first task
            if (rl(rtask, m - 1) == ntask) {
                // cut before the group of the chunk's last record; a group of 64 records or
                // more does not fit
                nc = 63u - (uint32_t)__builtin_clzll(__ballot(lane < m && (lane == 0 || ptask != rtask)));
                if (nc == 0) {
                    ok = false;
                    break;
                }
            }
        }
        const uint64_t Wm = __ballot(lane < nc && (rmode & 2u));                   // writer-rule LINKs
        const uint64_t GS = __ballot(lane < nc && (lane == 0 || ptask != rtask));  // task group starts
        const uint64_t RD = __ballot(lane < nc && rmode == 1u);                    // exact readers (logged)
        // records whose task has no other record in this bin
        const uint64_t SG = GS & ((GS >> 1) | (1ull << (nc - 1)));
        uint32_t i = 0;
        while (ok && i < nc) {
            const uint64_t wr = Wm & (~0ull << i);
            uint32_t end = nc;
            if (wr) {
                const uint32_t w = ctz(wr);
                end = 63u - (uint32_t)__builtin_clzll(GS & lanes_lt(w + 1));  // the writer's group start
            }
            if (end > i) {
                // Whole task groups with no writer: LINK reads only the writer map, PUT only
                // appends readers, so every record links and the exact readers are then logged.
                const uint64_t run = lanes_lt(end) & ~lanes_lt(i);
                if (nf <= end - i) {
                    // fragments in turn, records in lanes.  A writer's fragments are adjacent in
                    // the map (a paint keeps the order of the pieces and appends the new writer,
                    // whose id is the largest), so `last` emits each writer once per record.
                    const bool in = lane_in(run);
                    uint32_t last = NONE;  // the writer this record emitted last
                    for (uint32_t f = 0; f < nf; ++f) {
                        const uint32_t lw = rlf<KF>(b.flw, f);
                        if (lw == NONE) continue;
                        const bool ov = in && lw != last &&
                                        overlaps(rlf<KF>(b.fx0, f), rlf<KF>(b.fx1, f), rlf<KF>(b.fy0, f),
                                                 rlf<KF>(b.fy1, f), rx0, rx1, ry0, ry1);
                        const uint64_t mm = __ballot(ov);
                        reg_append<DIR>(a, s, cc, ec, mm, rlf<KF>(b.fo, f) ? mm & SG : 0ull, lw, racc, rtask | tdom);
                        last = ov ? lw : last;
                    }
                } else {
                    for (uint32_t q = i; q < end; ++q)
                        reg_link_reader<DIR, KF>(a, s, cc, ec, b, nf, rl(rx0, q), rl(rx1, q), rl(ry0, q), rl(ry1, q),
                                                 rl(racc, q), (rl(rtm, q) >> 8) | tdom, (SG >> q) & 1ull);
                }
                const uint64_t rd = RD & run;
                const uint32_t nr = popc(rd);
                if (ng + nr > 64u) {
                    reg_prune<KF>(s, b, nf, ng, gdirty);
                    if (ng + nr > 64u) {
                        ok = false;
                        break;
                    }
                }
                if (nr) {
                    if (lane_in(rd)) {
                        const uint32_t o = ng + popc_below(rd);
                        s.X[0][o] = rx0; s.X[1][o] = rx1; s.X[2][o] = ry0; s.X[3][o] = ry1; s.X[4][o] = racc;
                        s.X[5][o] = rone;
                    }
                    __syncthreads();
                    if (lane >= ng && lane < ng + nr) {
                        b.gx0 = s.X[0][lane]; b.gx1 = s.X[1][lane]; b.gy0 = s.X[2][lane]; b.gy1 = s.X[3][lane];
                        b.gacc = s.X[4][lane]; b.go = s.X[5][lane];
                    }
                    __syncthreads();
                    ng += nr;
                }
                visits += (unsigned long long)nf * (end - i);
                i = end;
                continue;
            }
            // the task group at i writes: LINK each record, then PUT each, in slot order
            const uint64_t nx = GS & ~lanes_lt(i + 1);
            end = nx ? min(ctz(nx), nc) : nc;
            const bool single = end == i + 1;  // the group's only record in this bin
            for (uint32_t q = i; ok && q < end; ++q) {
                const uint32_t X0 = rl(rx0, q), X1 = rl(rx1, q), Y0 = rl(ry0, q), Y1 = rl(ry1, q);
                const uint32_t acc = rl(racc, q), tm = rl(rtm, q), task = (tm >> 8) | tdom;
                if (tm & 2u)
                    ok = reg_link_writer<DIR, KF>(a, s, cc, ec, b, nf, ng, X0, X1, Y0, Y1, acc, task, single);
                else
                    reg_link_reader<DIR, KF>(a, s, cc, ec, b, nf, X0, X1, Y0, Y1, acc, task, single);
            }
            for (uint32_t q = i; ok && q < end; ++q) {
                const uint32_t mode = rl(rtm, q) & 0xffu;
                if (mode & 0x80u) continue;  // a segment seen by this domain: LINK only
                const uint32_t X0 = rl(rx0, q), X1 = rl(rx1, q), Y0 = rl(ry0, q), Y1 = rl(ry1, q), acc = rl(racc, q);
                const uint32_t one = rl(rone, q);
                if (mode & 2u) {
                    ok = reg_paint<KF>(s, b, nf, X0, X1, Y0, Y1, acc, one,
                                       X0 == BX0 && X1 == BX1 && Y0 == BY0 && Y1 == BY1);
                    grew = !ok;
                    fdirty = painted = 1;
                } else if (mode == 1u) {
                    if (ng >= 64u) {
                        reg_prune<KF>(s, b, nf, ng, gdirty);
                        if (ng >= 64u) {
                            ok = false;
                            break;
                        }
                    }
                    if (lane == ng) { b.gx0 = X0; b.gx1 = X1; b.gy0 = Y0; b.gy1 = Y1; b.gacc = acc; b.go = one; }
                    ++ng;
                }
            }
            visits += (unsigned long long)nf * (end - i);
            i = end;
        }
        c0 += nc;
        }
        // a log past half the registers is pruned before it is stored, when a paint may have
        // left readers that no point sees
        if (ok && painted && ng > 32u) reg_prune<KF>(s, b, nf, ng, gdirty);
        bin_pre_load(a, e_cur, pre, true);  // before the write-back stores
        e_nxt = entry_k(k + 2);
        if (!ok) {  // the next replay takes the bin from its HBM state
            ec = ec0;
            if (lane == 0) {
                const uint32_t to = wide && grew ? 0u : next;
                const uint32_t pos = atomicAdd(&a.ctr->ntier[to], 1u);
                a.active[(size_t)to * a.nbins + pos] = e;
            }
            continue;
        }
        uint32_t *fw = a.frag + (size_t)slot * a.fcap * 5;
        uint32_t *gw = a.rs + (size_t)slot * a.rcap * 5;
#pragma unroll
        for (int h = 0; h < KF; ++h) {
            const uint32_t f = lane + 64 * h;
            if (fdirty && f < nf) {
                uint32_t *q = fw + 5 * f;
                q[0] = b.fx0[h]; q[1] = b.fx1[h]; q[2] = b.fy0[h]; q[3] = b.fy1[h]; q[4] = b.flw[h];
            }
        }
        if (lane >= gdirty && lane < ng) {
            uint32_t *q = gw + 5 * lane;
            q[0] = b.gx0; q[1] = b.gx1; q[2] = b.gy0; q[3] = b.gy1; q[4] = b.gacc;
        }
        const uint64_t fm = __ballot(lane < nf && b.fo[0]), gm = __ballot(lane < ng && b.go);
        if (lane == 0) {
            a.bin_nf[slot] = nf;
            a.bin_nrs[slot] = ng;
            a.bin_cnt[slot] = 0;
            a.bin_one[slot] = make_uint4((uint32_t)fm, (uint32_t)(fm >> 32), (uint32_t)gm, (uint32_t)(gm >> 32));
        }
        maxf = max(maxf, nf);
        maxg = max(maxg, ng);
    }
    if (lane == 0) s.ccount = cc;
    flush_cands_to<DIR>(a, s);
    if (ec) flush_direct(a, s, ec);
    if (lane == 0 && visits) {  // statistics, in the first tier's slots
        uint4 *o = &a.blk[next * BLKSLOTS + blockIdx.x % BLKSLOTS];
        atomicMax(&o->x, maxf);
        atomicMax(&o->y, maxg);
        atomicAdd(reinterpret_cast<unsigned long long *>(&o->z), visits);
    }
}

template <int FCAP, int GCAP, int LCAP, bool DIR>
__global__ void __launch_bounds__(64) k_resolve(KArgs a, uint32_t tier)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    BinLds<FCAP, GCAP, LCAP, DIR> &s = *reinterpret_cast<BinLds<FCAP, GCAP, LCAP, DIR> *>(lds_raw);
    const uint32_t lane = lane_id();
    if (a.ctr->err) return;  // a poisoned pool touches no state
    // each tier walks its list: the bins the register replay or the tier below passed on
    const uint32_t nlist = min(a.ctr->ntier[tier], a.nbins);
    const uint4 *list = a.active + (size_t)tier * a.nbins;
    unsigned long long visits = 0;
    uint32_t maxf = 0, maxg = 0;
    if (lane == 0) s.ccount = 0;
    __syncthreads();
    // Software pipeline over this block's bins b, b + G, b + 2G, ...: the
    // first chunk of the next bin's state is loaded before this bin's
    // write-back.  The list entries of the block's bins are read 64 at a time,
    // one per lane, and taken from their lane when a bin comes up: loaded per
    // bin, each entry sat on that bin's path behind the block's earlier loads
    // and stores.
    const uint32_t G = gridDim.x;
    auto entry = [&](uint32_t x) { return x < nlist ? list[x] : make_uint4(0, 0, 0, 0); };
    uint32_t wbase = 0;  // the window holds the entries of bins k in [wbase, wbase + 64)
    uint4 ew = entry(blockIdx.x + lane * G);
    // the entry of the block's k-th bin; k grows by one from call to call
    auto entry_k = [&](uint32_t k) {
        if (k >= wbase + 64) {
            wbase += 64;
            ew = entry(blockIdx.x + (wbase + lane) * G);
        }
        const uint32_t l = k - wbase;
        return make_uint4(__builtin_amdgcn_readlane(ew.x, l), __builtin_amdgcn_readlane(ew.y, l),
                          __builtin_amdgcn_readlane(ew.z, l), __builtin_amdgcn_readlane(ew.w, l));
    };
    uint4 e_cur = entry_k(0), e_nxt = entry_k(1);
    BinPre pre;
    bin_pre_load(a, e_cur, pre, false);
    for (uint32_t b = blockIdx.x, k = 0; b < nlist; b += G, ++k) {
        const uint4 e = e_cur;
        const BinPre cur = pre;
        auto advance = [&]() {
            e_cur = e_nxt;
            bin_pre_load(a, e_cur, pre, false);
            e_nxt = entry_k(k + 2);
        };
        const uint32_t slot = e.x, cnt = e.z, nf0 = e.w & 0x7ffu, ng0 = (e.w >> 11) & 0xfffu;
        const size_t start = (size_t)slot * a.lmax;  // the bin's incidence bucket
        // a failed domain touches no state: its bins are dropped (their counts cleared)
        if (a.dom_err[e.w >> 24]) {
            advance();
            if (lane == 0) a.bin_cnt[slot] = 0;
            continue;
        }
        const uint32_t dm = e.w >> 24;
        const DomDesc g = a.dom[dm];
        const uint32_t local = slot - g.bin0;
        const uint32_t bxi = local % g.nbx, byi = local / g.nbx;
        const uint32_t BX0 = bxi << g.bxs, BX1 = min((bxi + 1) << g.bxs, g.xmax);
        const uint32_t BY0 = g.ybin0 + (byi << g.bys), BY1 = BY0 + (1u << g.bys);
        // a bin goes on to the next tier when it does not fit this one, or when its map or
        // log fill more than half of this tier's room (they grow during the replay)
        if (cnt > (uint32_t)LCAP || nf0 > (uint32_t)FCAP || ng0 > (uint32_t)GCAP ||
            (tier < a.max_tier && (2 * nf0 > (uint32_t)FCAP || 2 * ng0 > (uint32_t)GCAP))) {
            advance();
            escalate<FCAP, GCAP, LCAP, DIR>(a, tier, e, ERR_LIST);
            continue;
        }
        if (lane == 0) s.tdom = dm << 24;

        // record ids, fragments and reader log of the bin, loads issued together
        // (AoS in HBM: 5 words per fragment and per log entry)
        BinState st;
        st.fb = 0; st.err = 0;
        st.slot = slot; st.ng0 = ng0; st.gload = 0; st.gdirty = ng0; st.fdirty = nf0 == 0 ? 1u : 0u;
        uint32_t n2 = 1;
        while (n2 < cnt) n2 <<= 1;
        {
            const uint32_t *fr = a.frag + (size_t)slot * a.fcap * 5;  // AoS, 5 words per fragment
            const uint4 *ir = a.binrec + 2 * start;
            const uint32_t nmax = max(n2, nf0);
            for (uint32_t i = lane; i < nmax; i += 64) {
                uint4 q0 = cur.q0, q1 = cur.q1;  // the first chunk was loaded ahead
                uint32_t f0 = cur.f0, f1 = cur.f1, f2 = cur.f2, f3 = cur.f3, f4 = cur.f4;
                if (i >= 64) {
                    q0 = make_uint4(0, 0, 0, 0); q1 = make_uint4(NONE, 0, NONE, 0);
                    f0 = f1 = f2 = f3 = f4 = 0;
                    if (i < cnt) { q0 = ir[2 * i]; q1 = ir[2 * i + 1]; }
                    if (i < nf0) {
                        const uint32_t *q = fr + 5 * i;
                        f0 = q[0]; f1 = q[1]; f2 = q[2]; f3 = q[3]; f4 = q[4];
                    }
                }
                if (i < n2) {
                    // unsorted staging: key = record index, payload in the record arrays
                    s.L[i] = i < cnt ? (unsigned long long)q1.z : EMPTY64;
                    s.rx0[i] = max(q0.x, BX0); s.rx1[i] = min(q0.y, BX1);
                    s.ry0[i] = max(q0.z, BY0); s.ry1[i] = min(q0.w, BY1);
                    s.racc[i] = q1.x; s.rtm[i] = q1.y;
                }
                if (i < nf0) { s.fx0[0][i] = f0; s.fx1[0][i] = f1; s.fy0[0][i] = f2; s.fy1[0][i] = f3; s.flw[0][i] = f4; }
            }
        }
        if (nf0 == 0) {
            if (lane == 0) {
                s.fx0[0][0] = BX0; s.fx1[0][0] = BX1; s.fy0[0][0] = BY0; s.fy1[0][0] = BY1;
                s.flw[0][0] = NONE;
            }
            st.nf = 1;
        } else {
            st.nf = nf0;
        }
        st.ng = ng0;
        __syncthreads();

        // records in spawn order (record index): short lists move each lane's
        // record straight to its rank; longer ones sort (key, position) pairs and
        // re-read the incidences (L2-resident) in sorted order
        if (n2 <= 64) {
            const unsigned long long key = lane < n2 ? s.L[lane] : EMPTY64;
            uint32_t rank = 0;
            for (uint32_t j = 0; j < n2; ++j) {
                const unsigned long long o = s.L[j];
                rank += (o < key) || (o == key && j < lane);
            }
            uint32_t v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
            if (lane < cnt) {
                v0 = s.rx0[lane]; v1 = s.rx1[lane]; v2 = s.ry0[lane]; v3 = s.ry1[lane]; v4 = s.racc[lane]; v5 = s.rtm[lane];
            }
            __syncthreads();
            if (lane < cnt) {
                s.rx0[rank] = v0; s.rx1[rank] = v1; s.ry0[rank] = v2; s.ry1[rank] = v3; s.racc[rank] = v4;
                s.rtm[rank] = v5;
            }
            __syncthreads();
        } else {
            // keys (record index << 32 | staging position)
            for (uint32_t i = lane; i < n2; i += 64) s.L[i] = i < cnt ? (s.L[i] << 32) | i : EMPTY64;
            __syncthreads();
            for (uint32_t k = 2; k <= n2; k <<= 1)
                for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                    for (uint32_t i = lane; i < n2; i += 64) {
                        const uint32_t ixj = i ^ j;
                        if (ixj > i) {
                            const unsigned long long x = s.L[i], y = s.L[ixj];
                            const bool up = (i & k) == 0;
                            if ((x > y) == up) { s.L[i] = y; s.L[ixj] = x; }
                        }
                    }
                    __syncthreads();
                }
            const uint4 *ir = a.binrec + 2 * start;
            for (uint32_t i = lane; i < cnt; i += 64) {
                const uint32_t pos = (uint32_t)s.L[i];
                const uint4 q0 = ir[2 * pos], q1 = ir[2 * pos + 1];
                s.rx0[i] = max(q0.x, BX0); s.rx1[i] = min(q0.y, BX1);
                s.ry0[i] = max(q0.z, BY0); s.ry1[i] = min(q0.w, BY1);
                s.racc[i] = q1.x; s.rtm[i] = q1.y;
            }
            __syncthreads();
        }
        if (lane == 0) s.dcount = 0;
        for (uint32_t q = lane; q < BinLds<FCAP, GCAP, LCAP, DIR>::DSZ; q += 64) {
            s.D[q] = EMPTY64;
            s.Dv[q] = NONE;
        }
        __syncthreads();

        uint32_t i = 0;
        while (i < cnt && !st.err) {
            // first writer record at or after i
            uint32_t w = cnt;
            for (uint32_t c = i; c < cnt && w == cnt; c += 64) {
                const uint32_t q = c + lane;
                const uint64_t m = __ballot(q < cnt && (s.rtm[q] & 2u));
                if (m) w = c + (uint32_t)__builtin_ctzll(m);
            }
            // start of the writer's task group
            uint32_t k = w;
            if (w < cnt)
                while (k > i && (s.rtm[k - 1] >> 8) == (s.rtm[w] >> 8)) --k;
            if (k > i) {
                do_run(a, s, st, i, k);
                if (s.dcount > BinLds<FCAP, GCAP, LCAP, DIR>::DSZ / 2) drain_dedup(a, s);
                visits += (unsigned long long)st.nf * (k - i);
                i = k;
                continue;
            }
            // the task group at i writes: LINK each record, then PUT each, in slot order
            const uint32_t task = s.rtm[i] >> 8;
            uint32_t j = i + 1;
            while (j < cnt && (s.rtm[j] >> 8) == task) ++j;
            // a chain of one-record writer groups on the same rect: each links to
            // the previous one (it wrote the whole rect and nobody read since), so
            // only the first needs the full LINK and the last one's PUT stands
            if (j == i + 1 && !(s.rtm[i] & 0x80u)) {
                uint32_t m = 1;
                for (uint32_t c = i + 1; c < cnt; c += 64) {
                    const uint32_t q = c + lane;
                    bool ok = false;
                    if (q < cnt) {
                        const uint32_t tq = s.rtm[q] >> 8;
                        ok = (s.rtm[q] & 2u) && !(s.rtm[q] & 0x80u) && s.rx0[q] == s.rx0[i] && s.rx1[q] == s.rx1[i] && s.ry0[q] == s.ry0[i] &&
                             s.ry1[q] == s.ry1[i] && tq != (s.rtm[q - 1] >> 8) && (q + 1 == cnt || (s.rtm[q + 1] >> 8) != tq);
                    }
                    const uint64_t bad = __ballot(!ok);
                    if (bad) {
                        m = c + (uint32_t)__builtin_ctzll(bad) - i;
                        break;
                    }
                    m = c + 64 - i;
                }
                if (m >= 2) {
                    do_link(a, s, st, s.rx0[i], s.rx1[i], s.ry0[i], s.ry1[i], s.rtm[i] & 0xffu, s.racc[i], task);
                    for (uint32_t q = i + 1 + lane; q < i + m; q += 64) emit(a, s, s.racc[q - 1], s.racc[q], s.rtm[q] >> 8);
                    __syncthreads();
                    if (s.dcount > BinLds<FCAP, GCAP, LCAP, DIR>::DSZ / 2) drain_dedup(a, s);
                    const uint32_t last = i + m - 1;
                    do_paint(s, st, s.rx0[last], s.rx1[last], s.ry0[last], s.ry1[last], s.racc[last]);
                    visits += (unsigned long long)st.nf * (m + 1);
                    i += m;
                    continue;
                }
            }
            for (uint32_t q = i; q < j && !st.err; ++q) {
                do_link(a, s, st, s.rx0[q], s.rx1[q], s.ry0[q], s.ry1[q], s.rtm[q] & 0xffu, s.racc[q], task);
                visits += st.nf;
            }
            __syncthreads();
            if (s.dcount > BinLds<FCAP, GCAP, LCAP, DIR>::DSZ / 2) drain_dedup(a, s);
            for (uint32_t q = i; q < j && !st.err; ++q) {
                const uint32_t mode = s.rtm[q] & 0xffu;
                visits += st.nf;
                if (mode & 0x80u) continue;  // a segment seen by this domain: LINK only
                if (mode & 2u)
                    do_paint(s, st, s.rx0[q], s.rx1[q], s.ry0[q], s.ry1[q], s.racc[q]);
                else if (mode == 1u)
                    do_log(a, s, st, s.rx0[q], s.rx1[q], s.ry0[q], s.ry1[q], s.racc[q]);
            }
            i = j;
        }
        if (!st.err) {
            drain_dedup(a, s);
            if (st.gload) prune_log(a, s, st);  // a log never loaded only grew by appends
        }
        advance();  // before the write-back stores
        if (st.err) {
            escalate<FCAP, GCAP, LCAP, DIR>(a, tier, e, st.err);
            __syncthreads();
            continue;
        }

        // write back
        {
            const uint32_t fb = st.fb;
            uint32_t *fw = a.frag + (size_t)slot * a.fcap * 5;
            uint32_t *gw = a.rs + (size_t)slot * a.rcap * 5;
            for (uint32_t f = lane; f < (st.fdirty ? st.nf : 0u); f += 64) {
                uint32_t *q = fw + 5 * f;
                q[0] = s.fx0[fb][f]; q[1] = s.fx1[fb][f]; q[2] = s.fy0[fb][f]; q[3] = s.fy1[fb][f]; q[4] = s.flw[fb][f];
            }
            for (uint32_t g = st.gdirty + lane; g < st.ng; g += 64) {
                uint32_t *q = gw + 5 * g;
                q[0] = s.gx0[g]; q[1] = s.gx1[g]; q[2] = s.gy0[g]; q[3] = s.gy1[g]; q[4] = s.gacc[g];
            }
            if (lane == 0) {
                a.bin_nf[slot] = st.nf;
                a.bin_nrs[slot] = st.ng;
                a.bin_cnt[slot] = 0;
                a.bin_one[slot] = make_uint4(0, 0, 0, 0);  // this replay keeps no single-bin marks
            }
            maxf = max(maxf, st.nf);
            maxg = max(maxg, st.ng);
        }
        __syncthreads();
    }
    flush_cands(a, s);
    if (lane == 0 && visits) {  // statistics: per-block slots, shared past BLKSLOTS blocks
        uint4 *o = &a.blk[tier * BLKSLOTS + blockIdx.x % BLKSLOTS];
        atomicMax(&o->x, maxf);
        atomicMax(&o->y, maxg);
        atomicAdd(reinterpret_cast<unsigned long long *>(&o->z), visits);  // (z, w) = 64-bit count
    }
}

// ---------------------------------------------------------------- dedup

// k_dedup's LDS: where each candidate stripe starts in the launch's candidate
// numbering, and the block's new-edge counts per wave
constexpr uint32_t DEDUP_PER_THREAD = 4;  // candidates per thread and pass
constexpr uint32_t DEDUP_COMPACT_MIN = 1u << 16;  // candidates from which a launch hashes into 2n slots
constexpr uint32_t DEDUP_BLOCKS = 1024;   // k_dedup's and k_final's grids (staged form)
struct DedupLds {
    uint32_t pre[CSTRIPES + 1];
    uint32_t wave[4];
};

// Inserts the staged candidates into the dedup table.  Block b lists the table
// slots of the new keys it made in its own part of eslot (entries
// [b * R, b * R + dcnt[b]), R = eslot size / DEDUP_BLOCKS) and k_final gives
// them edge positions by a prefix over dcnt: no counter is shared (one edge
// counter update per block and pass had every block of a large launch queue
// on that one address).
__global__ void __launch_bounds__(256) k_dedup(KArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    DedupLds &s = *reinterpret_cast<DedupLds *>(lds_raw);
    if (a.ctr->err & ERR_CAND) {  // candidates are incomplete (k_final sees the error)
        if (threadIdx.x == 0) a.dcnt[blockIdx.x] = 0;
        return;
    }
    // candidates: stripe q holds cstr[q] of them from index q * cand_cap
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t q = 0; q < CSTRIPES; ++q) {
            s.pre[q] = t;
            t += min(a.cstr[16 * q], a.cand_cap);
        }
        s.pre[CSTRIPES] = t;
    }
    __syncthreads();
    const uint32_t n = s.pre[CSTRIPES];
    // A launch of many candidates hashes into the first 2^k >= 2n slots: the table is sized
    // for the pool's largest launch (up to GBs), and a launch spread over all of it takes
    // an HBM line per atomic; at 2n slots the table stays in the last-level cache.  Small
    // launches keep the whole table (spread keys keep their few atomics on distinct lines).
    uint32_t mask = a.hmask;
    if (n >= DEDUP_COMPACT_MIN) {
        uint32_t m = 1;
        while (m < 2u * n) m <<= 1;
        mask = min(mask, m - 1);
    }
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint32_t region = (a.hmask + 1) / DEDUP_BLOCKS;  // this block's part of eslot
    uint32_t *mine = a.eslot + (size_t)blockIdx.x * region;
    uint32_t bcount = 0;  // new keys this block made (block-uniform)
    // several candidates per thread only when the grid has more than one each: a
    // thread's inserts run one after the other
    const uint32_t nthreads = gridDim.x * blockDim.x;
    const uint32_t dpt = min(DEDUP_PER_THREAD, max(1u, (uint32_t)(((uint64_t)n + nthreads - 1) / nthreads)));
    const uint32_t per_pass = blockDim.x * dpt;
    // block-uniform trip count: the barriers below are reached by every thread
    for (uint32_t i0 = blockIdx.x * per_pass; i0 < n; i0 += gridDim.x * per_pass) {
        uint32_t h[DEDUP_PER_THREAD], acc[DEDUP_PER_THREAD];
        unsigned long long key[DEDUP_PER_THREAD], old[DEDUP_PER_THREAD];
        bool live[DEDUP_PER_THREAD], direct[DEDUP_PER_THREAD];
        // the first probe of each of the thread's candidates is issued before any
        // result is needed, so their atomics are in flight together
#pragma unroll
        for (uint32_t k = 0; k < DEDUP_PER_THREAD; ++k) {
            const uint32_t i = i0 + k * blockDim.x + threadIdx.x;
            live[k] = k < dpt && i < n;
            direct[k] = false;
            h[k] = 0;
            acc[k] = 0;
            key[k] = EMPTY64;
            old[k] = EMPTY64;
            if (live[k]) {
                uint32_t q = 0;
                while (i >= s.pre[q + 1]) ++q;
                const size_t ci = (size_t)q * a.cand_cap + (i - s.pre[q]);
                if (q >= CSTRIPES_C) {
                    // an edge known unique: its own position, no table entry
                    h[k] = DIRECT_SLOT | (uint32_t)ci;
                    live[k] = false;
                    direct[k] = true;
                } else {
                    const uint32_t *c = a.cand + 3 * ci;
                    key[k] = ((unsigned long long)c[0] << 32) | c[2];
                    acc[k] = c[1];
                    h[k] = (uint32_t)mix64(key[k]) & mask;
                    old[k] = atomicCAS(&a.hkeys[h[k]], EMPTY64, key[k]);
                }
            }
        }
        uint64_t nm[DEDUP_PER_THREAD];
        uint32_t wc = 0;
#pragma unroll
        for (uint32_t k = 0; k < DEDUP_PER_THREAD; ++k) {
            bool isnew = direct[k];
            if (live[k]) {
                // linear probing from the first slot on
                uint32_t hh = h[k], probe = 0;
                unsigned long long o = old[k];
                while (o != EMPTY64 && o != key[k]) {
                    hh = (hh + 1) & mask;
                    if (++probe > mask) {
                        atomicOr(&a.ctr->err, ERR_HASH);
                        break;
                    }
                    o = atomicCAS(&a.hkeys[hh], EMPTY64, key[k]);
                }
                if (o == EMPTY64 || o == key[k]) {
                    atomicMin(&a.hvals[hh], acc[k]);
                    isnew = o == EMPTY64;
                }
                h[k] = hh;
            }
            nm[k] = __ballot(isnew);
            wc += (uint32_t)__popcll(nm[k]);
        }
        if (lane == 0) s.wave[w] = wc;
        __syncthreads();
        uint32_t pos = bcount, total = 0;
        for (uint32_t u = 0; u < (blockDim.x >> 6); ++u) {
            pos += u < w ? s.wave[u] : 0u;
            total += s.wave[u];
        }
#pragma unroll
        for (uint32_t k = 0; k < DEDUP_PER_THREAD; ++k) {
            if ((nm[k] >> lane) & 1ull) {
                const uint32_t e = pos + (uint32_t)__popcll(nm[k] & ((1ull << lane) - 1ull));
                if (e >= region) atomicOr(&a.ctr->err, ERR_EDGES);
                else mine[e] = h[k];
            }
            pos += (uint32_t)__popcll(nm[k]);
        }
        bcount += total;
        __syncthreads();  // s.wave is rewritten by the next pass
    }
    if (threadIdx.x == 0) a.dcnt[blockIdx.x] = bcount;
}

// The staged form's edge count of the launch: the new keys of the k_dedup
// blocks, and where block b's start (an exclusive prefix over dcnt, in LDS)
struct FinalLds {
    uint32_t part[4];
    uint32_t start;      // this block's first edge, relative to the launch's first
    uint32_t total;      // the launch's new edges
};

__global__ void __launch_bounds__(256) k_final(KArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    FinalLds &s = *reinterpret_cast<FinalLds *>(lds_raw);
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const bool staged = a.cand_cap != 0;
    const unsigned long long ebase = a.ctr->edge_base;
    if (staged) {
        // prefix over the k_dedup blocks' counts: 4 per thread, then over the 256 threads
        uint32_t c[4], t = 0;
        for (uint32_t k = 0; k < 4; ++k) {
            c[k] = a.dcnt[4 * threadIdx.x + k];
            t += c[k];
        }
        const uint32_t incl = wave_incl_scan(t);
        if (lane_id() == 63) s.part[threadIdx.x >> 6] = incl;  // the four waves' totals
        __syncthreads();
        if ((blockIdx.x >> 2) == threadIdx.x) {
            uint32_t st = incl - t;
            for (uint32_t u = 0; u < (threadIdx.x >> 6); ++u) st += s.part[u];
            for (uint32_t k = 0; k < (blockIdx.x & 3u); ++k) st += c[k];
            s.start = st;
        }
        if (threadIdx.x == 0) s.total = s.part[0] + s.part[1] + s.part[2] + s.part[3];
        __syncthreads();
        // the launch's edges are [edge_base, edge_base + total): block 0's thread 0 stores the
        // count below, before it writes the launch outcome (no fence: the same thread reads it)
        if (tid == 0) {
            const unsigned long long et = ebase + s.total;
            if (et > a.edge_cap) atomicOr(&a.ctr->err, ERR_EDGES);
            a.ctr->edge_total = et;
        }
    }
    if (blockIdx.x == 0) {
        // the launch's outcome to pinned host memory (seen once the launch's completion is):
        // the per-domain error bits first, four domain slots per word, then the counters
        const unsigned long long seq = a.ctr->nlaunch + 1;
        LaunchSnap *o = a.snap + seq % NSNAP;
        if (threadIdx.x < MAX_DOMS / 4) {
            const uint32_t *d = a.dom_err + 4 * threadIdx.x;
            const uint32_t w = (d[0] & 0xffu) | (d[1] & 0xffu) << 8 | (d[2] & 0xffu) << 16 | (d[3] & 0xffu) << 24;
            __hip_atomic_store(&o->derr[threadIdx.x], w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            a.ctr->ntouched = 0;  // the touched list of the next launch starts empty
            a.ctr->nlaunch = seq;
            __hip_atomic_store(&o->edge_total, a.ctr->edge_total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&o->err, a.ctr->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&o->launch, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    if (tid == 0) {
        unsigned long long nc = 0;
        for (uint32_t q = 0; q < CSTRIPES; ++q) nc += a.cstr[16 * q];
        a.ctr->cand_total += nc;
    }
    if (staged) {
        if (a.ctr->err & (ERR_EDGES | ERR_HASH | ERR_CAND)) return;  // eslot is incomplete
        // this block writes the edges of k_dedup block blockIdx.x: its slots, at its start
        const uint32_t region = (a.hmask + 1) / DEDUP_BLOCKS;
        const uint32_t *mine = a.eslot + (size_t)blockIdx.x * region;
        const uint32_t n = a.dcnt[blockIdx.x];
        const unsigned long long e0 = ebase + s.start;
        for (uint32_t i0 = threadIdx.x; i0 < n; i0 += 4 * blockDim.x) {
            uint32_t h[4], v[4];
            unsigned long long key[4];
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                const uint32_t i = i0 + k * blockDim.x;
                h[k] = i < n ? mine[i] : 0u;
            }
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                const uint32_t i = i0 + k * blockDim.x;
                key[k] = 0ull;
                v[k] = 0u;
                if (i < n && (h[k] & DIRECT_SLOT)) {  // an edge known unique: its staged words
                    const uint32_t *c = a.cand + 3ull * (h[k] & ~DIRECT_SLOT);
                    key[k] = ((unsigned long long)c[0] << 32) | c[2];
                    v[k] = c[1];
                } else if (i < n) {
                    key[k] = a.hkeys[h[k]];
                    v[k] = a.hvals[h[k]];
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                const uint32_t i = i0 + k * blockDim.x;
                if (i < n) {
                    const unsigned long long e = e0 + i;
                    a.edges[2 * e] = (uint32_t)(key[k] >> 32);
                    a.edges[2 * e + 1] = v[k];
                    a.edom[e] = (uint32_t)(key[k] >> 24) & 0xffu;
                    if (!(h[k] & DIRECT_SLOT)) {
                        a.hkeys[h[k]] = EMPTY64;
                        a.hvals[h[k]] = NONE;
                    }
                }
            }
        }
        return;
    }
    // direct form: the resolve kernels gave each new key its position (eslot[e - edge_base])
    if (a.ctr->err & (ERR_EDGES | ERR_HASH)) return;  // eslot is incomplete
    unsigned long long et = a.ctr->edge_total;
    if (et > a.edge_cap) et = a.edge_cap;
    unsigned long long n = et > ebase ? et - ebase : 0ull;
    if (n > (unsigned long long)a.hmask + 1) n = (unsigned long long)a.hmask + 1;
    // four edges per thread and pass, their table reads issued before any store (each
    // edge is a chain: position -> table slot -> key and successor)
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long e0 = tid; e0 < n; e0 += 4 * stride) {
        uint32_t h[4], v[4];
        unsigned long long key[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            const unsigned long long e = e0 + k * stride;
            h[k] = e < n ? a.eslot[e] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            const unsigned long long e = e0 + k * stride;
            key[k] = e < n ? a.hkeys[h[k]] : 0ull;
            v[k] = e < n ? a.hvals[h[k]] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            const unsigned long long e = e0 + k * stride;
            if (e < n) {
                a.edges[2 * (ebase + e)] = (uint32_t)(key[k] >> 32);
                a.edges[2 * (ebase + e) + 1] = v[k];
                a.edom[ebase + e] = (uint32_t)(key[k] >> 24) & 0xffu;
                a.hkeys[h[k]] = EMPTY64;
                a.hvals[h[k]] = NONE;
            }
        }
    }
}

// ---------------------------------------------------------------- maintenance

// forget the state of pool bins [b0, b0 + n); with c, also the pool's counters
__global__ void k_reset(uint32_t *nf, uint32_t *nrs, uint32_t *cnt, uint4 *one, uint32_t b0, uint32_t n,
                        DevCounters *c, uint4 *blk)
{
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t i = tid; i < n; i += gridDim.x * blockDim.x) {
        nf[b0 + i] = 0;
        nrs[b0 + i] = 0;
        cnt[b0 + i] = 0;
        one[b0 + i] = make_uint4(0, 0, 0, 0);
    }
    if (c && blk)
        for (uint32_t i = tid; i < 4 * BLKSLOTS; i += gridDim.x * blockDim.x) blk[i] = make_uint4(0, 0, 0, 0);
    if (tid == 0 && c) {
        c->err = 0;
        c->ntouched = 0;
        c->edge_total = 0;
        c->maxfrag = 0;
        c->maxrs = 0;
        c->cand_total = 0;
        c->frag_visits = 0;
    }
}

__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// plain streaming copy, the reference point for HBM byte counters
__global__ void k_copy(uint4 *dst, const uint4 *src, uint64_t n16)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ---------------------------------------------------------------- launch gate
//
// A launch enqueued before its records exist.  k_gate (one lane) waits until
// the host writes the launch's sequence number into the pinned header of the
// staging buffer -- after the records and their count -- and hands the count
// to the kernels queued behind it, so the host's part of such a launch is
// writing memory.  The wait is bounded: after `limit` wall-clock ticks the gate
// gives up and the launch runs empty.  Header words: [0] record count,
// [1] sequence number, [2] the gate's result (written here).
enum : uint32_t { GATE_WAIT = 0, GATE_OPEN = 1, GATE_EXPIRED = 2 };

__global__ void k_gate(uint32_t *hdr, uint32_t seq, unsigned long long limit, uint32_t *count)
{
    if (threadIdx.x != 0) return;
    const unsigned long long t0 = wall_clock64();
    uint32_t result = GATE_OPEN;
    while (__hip_atomic_load(&hdr[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) != seq) {
        if (wall_clock64() - t0 > limit) {
            result = GATE_EXPIRED;
            break;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    *count = result == GATE_OPEN ? __hip_atomic_load(&hdr[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) : 0u;
    __hip_atomic_store(&hdr[2], result, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void k_consume(DevCounters *c)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) c->edge_total = 0;
}

// retire the accesses with id < h of the bins [b0, b0 + n): a fragment last
// written by one becomes unwritten, its reader-log entries go (the log is
// compacted in order), and the remaining ids drop by h.  One wave per bin.
__global__ void k_retire(uint32_t *nf, uint32_t *nrs, uint4 *one, uint32_t *fr, uint32_t *rs, uint32_t fcap,
                         uint32_t rcap, uint32_t b0, uint32_t n, uint32_t h)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t b = wid; b < n; b += nw) {
        const uint32_t slot = b0 + b;
        const uint32_t nfr = nf[slot];
        uint32_t *f = fr + (size_t)slot * fcap * 5;
        for (uint32_t k = lane; k < nfr; k += 64) {
            const uint32_t lw = f[5 * k + 4];
            f[5 * k + 4] = (lw == NONE || lw < h) ? NONE : lw - h;
        }
        const uint32_t ng = nrs[slot];
        uint32_t *g = rs + (size_t)slot * rcap * 5;
        uint32_t o = 0;
        for (uint32_t c = 0; c < ng; c += 64) {
            const uint32_t k = c + lane;
            uint32_t e[5] = {0, 0, 0, 0, 0};
            if (k < ng)
                for (int w = 0; w < 5; ++w) e[w] = g[5 * k + w];
            const bool keep = k < ng && e[4] >= h;
            const uint64_t m = __ballot(keep);
            // every lane holds its entry before any store: the stores go to
            // positions at or below the loads of this and later chunks
            if (keep) {
                const uint32_t d = o + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                e[4] -= h;
                for (int w = 0; w < 5; ++w) g[5 * d + w] = e[w];
            }
            o += (uint32_t)__popcll(m);
        }
        if (lane == 0) {
            nrs[slot] = o;
            one[slot] = make_uint4(0, 0, 0, 0);  // entries moved: their single-bin marks are dropped
        }
    }
}

// move a domain's grid to another range of the pool: old bin (bx, by) ->
// new bin (bx, by + dy); the ranges are disjoint
__global__ void k_regrid(uint32_t *nf, uint32_t *nrs, uint4 *one, uint32_t *fr, uint32_t *rs, uint32_t fcap,
                         uint32_t rcap, uint32_t ob0, uint32_t onbx, uint32_t onby, uint32_t nb0, uint32_t nnbx,
                         uint32_t dy)
{
    const uint32_t ob = blockIdx.x;
    if (ob >= onbx * onby) return;
    const uint32_t bx = ob % onbx, by = ob / onbx;
    const size_t o = (size_t)ob0 + ob, n = (size_t)nb0 + (size_t)(by + dy) * nnbx + bx;
    const uint32_t f = nf[o], r = nrs[o];
    for (uint32_t i = threadIdx.x; i < 5 * f; i += blockDim.x) fr[n * fcap * 5 + i] = fr[o * fcap * 5 + i];
    for (uint32_t i = threadIdx.x; i < 5 * r; i += blockDim.x) rs[n * rcap * 5 + i] = rs[o * rcap * 5 + i];
    if (threadIdx.x == 0) {
        nf[n] = f;
        nrs[n] = r;
        one[n] = one[o];
    }
}

// ---------------------------------------------------------------- host side

namespace {

struct KClass {
    int fcap, rcap, lcap;
};
// fragments, reader-log entries, records per bin and launch
const KClass KCLASSES[3] = {{64, 128, 256}, {256, 512, 512}, {1024, 2048, 1024}};
// Tiers by (fragments, log entries, records): tier 0 is the register replay's room
// (k_resolve_reg), tiers 1-3 the LDS kernels; tier t + 1 is class t's room
const KClass TIERS[4] = {{64, 64, 256}, {64, 128, 256}, {256, 512, 512}, {1024, 2048, 1024}};

template <int F, int R, int L>
size_t lds_bytes() { return sizeof(BinLds<F, R, L, false>); }  // the same for both forms
static_assert(sizeof(BinLds<1024, 2048, 1024, false>) <= 160u * 1024u, "largest tier exceeds the 160 KiB LDS of a CU");
// the 64-bit LDS arrays sit at 8-byte offsets of the (16-byte aligned) dynamic LDS base
template <class B>
constexpr bool lds_aligned() { return offsetof(B, L) % 8 == 0 && offsetof(B, D) % 8 == 0; }
static_assert(lds_aligned<BinLds<64, 128, 64, false>>() && lds_aligned<BinLds<64, 128, 256, true>>() &&
              lds_aligned<BinLds<256, 512, 512, false>>() && lds_aligned<BinLds<1024, 2048, 1024, true>>(),
              "64-bit LDS arrays must be 8-byte aligned");

constexpr uint32_t NDRING = 4;

int hip_err(hipError_t e)
{
    if (e == hipSuccess) return 0;
    if (e == hipErrorOutOfMemory) return -ENOMEM;
    return -ENODEV;
}

#define HIPCHK(x) do { hipError_t _e = (x); if (_e != hipSuccess) return hip_err(_e); } while (0)

uint32_t ilog2(uint64_t v)
{
    uint32_t r = 0;
    while ((1ull << (r + 1)) <= v) ++r;
    return r;
}

bool is_pow2(uint64_t v) { return v && !(v & (v - 1)); }

} // namespace

struct xkrt_hip_pool {
    xkrt_hip_pool_config_t cfg{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;
    uint32_t kclass = 0, fcap = 0, rcap = 0;
    // bin arrays (cap_bins bins) and their free ranges [start, len), sorted
    uint32_t cap_bins = 0;
    std::vector<std::pair<uint32_t, uint32_t>> free_ranges;
    uint32_t *bin_nf = nullptr, *bin_nrs = nullptr, *frag = nullptr, *rs = nullptr;
    uint4 *bin_one = nullptr;  // per bin: single-bin marks of fragments and log entries
    uint32_t *bin_cnt = nullptr, *bin_dom = nullptr, *touched = nullptr;
    uint4 *active = nullptr;
    // domains
    DomDesc h_dom[MAX_DOMS] = {};
    xkrt_hip_dep_domain *doms[MAX_DOMS] = {};
    uint32_t ndoms = 0;
    DomDesc *d_dom = nullptr;
    bool dom_dirty = true;
    // launch scratch
    uint4 *binrec = nullptr;  // per bin, lmax incidences (32 B each) of the current launch
    uint32_t lmax = 0;
    uint32_t *cand = nullptr;
    uint32_t cand_cap = 0;  // per stripe; 0 in the direct form
    uint32_t *cstr = nullptr;
    uint4 *blk = nullptr;
    unsigned long long *hkeys = nullptr;
    uint32_t *hvals = nullptr;
    uint32_t hmask = 0;
    uint32_t *eslot = nullptr;
    uint32_t *dcnt = nullptr;  // new keys per k_dedup block
    uint32_t *edges = nullptr, *edom = nullptr;  // device addresses (of host memory with host_edges)
    uint32_t *edges_h = nullptr, *edom_h = nullptr;  // host_edges: the same buffers' host addresses
    uint64_t edge_cap = 0;
    DevCounters *ctr = nullptr;
    DevCounters *ctr_host = nullptr;  // pinned
    uint32_t *dom_err = nullptr;       // device: MAX_DOMS per-domain error words
    uint32_t *dom_err_host = nullptr;  // pinned copy (read_counters)
    // staging of the open launch: pinned host records + the record count
    // (one upload per launch: a 32 B header holding the record count, then the
    // records; two pinned host buffers in turn, so staging the next launch does
    // not wait for the upload of the previous one)
    uint4 *rec_dev_hdr = nullptr, *rec_dev = nullptr;
    uint4 *rec_host_hdr[2] = {}, *rec_host = nullptr;  // rec_host: records of buffer cur_host
    uint4 *rec_host_dev[2] = {};   // device addresses of the host buffers (small launches read them in place)
    uint64_t buf_ticket[2] = {};   // the last launch that read a host buffer in place
    uint32_t cur_host = 0;
    uint64_t rec_cap = 0;
    hipEvent_t staged[2] = {};
    bool staged_pending[2] = {};
    // domain table uploads from a pinned ring, so a change never waits for the stream
    DomDesc *dom_ring = nullptr;
    hipEvent_t dom_ev[NDRING] = {};
    bool dom_ev_pending[NDRING] = {};
    uint32_t dom_slot = 0;
    // each launch's outcome, written by its k_final (slot = ticket % NSNAP)
    LaunchSnap *snap = nullptr;
    // launch gate: a small launch enqueued ahead of its records (serving engine)
    bool gate_armed = false;
    uint32_t gate_hb = 0, gate_seq = 0;
    // per staging buffer: the gate result of its previous launch, kept when the buffer is armed
    // again (a second launch in flight may re-arm it before the engine retires that launch)
    uint64_t gate_prev_ticket[2] = {0, 0};
    uint32_t gate_prev_result[2] = {0, 0};
    uint64_t gate_ticket = 0;
    uint32_t *gate_count = nullptr;      // device word: the gated launch's record count
    unsigned long long gate_limit = 0;   // wall-clock ticks a gate waits (1 s)
    struct xkrt_hip_server *server = nullptr;
    uint32_t open_nrec = 0, open_tasks = 0;
    uint64_t open_acc = 0, open_batches = 0;
    // completion tickets: one event per launch, oldest first
    std::deque<std::pair<uint64_t, hipEvent_t>> tickets;
    std::vector<hipEvent_t> ticket_pool;
    uint64_t done = 0;
    // captured launches, one per launch-size class (the launch settings follow the record
    // count: tier-0 blocks per slot below / from 64K records, the tier-1 kernel below 128K)
    static constexpr int NGRAPH = 3;
    hipGraph_t graph[NGRAPH] = {};
    hipGraphExec_t gexec[NGRAPH] = {};
    bool graph_valid[NGRAPH] = {};
    // accounting
    bool poisoned = false;
    uint64_t n_launch = 0;
    uint64_t bytes_scratch = 0;
    // optional HIP-event timing of every launch
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    // per timed launch, in launch order: (kernel slot, start, stop) of every kernel;
    // a graph launch has one span with slot -1
    std::vector<std::vector<std::tuple<int, hipEvent_t, hipEvent_t>>> ev_pending;
    double kernel_ms[6] = {0, 0, 0, 0, 0, 0};
    uint64_t timed_launches = 0;
    std::vector<float> launch_ms;        // per timed launch: first kernel start -> last kernel end
};

struct xkrt_hip_dep_domain {
    xkrt_hip_pool *pool = nullptr;
    bool private_pool = false;
    uint32_t idx = 0;
    uint64_t ld = 0, s = 0, LDs = 0;
    uint32_t bxs = 0, bys = 0;
    uint32_t nbx = 0, nby = 0, ybin0 = 0, bin0 = 0;
    bool fixed_frame = false;  // pool domains: the column range is given at creation
    bool framed = false;
    uint64_t ybase = 0;
    uint64_t next_id = 0;
    uint64_t id_base = 0;  // ids retired so far: reported ids + id_base = ids over the domain's life
    uint64_t max_xlen = 0;
    bool col0 = false;  // some rect started at absolute column 0 (addresses below ld * sizeof_type)
    // a segment domain (interval/dependency-tree.hpp:187-363): segments only, each cut into
    // the rects of a frame of ld-byte columns and PUT exactly
    bool segments = false;
    bool any_id = false;   // batches staged with caller access ids: the last one given
    uint32_t last_id = 0;
    uint64_t n_acc = 0, n_rect = 0;
    uint32_t err = 0;   // device error bits of this domain seen by the host (sticky until reset)
};

namespace {

uint64_t bin_bytes(const xkrt_hip_pool *p)
{
    return 4ull * 5 + 64ull + 16ull + 20ull * p->fcap + 20ull * p->rcap + 32ull * p->lmax;
}

hipEvent_t take_event(xkrt_hip_pool *p)
{
    hipEvent_t e = nullptr;
    if (!p->ev_pool.empty()) {
        e = p->ev_pool.back();
        p->ev_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        e = nullptr;
    }
    return e;
}

// accumulate the elapsed times of completed launches (stream synchronised):
// per kernel, the span of its own start / stop events; per launch, the
// first kernel's start to the last kernel's end
void harvest_events(xkrt_hip_pool *p)
{
    for (auto &spans : p->ev_pending) {
        if (spans.empty()) continue;
        for (auto &[slot, a, b] : spans) {
            float ms = 0.f;
            if (slot >= 0 && hipEventElapsedTime(&ms, a, b) == hipSuccess) p->kernel_ms[slot] += ms;
        }
        float tot = 0.f;
        (void)hipEventElapsedTime(&tot, std::get<1>(spans.front()), std::get<2>(spans.back()));
        p->launch_ms.push_back(tot);
        p->timed_launches++;
        for (auto &[slot, a, b] : spans) {
            p->ev_pool.push_back(a);
            p->ev_pool.push_back(b);
        }
    }
    p->ev_pending.clear();
}

void drop_graph(xkrt_hip_pool *p)
{
    for (int k = 0; k < xkrt_hip_pool::NGRAPH; ++k) {
        if (p->gexec[k]) (void)hipGraphExecDestroy(p->gexec[k]);
        if (p->graph[k]) (void)hipGraphDestroy(p->graph[k]);
        p->gexec[k] = nullptr;
        p->graph[k] = nullptr;
        p->graph_valid[k] = false;
    }
}

// launch-size class of a graphed launch of nrec records, and the record count its
// graph is captured for (the largest of the class: k_count's grid strides over fewer)
int graph_class(const xkrt_hip_pool *p, uint64_t nrec, uint64_t *capture_nrec)
{
    const uint64_t lim[2] = {1ull << 16, 1ull << 17};
    int k = nrec < lim[0] ? 0 : (nrec < lim[1] ? 1 : 2);
    *capture_nrec = std::min<uint64_t>(k < 2 ? lim[k] - 1 : p->rec_cap, p->rec_cap);
    return k;
}

// grow the bin arrays to at least `need` bins (stream synchronised, contents kept)
int grow_bins(xkrt_hip_pool *p, uint32_t need)
{
    uint64_t nc = std::max<uint64_t>(need, (uint64_t)p->cap_bins + p->cap_bins / 2);
    nc = std::max<uint64_t>(nc, 64);
    if (nc > 0xFFFFFFull) nc = std::max<uint64_t>(need, 0xFFFFFFull);
    if (nc > 0xFFFFFFull) return -ERANGE;
    if (p->cfg.budget_bytes && nc * bin_bytes(p) + p->bytes_scratch > p->cfg.budget_bytes) {
        nc = need;  // the smallest growth that admits the request
        if (nc * bin_bytes(p) + p->bytes_scratch > p->cfg.budget_bytes) return -ENOMEM;
    }
    HIPCHK(hipStreamSynchronize(p->stream));
    uint32_t *nf, *nrs, *fr, *rs, *cnt, *dom, *tch;
    uint4 *act, *brec, *one;
    HIPCHK(hipMalloc(&nf, nc * 4));
    HIPCHK(hipMalloc(&one, nc * 16));
    HIPCHK(hipMalloc(&nrs, nc * 4));
    HIPCHK(hipMalloc(&fr, nc * p->fcap * 20ull));
    HIPCHK(hipMalloc(&rs, nc * p->rcap * 20ull));
    HIPCHK(hipMalloc(&cnt, nc * 4));
    HIPCHK(hipMalloc(&dom, nc * 4));
    HIPCHK(hipMalloc(&brec, nc * p->lmax * 32ull));  // launch scratch: nothing to copy
    HIPCHK(hipMalloc(&tch, nc * 4));
    HIPCHK(hipMalloc(&act, nc * 16 * 4));
    HIPCHK(hipMemsetAsync(nf, 0, nc * 4, p->stream));
    HIPCHK(hipMemsetAsync(nrs, 0, nc * 4, p->stream));
    HIPCHK(hipMemsetAsync(one, 0, nc * 16, p->stream));
    HIPCHK(hipMemsetAsync(cnt, 0, nc * 4, p->stream));
    HIPCHK(hipMemsetAsync(dom, 0, nc * 4, p->stream));
    const uint64_t oc = p->cap_bins;
    if (oc) {
        HIPCHK(hipMemcpyAsync(nf, p->bin_nf, oc * 4, hipMemcpyDeviceToDevice, p->stream));
        HIPCHK(hipMemcpyAsync(nrs, p->bin_nrs, oc * 4, hipMemcpyDeviceToDevice, p->stream));
        HIPCHK(hipMemcpyAsync(one, p->bin_one, oc * 16, hipMemcpyDeviceToDevice, p->stream));
        HIPCHK(hipMemcpyAsync(fr, p->frag, oc * p->fcap * 20ull, hipMemcpyDeviceToDevice, p->stream));
        HIPCHK(hipMemcpyAsync(rs, p->rs, oc * p->rcap * 20ull, hipMemcpyDeviceToDevice, p->stream));
        HIPCHK(hipMemcpyAsync(cnt, p->bin_cnt, oc * 4, hipMemcpyDeviceToDevice, p->stream));
        HIPCHK(hipMemcpyAsync(dom, p->bin_dom, oc * 4, hipMemcpyDeviceToDevice, p->stream));
        HIPCHK(hipStreamSynchronize(p->stream));
        (void)hipFree(p->bin_nf); (void)hipFree(p->bin_nrs); (void)hipFree(p->bin_one); (void)hipFree(p->frag);
        (void)hipFree(p->rs);
        (void)hipFree(p->bin_cnt); (void)hipFree(p->bin_dom); (void)hipFree(p->active);
        (void)hipFree(p->touched); (void)hipFree(p->binrec);
    }
    p->bin_nf = nf; p->bin_nrs = nrs; p->bin_one = one; p->frag = fr; p->rs = rs;
    p->bin_cnt = cnt; p->bin_dom = dom; p->active = act; p->touched = tch; p->binrec = brec;
    // the new tail is free
    if (!p->free_ranges.empty() && p->free_ranges.back().first + p->free_ranges.back().second == oc)
        p->free_ranges.back().second += (uint32_t)(nc - oc);
    else
        p->free_ranges.push_back({(uint32_t)oc, (uint32_t)(nc - oc)});
    p->cap_bins = (uint32_t)nc;
    drop_graph(p);  // the arrays moved
    return 0;
}

// first-fit range of n bins, cleared and tagged with domain `dom`
int alloc_bins(xkrt_hip_pool *p, uint32_t n, uint32_t dom, uint32_t *b0)
{
    for (int attempt = 0; attempt < 2; ++attempt) {
        for (size_t i = 0; i < p->free_ranges.size(); ++i) {
            auto &r = p->free_ranges[i];
            if (r.second < n) continue;
            *b0 = r.first;
            r.first += n;
            r.second -= n;
            if (r.second == 0) p->free_ranges.erase(p->free_ranges.begin() + (long)i);
            hipLaunchKernelGGL(k_reset, dim3(std::min<uint32_t>((n + 255) / 256, 4096u)), dim3(256), 0, p->stream,
                               p->bin_nf, p->bin_nrs, p->bin_cnt, p->bin_one, *b0, n, (DevCounters *)nullptr,
                               (uint4 *)nullptr);
            hipLaunchKernelGGL(k_fill_u32, dim3(std::min<uint32_t>((n + 255) / 256, 4096u)), dim3(256), 0,
                               p->stream, p->bin_dom + *b0, n, dom);
            return hip_err(hipGetLastError());
        }
        // grow so that the trailing free range (if any) can hold n bins
        uint32_t tail = 0;
        if (!p->free_ranges.empty() && p->free_ranges.back().first + p->free_ranges.back().second == p->cap_bins)
            tail = p->free_ranges.back().second;
        const int rc = grow_bins(p, p->cap_bins + (n - tail));
        if (rc) return rc;
    }
    return -ENOMEM;
}

void free_bins(xkrt_hip_pool *p, uint32_t b0, uint32_t n)
{
    if (!n) return;
    auto &v = p->free_ranges;
    auto it = std::lower_bound(v.begin(), v.end(), std::make_pair(b0, 0u));
    it = v.insert(it, {b0, n});
    size_t i = (size_t)(it - v.begin());
    if (i + 1 < v.size() && v[i].first + v[i].second == v[i + 1].first) {
        v[i].second += v[i + 1].second;
        v.erase(v.begin() + (long)i + 1);
    }
    if (i > 0 && v[i - 1].first + v[i - 1].second == v[i].first) {
        v[i - 1].second += v[i].second;
        v.erase(v.begin() + (long)i);
    }
}

void publish(xkrt_hip_pool *p, xkrt_hip_dep_domain *d)
{
    DomDesc &g = p->h_dom[d->idx];
    g.bin0 = d->bin0; g.nbx = d->nbx; g.nby = d->nby; g.bxs = d->bxs; g.bys = d->bys;
    g.ybin0 = d->ybin0; g.xmax = (uint32_t)d->LDs; g.pad = 0;
    p->dom_dirty = true;
}

// make the domain's grid cover [ylo, yhi) (frame coordinates); grows by half
// the current span in the direction of growth so repeated growth stays cheap
int ensure_grid(xkrt_hip_dep_domain *d, uint32_t ylo, uint32_t yhi)
{
    xkrt_hip_pool *p = d->pool;
    const uint64_t BY = 1ull << d->bys;
    const uint64_t lo = (uint64_t)ylo & ~(BY - 1), hi = ((uint64_t)yhi + BY - 1) & ~(BY - 1);
    const uint64_t olo = d->ybin0, ohi = (uint64_t)d->ybin0 + (uint64_t)d->nby * BY;
    if (d->nby && lo >= olo && hi <= ohi) return 0;
    if (d->fixed_frame) return -ERANGE;
    uint64_t nlo = lo, nhi = hi;
    if (d->nby) {
        const uint64_t extra = ((ohi - olo) / 2) & ~(BY - 1);
        nlo = std::min(lo, olo);
        nhi = std::max(hi, ohi);
        if (lo < olo) nlo = nlo > extra ? nlo - extra : 0;
        if (hi > ohi) nhi = std::min<uint64_t>(nhi + extra, (1ull << 32) - BY);
    }
    const uint64_t rows = (nhi - nlo) / BY;
    const uint64_t nb_new = (uint64_t)d->nbx * rows;
    if (nb_new > 0xFFFFFFull) return -ERANGE;
    uint32_t nb0 = 0;
    int rc = alloc_bins(p, (uint32_t)nb_new, d->idx, &nb0);
    if (rc) return rc;
    if (d->nby) {
        const uint32_t dy = (uint32_t)((olo - nlo) / BY);
        hipLaunchKernelGGL(k_regrid, dim3(d->nbx * d->nby), dim3(256), 0, p->stream, p->bin_nf, p->bin_nrs,
                           p->bin_one, p->frag, p->rs, p->fcap, p->rcap, d->bin0, d->nbx, d->nby, nb0, d->nbx, dy);
        HIPCHK(hipGetLastError());
        free_bins(p, d->bin0, d->nbx * d->nby);
    }
    d->bin0 = nb0;
    d->ybin0 = (uint32_t)nlo;
    d->nby = (uint32_t)rows;
    publish(p, d);
    return 0;
}

KArgs make_args(xkrt_hip_pool *p, const uint4 *rec, uint32_t nrec_arg)
{
    KArgs a;
    a.rec = rec;
    a.nrec_arg = nrec_arg;
    a.nrec_ptr = (const uint32_t *)p->rec_dev_hdr;
    a.dom = p->d_dom;
    a.bin_dom = p->bin_dom;
    a.bin_nf = p->bin_nf; a.bin_nrs = p->bin_nrs; a.bin_one = p->bin_one; a.frag = p->frag; a.rs = p->rs;
    a.fcap = p->fcap; a.rcap = p->rcap;
    a.bin_cnt = p->bin_cnt; a.active = p->active; a.binrec = p->binrec;
    a.touched = p->touched;
    a.lmax = p->lmax;
    a.nbins = p->cap_bins;
    a.max_tier = p->kclass + 1;  // tier t + 1 is class t
    a.first_tier = p->kclass >= 1 ? 2u : 1u;
    a.krpw = KRPW_SMALL;
    a.cand = p->cand; a.cand_cap = p->cand_cap; a.cstr = p->cstr; a.blk = p->blk;
    a.hkeys = p->hkeys; a.hvals = p->hvals; a.hmask = p->hmask;
    a.eslot = p->eslot; a.dcnt = p->dcnt; a.edges = p->edges; a.edom = p->edom; a.edge_cap = p->edge_cap;
    a.ctr = p->ctr;
    a.dom_err = p->dom_err;
    a.snap = p->snap;
    return a;
}

// launches a kernel on stream st; when `spans` is given, with start / stop
// events around that kernel alone (hipExtLaunchKernelGGL), filed under kernel
// slot `slot`
template <typename K, typename... Args>
void launch_on(xkrt_hip_pool *p, hipStream_t st, std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> *spans,
               int slot, K kernel, dim3 grid, dim3 block, uint32_t lds, Args... args)
{
    if (spans) {
        hipEvent_t a = take_event(p), b = take_event(p);
        if (a && b) {
            hipExtLaunchKernelGGL(kernel, grid, block, lds, st, a, b, 0, args...);
            spans->push_back({slot, a, b});
            return;
        }
        if (a) p->ev_pool.push_back(a);
        if (b) p->ev_pool.push_back(b);
    }
    hipLaunchKernelGGL(kernel, grid, block, lds, st, args...);
}

template <typename K, typename... Args>
void launch(xkrt_hip_pool *p, std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> *spans, int slot, K kernel,
            dim3 grid, dim3 block, uint32_t lds, Args... args)
{
    launch_on(p, p->stream, spans, slot, kernel, grid, block, lds, args...);
}

template <int F, int R, int L>
void launch_resolve(xkrt_hip_pool *p, const KArgs &a, uint32_t tier,
                    std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> *spans, hipStream_t st = nullptr,
                    uint32_t over = 1)
{
    if (!st) st = p->stream;
    // enough blocks to fill the chip at this tier's LDS footprint (times `over`); each walks
    // its share of the list
    const size_t lds = lds_bytes<F, R, L>();
    const uint32_t per_cu = std::max<uint32_t>(1u, (uint32_t)(160u * 1024u / lds));
    const uint32_t blocks = std::min<uint32_t>(p->cap_bins, 256u * std::min<uint32_t>(per_cu, 16u) * over);
    if (p->cfg.direct_dedup)
        launch_on(p, st, spans, 3, k_resolve<F, R, L, true>, dim3(blocks), dim3(64), (uint32_t)lds, a, tier);
    else
        launch_on(p, st, spans, 3, k_resolve<F, R, L, false>, dim3(blocks), dim3(64), (uint32_t)lds, a, tier);
}

// k_resolve_reg blocks per CU: 64 fragments per bin, 6 waves per SIMD by its launch bound (its
// LDS, 6.6 KB, allows 24 blocks); 128 fragments, 4 waves per SIMD (8.3 KB)
constexpr uint32_t REG_PER_CU = 24, REG2_PER_CU = 16;

template <int KF>
void launch_reg_kf(xkrt_hip_pool *p, const KArgs &a, std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> *spans,
                   uint32_t blocks)
{
    if (p->cfg.direct_dedup)
        launch(p, spans, 3, k_resolve_reg<true, KF>, dim3(blocks), dim3(64), (uint32_t)sizeof(RegLds<KF>), a);
    else
        launch(p, spans, 3, k_resolve_reg<false, KF>, dim3(blocks), dim3(64), (uint32_t)sizeof(RegLds<KF>), a);
}

// The register replay: 64 fragments per bin on the touched list (its blocks stride over it: a
// small launch gets at most one block per record), then, in classes with more than 64 fragments
// per bin, 128 on the bins whose map outgrew 64.
void launch_reg(xkrt_hip_pool *p, const KArgs &a, std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> *spans,
                uint32_t over, uint64_t nrec)
{
    const uint64_t cap = std::max<uint64_t>(nrec, 256);
    launch_reg_kf<1>(p, a, spans,
                     (uint32_t)std::min<uint64_t>({(uint64_t)p->cap_bins, 256ull * REG_PER_CU * over, cap}));
    if (a.first_tier >= 2)
        launch_reg_kf<2>(p, a, spans, (uint32_t)std::min<uint64_t>({(uint64_t)p->cap_bins, 256ull * REG2_PER_CU, cap}));
}

// the kernels of one launch (timing slots: count, -, -, resolve, dedup, final)
// k_count blocks (4 waves of krpw records each) for n records; it strides over more
uint32_t count_blocks(uint64_t n, uint32_t krpw)
{
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 4 * krpw - 1) / (4 * krpw), 8192));
}

void launch_kernels(xkrt_hip_pool *p, const KArgs &a, uint64_t nrec, bool small,
                    std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> *spans)
{
    KArgs ak = a;
    ak.krpw = nrec >= KRPW_LARGE_FROM ? KRPW_LARGE : KRPW_SMALL;
    launch(p, spans, 0, k_count, dim3(count_blocks(nrec, ak.krpw)), dim3(256), COUNT_LDS, ak);
    // The register replay takes the touched bins.  Several blocks per resident slot: each
    // walks fewer bins, and a waiting block starts wherever one finished, so the bins that take
    // several times the mean do not pile up on some blocks while the others idle.  Launches of
    // many records (tens of records per bin) gain up to 8 blocks per slot; short ones do best
    // with 3 and lose to the extra block starts beyond that; a small launch takes one.
    const uint32_t over = small ? 1u : (nrec >= (1u << 16) ? 8u : 3u);
    launch_reg(p, a, spans, over, nrec);
    // The bins it passes on hold more than 64 records, fragments or log entries: tier 0 (64
    // fragments and records) cannot take them, so they go to the first tier with more room --
    // tier 1 (128 log entries, 256 records) in class 0, else tier 2 -- and up from there.
    if (a.first_tier <= 1) launch_resolve<64, 128, 256>(p, a, 1, spans);
    if (p->kclass >= 1) launch_resolve<256, 512, 512>(p, a, 2, spans);
    if (p->kclass >= 2) launch_resolve<1024, 2048, 1024>(p, a, 3, spans);
    if (!p->cfg.direct_dedup)
        launch(p, spans, 4, k_dedup, dim3(DEDUP_BLOCKS), dim3(256), (uint32_t)sizeof(DedupLds), a);
    launch(p, spans, 5, k_final, dim3(DEDUP_BLOCKS), dim3(256), (uint32_t)sizeof(FinalLds), a);
}

int push_domains(xkrt_hip_pool *p)
{
    if (!p->dom_dirty) return 0;
    // stream ordered, so launches in flight still read the table they were enqueued with;
    // the pinned ring slot stays untouched until its upload is done
    const uint32_t k = p->dom_slot;
    if (p->dom_ev_pending[k]) HIPCHK(hipEventSynchronize(p->dom_ev[k]));
    DomDesc *src = p->dom_ring + (size_t)k * MAX_DOMS;
    memcpy(src, p->h_dom, sizeof(p->h_dom));
    HIPCHK(hipMemcpyAsync(p->d_dom, src, sizeof(p->h_dom), hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipEventRecord(p->dom_ev[k], p->stream));
    p->dom_ev_pending[k] = true;
    p->dom_slot = (k + 1) % NDRING;
    p->dom_dirty = false;
    return 0;
}

// one launch over nrec records at rec (device); the record count is already
// in the counters when `counted`, else it is set by a kernel
int ticket(xkrt_hip_pool *p);

// a small launch: records read in place from the pinned host buffer at hdr + 32 B
// (header = record count), no graph, the register replay on one block per resident slot
int enqueue_small(xkrt_hip_pool *p, const uint4 *hdr, uint32_t nrec)
{
    int rc = push_domains(p);
    if (rc) return rc;
    KArgs a = make_args(p, hdr + 2, NONE);
    a.nrec_ptr = (const uint32_t *)hdr;
    std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> spans;
    launch_kernels(p, a, nrec, true, p->timing ? &spans : nullptr);
    if (!spans.empty()) p->ev_pending.push_back(std::move(spans));
    return ticket(p);
}

int enqueue(xkrt_hip_pool *p, const uint4 *rec, uint32_t nrec, bool counted)
{
    int rc = push_domains(p);
    if (rc) return rc;
    const bool graphed = p->cfg.graph && rec == p->rec_dev && counted;
    if (graphed) {
        uint64_t cap_nrec = 0;
        const int gk = graph_class(p, nrec, &cap_nrec);
        if (!p->graph_valid[gk]) {
            if (p->gexec[gk]) (void)hipGraphExecDestroy(p->gexec[gk]);
            if (p->graph[gk]) (void)hipGraphDestroy(p->graph[gk]);
            p->gexec[gk] = nullptr;
            p->graph[gk] = nullptr;
            const KArgs a = make_args(p, p->rec_dev, NONE);
            HIPCHK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
            launch_kernels(p, a, cap_nrec, false, nullptr);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(p->stream, &g);
            if (e != hipSuccess) return hip_err(e);
            p->graph[gk] = g;
            HIPCHK(hipGraphInstantiate(&p->gexec[gk], p->graph[gk], nullptr, nullptr, 0));
            p->graph_valid[gk] = true;
        }
        hipEvent_t ev[2] = {};
        if (p->timing) {
            ev[0] = take_event(p);
            ev[1] = take_event(p);
        }
        if (ev[0]) (void)hipEventRecord(ev[0], p->stream);
        HIPCHK(hipGraphLaunch(p->gexec[gk], p->stream));
        if (ev[1]) (void)hipEventRecord(ev[1], p->stream);
        if (ev[0] && ev[1]) {
            p->ev_pending.push_back({{-1, ev[0], ev[1]}});
        } else {
            for (hipEvent_t e : ev) if (e) p->ev_pool.push_back(e);
        }
    } else {
        const KArgs a = make_args(p, rec, counted ? NONE : nrec);
        std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> spans;
        launch_kernels(p, a, nrec, false, p->timing ? &spans : nullptr);
        if (!spans.empty()) p->ev_pending.push_back(std::move(spans));
    }
    return ticket(p);
}

// the completion ticket of the launch just enqueued
int ticket(xkrt_hip_pool *p)
{
    HIPCHK(hipGetLastError());
    p->n_launch++;
    // completion ticket of this launch
    hipEvent_t t = nullptr;
    if (!p->ticket_pool.empty()) {
        t = p->ticket_pool.back();
        p->ticket_pool.pop_back();
    } else {
        HIPCHK(hipEventCreateWithFlags(&t, hipEventDisableTiming));
    }
    HIPCHK(hipEventRecord(t, p->stream));
    p->tickets.push_back({p->n_launch, t});
    return 0;
}

// retire the tickets of completed launches (non-blocking unless `until`)
int retire(xkrt_hip_pool *p, uint64_t until)
{
    while (!p->tickets.empty()) {
        auto &f = p->tickets.front();
        if (f.first <= until) {
            HIPCHK(hipEventSynchronize(f.second));
        } else {
            const hipError_t q = hipEventQuery(f.second);
            if (q == hipErrorNotReady) break;
            if (q != hipSuccess) return hip_err(q);
        }
        p->done = f.first;
        p->ticket_pool.push_back(f.second);
        p->tickets.pop_front();
    }
    return 0;
}

int err_to_errno(uint32_t e)
{
    if (e & ERR_RANGE) return -ERANGE;
    if (e & (ERR_LIST | ERR_FRAG | ERR_READS)) return -ENOSPC;
    if (e & (ERR_CAND | ERR_HASH | ERR_EDGES)) return -EOVERFLOW;
    return 0;
}

// the device error bits of each admitted domain into its host record (sticky)
void fold_domain_errors(xkrt_hip_pool *p, const uint32_t *derr)
{
    for (uint32_t i = 0; i < MAX_DOMS; ++i)
        if (p->doms[i] && derr[i]) p->doms[i]->err |= derr[i] & ERR_DOMAIN_BITS;
}

// Synchronises the pool's stream and reads its counters and the per-domain
// error words.  Returns a pool-level error (the pool is poisoned); a domain's
// own failure is kept in its record and reported by the calls on that domain.
int read_counters(xkrt_hip_pool *p)
{
    HIPCHK(hipMemcpyAsync(p->ctr_host, p->ctr, sizeof(DevCounters), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipMemcpyAsync(p->dom_err_host, p->dom_err, 4ull * MAX_DOMS, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    harvest_events(p);
    fold_domain_errors(p, p->dom_err_host);
    if (p->ctr_host->err) {
        p->poisoned = true;
        return err_to_errno(p->ctr_host->err);
    }
    return 0;
}

// the domain's own failure as a negative errno (0: healthy)
int domain_error(const xkrt_hip_dep_domain *d) { return err_to_errno(d->err); }

// clears the device error word of domain slot idx (stream ordered)
int clear_domain_error(xkrt_hip_pool *p, uint32_t idx)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, p->stream, p->dom_err + idx, 1u, 0u);
    return hip_err(hipGetLastError());
}

void destroy_pool(xkrt_hip_pool *p)
{
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    drop_graph(p);
    for (auto &t : p->tickets) (void)hipEventDestroy(t.second);
    for (hipEvent_t e : p->ticket_pool) (void)hipEventDestroy(e);
    for (auto &spans : p->ev_pending)
        for (auto &[slot, a, b] : spans) {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    for (hipEvent_t e : p->ev_pool) (void)hipEventDestroy(e);
    void *dev[] = {p->bin_nf, p->bin_nrs, p->bin_one, p->frag, p->rs, p->bin_cnt, p->bin_dom, p->active, p->touched,
                   p->d_dom, p->dom_err,
                   p->binrec, p->cand, p->cstr, p->blk, p->hkeys, p->hvals, p->eslot, p->dcnt, p->ctr,
                   p->gate_count,
                   p->rec_dev_hdr};
    for (void *q : dev)
        if (q) (void)hipFree(q);
    if (p->edges_h || p->edom_h) {
        if (p->edges_h) (void)hipHostFree(p->edges_h);
        if (p->edom_h) (void)hipHostFree(p->edom_h);
    } else {
        if (p->edges) (void)hipFree(p->edges);
        if (p->edom) (void)hipFree(p->edom);
    }
    if (p->ctr_host) (void)hipHostFree(p->ctr_host);
    if (p->dom_err_host) (void)hipHostFree(p->dom_err_host);
    for (int b = 0; b < 2; ++b) {
        if (p->rec_host_hdr[b]) (void)hipHostFree(p->rec_host_hdr[b]);
        if (p->staged[b]) (void)hipEventDestroy(p->staged[b]);
    }
    if (p->dom_ring) (void)hipHostFree(p->dom_ring);
    for (hipEvent_t e : p->dom_ev)
        if (e) (void)hipEventDestroy(e);
    if (p->snap) (void)hipHostFree(p->snap);
    if (p->own_stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

// dynamic LDS limit of one tier's kernels (both dedup forms).  The kernels
// cast the dynamic LDS base to BinLds, whose 64-bit arrays need it 8-byte
// aligned: that holds only while no static __shared__ variable sits before it,
// so a kernel with static LDS is refused here rather than launched.
template <int F, int R, int L>
hipError_t set_lds()
{
    const void *k[2] = {(const void *)k_resolve<F, R, L, false>, (const void *)k_resolve<F, R, L, true>};
    for (const void *f : k) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes<F, R, L>());
        if (e != hipSuccess) return e;
        hipFuncAttributes at{};
        e = hipFuncGetAttributes(&at, f);
        if (e != hipSuccess) return e;
        if (at.sharedSizeBytes != 0) return hipErrorInvalidKernelFile;
    }
    return hipSuccess;
}

int create_pool(const xkrt_hip_pool_config_t *cfg, void *stream, xkrt_hip_pool **out)
{
    xkrt_hip_pool *p = new xkrt_hip_pool();
    if (cfg) p->cfg = *cfg;
    if (p->cfg.kernel > 2) { delete p; return -EINVAL; }
    p->kclass = p->cfg.kernel;
    p->fcap = KCLASSES[p->kclass].fcap;
    p->rcap = KCLASSES[p->kclass].rcap;
    if (hipGetDevice(&p->device) != hipSuccess) { delete p; return -ENODEV; }
    if (stream) {
        p->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) { delete p; return -ENODEV; }
        p->own_stream = true;
    }
    const uint64_t max_rects = p->cfg.max_rects ? p->cfg.max_rects : (1ull << 20);
    p->cfg.max_rects = max_rects;
    p->rec_cap = max_rects;
    p->lmax = TIERS[p->kclass + 1].lcap;  // records per bin and launch of the highest tier
    const uint64_t cand_total = std::min<uint64_t>(0x3FFFFFF0ull, std::max<uint64_t>(32 * max_rects, 1 << 20));
    uint64_t hsz = 1;
    while (hsz < 2ull * cand_total) hsz <<= 1;
    p->hmask = (uint32_t)(hsz - 1);
    p->edge_cap = p->cfg.max_edges ? p->cfg.max_edges : std::max<uint64_t>(16 * max_rects, 1 << 20);
    p->cand_cap = p->cfg.direct_dedup ? 0u : (uint32_t)(cand_total / CSTRIPES);
    p->bytes_scratch = 12ull * CSTRIPES * p->cand_cap + 12ull * hsz + 4ull * hsz +
                       (p->cfg.host_edges ? 0ull : 12ull * p->edge_cap) + 32ull * (p->rec_cap + 1) +
                       sizeof(DevCounters) + sizeof(p->h_dom);
    int rc = 0;
    auto chk = [&](hipError_t e) { if (!rc) rc = hip_err(e); };
    if (p->cfg.budget_bytes && p->bytes_scratch > p->cfg.budget_bytes) rc = -ENOMEM;
    if (p->cand_cap) chk(hipMalloc(&p->cand, 12ull * CSTRIPES * p->cand_cap));
    chk(hipMalloc(&p->cstr, 64ull * CSTRIPES));
    chk(hipMalloc(&p->blk, 16ull * 4 * BLKSLOTS));
    chk(hipMalloc(&p->hkeys, 8ull * hsz));
    chk(hipMalloc(&p->hvals, 4ull * hsz));
    chk(hipMalloc(&p->eslot, 4ull * hsz));
    chk(hipMalloc(&p->dcnt, 4ull * DEDUP_BLOCKS));
    if (p->cfg.host_edges) {
        // k_final's edge stores land in pinned host memory: a launch's edges are on the host
        // when the launch completes, with no copy
        chk(hipHostMalloc(&p->edges_h, 8ull * p->edge_cap, hipHostMallocCoherent));
        chk(hipHostMalloc(&p->edom_h, 4ull * p->edge_cap, hipHostMallocCoherent));
        if (p->edges_h) chk(hipHostGetDevicePointer((void **)&p->edges, p->edges_h, 0));
        if (p->edom_h) chk(hipHostGetDevicePointer((void **)&p->edom, p->edom_h, 0));
    } else {
        chk(hipMalloc(&p->edges, 8ull * p->edge_cap));
        chk(hipMalloc(&p->edom, 4ull * p->edge_cap));
    }
    chk(hipMalloc(&p->ctr, sizeof(DevCounters)));
    chk(hipMalloc(&p->dom_err, 4ull * MAX_DOMS));
    chk(hipHostMalloc(&p->dom_err_host, 4ull * MAX_DOMS, hipHostMallocDefault));
    chk(hipMalloc(&p->gate_count, 64));
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, p->device) != hipSuccess || khz <= 0)
            khz = 100000;
        p->gate_limit = 1000ull * (unsigned long long)khz;  // 1 s
    }
    chk(hipMalloc(&p->d_dom, sizeof(p->h_dom)));
    chk(hipMalloc(&p->rec_dev_hdr, 32ull * (p->rec_cap + 1)));
    if (p->rec_dev_hdr) p->rec_dev = p->rec_dev_hdr + 2;
    chk(hipHostMalloc(&p->ctr_host, sizeof(DevCounters), hipHostMallocDefault));
    for (int b = 0; b < 2; ++b) {
        chk(hipHostMalloc(&p->rec_host_hdr[b], 32ull * (p->rec_cap + 1), hipHostMallocCoherent));
        if (p->rec_host_hdr[b]) chk(hipHostGetDevicePointer((void **)&p->rec_host_dev[b], p->rec_host_hdr[b], 0));
        chk(hipEventCreateWithFlags(&p->staged[b], hipEventDisableTiming));
    }
    if (p->rec_host_hdr[0]) p->rec_host = p->rec_host_hdr[0] + 2;
    chk(hipHostMalloc(&p->dom_ring, sizeof(DomDesc) * MAX_DOMS * NDRING, hipHostMallocDefault));
    chk(hipHostMalloc(&p->snap, sizeof(LaunchSnap) * NSNAP, hipHostMallocCoherent));
    if (p->snap) memset(p->snap, 0, sizeof(LaunchSnap) * NSNAP);
    for (hipEvent_t &e : p->dom_ev) chk(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (!rc) chk(hipMemsetAsync(p->hkeys, 0xFF, 8ull * hsz, p->stream));
    if (!rc) chk(hipMemsetAsync(p->hvals, 0xFF, 4ull * hsz, p->stream));
    if (!rc) chk(hipMemsetAsync(p->ctr, 0, sizeof(DevCounters), p->stream));
    if (!rc) chk(hipMemsetAsync(p->dom_err, 0, 4ull * MAX_DOMS, p->stream));
    if (!rc) chk(hipMemsetAsync(p->cstr, 0, 64ull * CSTRIPES, p->stream));
    if (!rc) chk(hipMemsetAsync(p->blk, 0, 16ull * 4 * BLKSLOTS, p->stream));
    {
        hipFuncAttributes at{};
        for (const void *f : {(const void *)k_resolve_reg<false, 1>, (const void *)k_resolve_reg<true, 1>,
                              (const void *)k_resolve_reg<false, 2>, (const void *)k_resolve_reg<true, 2>}) {
            chk(hipFuncGetAttributes(&at, f));
            if (!rc && at.sharedSizeBytes != 0) rc = -ENOEXEC;  // dynamic LDS must start at offset 0
        }
    }
    chk(set_lds<64, 128, 256>());
    chk(set_lds<256, 512, 512>());
    chk(set_lds<1024, 2048, 1024>());
    if (!rc) chk(hipStreamSynchronize(p->stream));
    if (rc) {
        destroy_pool(p);
        return rc;
    }
    *out = p;
    return 0;
}

int add_domain(xkrt_hip_pool *p, uint64_t ld, uint64_t s, uint32_t BX, uint32_t BY, xkrt_hip_dep_domain **out)
{
    if (ld == 0 || s == 0) return -EINVAL;
    const uint64_t LDs = ld * s;
    if (LDs > 0xFFFFFFFFull) return -ERANGE;
    if (!is_pow2(BX) || !is_pow2(BY)) return -EINVAL;
    uint32_t idx = MAX_DOMS;
    for (uint32_t i = 0; i < MAX_DOMS; ++i)
        if (!p->doms[i]) { idx = i; break; }
    if (idx == MAX_DOMS) return -ENOSPC;
    xkrt_hip_dep_domain *d = new xkrt_hip_dep_domain();
    d->pool = p;
    d->idx = idx;
    d->ld = ld; d->s = s; d->LDs = LDs;
    d->bxs = ilog2(BX); d->bys = ilog2(BY);
    d->nbx = (uint32_t)((LDs + BX - 1) / BX);
    const int rc = clear_domain_error(p, idx);  // the slot may have held a failed domain
    if (rc) {
        delete d;
        return rc;
    }
    p->doms[idx] = d;
    p->ndoms++;
    publish(p, d);
    *out = d;
    return 0;
}

void remove_domain(xkrt_hip_dep_domain *d)
{
    xkrt_hip_pool *p = d->pool;
    if (d->nby) free_bins(p, d->bin0, d->nbx * d->nby);
    p->doms[d->idx] = nullptr;
    p->h_dom[d->idx] = DomDesc{};
    p->dom_dirty = true;
    p->ndoms--;
    delete d;
}

// A fixed set of host threads that build the records of large batches in
// chunks (the calling thread takes chunks too).  Calls from several threads
// take turns.  A worker reads the job only inside a generation it joined
// under the lock, and a call sets up its job only once no worker is still
// inside the previous one.  The set lives until the process ends (never
// destroyed: no join at exit); a forked child, which has none of the
// threads, runs every chunk itself.
class HostWorkers {
  public:
    static HostWorkers &get()
    {
        static HostWorkers *w = new HostWorkers();
        return *w;
    }
    uint32_t threads() const { return getpid() == pid_ ? (uint32_t)th_.size() + 1 : 1u; }
    // f(c) for c in [0, n): returns once every call has returned
    template <typename F>
    void run(uint32_t n, F &&f)
    {
        if (getpid() != pid_) {
            for (uint32_t c = 0; c < n; ++c) f(c);
            return;
        }
        std::lock_guard<std::mutex> turn(call_mu_);
        std::function<void(uint32_t)> job(std::forward<F>(f));
        std::unique_lock<std::mutex> g(mu_);
        done_.wait(g, [&] { return busy_ == 0; });
        job_ = &job;
        n_ = n;
        next_.store(0, std::memory_order_relaxed);
        left_ = n;
        ++gen_;
        g.unlock();
        cv_.notify_all();
        work();
        g.lock();
        done_.wait(g, [&] { return left_ == 0 && busy_ == 0; });
        job_ = nullptr;
    }

  private:
    HostWorkers() : pid_(getpid())
    {
        uint32_t t = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        if (const char *e = getenv("XKRT_HIP_HOST_THREADS")) t = (uint32_t)std::max(1, std::min(64, atoi(e)));
        for (uint32_t i = 1; i < t; ++i) th_.emplace_back([this] { loop(); });
    }
    void work()
    {
        uint32_t c;
        while ((c = next_.fetch_add(1, std::memory_order_relaxed)) < n_) {
            (*job_)(c);
            std::lock_guard<std::mutex> g(mu_);
            if (--left_ == 0) done_.notify_all();
        }
    }
    void loop()
    {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> g(mu_);
        while (true) {
            cv_.wait(g, [&] { return gen_ != seen; });
            seen = gen_;
            ++busy_;
            g.unlock();
            work();
            g.lock();
            if (--busy_ == 0) done_.notify_all();
        }
    }
    std::vector<std::thread> th_;
    std::mutex call_mu_, mu_;
    std::condition_variable cv_, done_;
    std::function<void(uint32_t)> *job_ = nullptr;
    uint32_t n_ = 0, left_ = 0, busy_ = 0;
    std::atomic<uint32_t> next_{0};
    uint64_t gen_ = 0;
    const pid_t pid_;
};

// What building the records of accesses [0, i) leaves for access i: records so far, the task
// number, the widest rect in dim 0 and whether a rect started at column 0 (the two feed the
// refusals below), and the column range.
struct BuildState {
    uint64_t k = 0;
    uint32_t tl = 0;
    uint64_t maxx = 0;
    bool col0 = false;
    uint64_t lo = UINT64_MAX, hi = 0;
};

// the records of accesses [i0, i1) of a batch from state st (out == nullptr: count only)
int build_range(const xkrt_hip_dep_domain *d, const xkrt_hip_access_t *acc, uint64_t i0, uint64_t i1,
                uint32_t first_id, uint64_t cap, xkrt_hip_rect_t *out, const uint32_t *ids, BuildState &st)
{
    const uint64_t LDs = d->LDs, s = d->s;
    uint64_t k = st.k, lo = st.lo, hi = st.hi, maxx = st.maxx;
    uint32_t tl = st.tl;
    bool col0 = st.col0;
    int rc = 0;
    for (uint64_t i = i0; i < i1; ++i) {
        const xkrt_hip_access_t &q = acc[i];
        if (i && q.task != acc[i - 1].task) {
            if (q.task < acc[i - 1].task) { rc = -EINVAL; break; }
            tl++;
        }
        if (q.m == 0 || (q.n == 0 && q.kind == XKRT_HIP_KIND_TILE) || q.kind > XKRT_HIP_KIND_SEGMENT) { rc = -EINVAL; break; }
        if (d->segments && q.kind != XKRT_HIP_KIND_SEGMENT) { rc = -EINVAL; break; }
        if (tl > 0xFFFFFFu) { rc = -ERANGE; break; }
        uint64_t r[3][4];
        int nr;
        uint32_t flag = 0;
        const uint64_t X0 = q.addr % LDs, Y0 = q.addr / LDs, ms = q.m * s;
        if (q.kind == XKRT_HIP_KIND_SEGMENT) {
            // access.hpp:77-139 interval_to_rects in the domain frame; empty rects
            // are skipped, as the tree's insert and intersect skip them.
            // In a segment domain the rects are PUT exactly: the interval tree marks
            // the nodes that meet the segment itself (dependency-tree.hpp:241-256
            // with K = 1), so each point of [addr, addr + m) is written or read once.
            // In a BLAS domain, PUT tests
            // the tree's nodes against the segment's own rects, built with
            // ld = SIZE_MAX (access.hpp:495-513): [addr, addr + m) x [0, 1) and two
            // empty rects, one of them [0, SIZE_MAX) x [1, 1).  Interval overlap is
            // strict (interval.hpp:110-119), so these meet exactly the nodes that
            // start at column 0 -- which exist only once some access touched the
            // first ld * sizeof_type bytes of memory.  Before that the PUT marks
            // nothing and the segment only links; after it, it is refused.
            if (!d->segments && (q.addr < LDs || col0)) { rc = -ENOTSUP; break; }
            const uint64_t size = q.m, x1 = std::min(X0 + size, LDs), dx10 = x1 - X0;
            const uint64_t x3 = (size - dx10) % LDs;
            const uint64_t y1 = Y0 + 1, y3 = Y0 + (size - dx10 - x3) / LDs + 2, y2 = y3 - 1;
            const uint64_t c[3][4] = {{X0, x1, Y0, y1}, {0, LDs, y1, y2}, {0, x3, y2, y3}};
            nr = 0;
            for (int j = 0; j < 3; ++j)
                if (c[j][0] < c[j][1] && c[j][2] < c[j][3]) {
                    for (int w = 0; w < 4; ++w) r[nr][w] = c[j][w];
                    ++nr;
                }
            if (!d->segments) flag = 0x80u;  // LINK only: its PUT marks no node of this domain
        } else if (X0 + ms <= LDs) {
            r[0][0] = X0; r[0][1] = X0 + ms; r[0][2] = Y0; r[0][3] = Y0 + q.n;
            nr = 1;
        } else {
            if (ms > LDs) { rc = -EINVAL; break; }
            r[0][0] = X0; r[0][1] = LDs; r[0][2] = Y0; r[0][3] = Y0 + q.n;
            r[1][0] = 0; r[1][1] = ms - (LDs - X0); r[1][2] = Y0 + 1; r[1][3] = Y0 + q.n + 1;
            nr = 2;
            // a prior rect spanning [x3 - 1, x0] in dim 0 makes the reference
            // mark split fragments beyond the tile
            if (maxx > X0 - r[1][1]) { rc = -ENOTSUP; break; }
        }
        for (int j = 0; j < nr && !rc; ++j) {
            if (r[j][2] < d->ybase || r[j][3] - d->ybase > 0xFFFFFFFFull) { rc = -ERANGE; break; }
            maxx = std::max(maxx, r[j][1] - r[j][0]);
            col0 = col0 || r[j][2] == 0;
            if (out && k < cap) {
                xkrt_hip_rect_t &o = out[k];
                o.x0 = (uint32_t)r[j][0]; o.x1 = (uint32_t)r[j][1];
                o.y0 = (uint32_t)(r[j][2] - d->ybase); o.y1 = (uint32_t)(r[j][3] - d->ybase);
                o.acc = ids ? ids[i] : first_id + (uint32_t)i;
                o.tm = (tl << 8) | q.mode | flag;
                o.dom = d->idx;
                o.pad = 0;
            }
            lo = std::min(lo, r[j][2] - d->ybase);
            hi = std::max(hi, r[j][3] - d->ybase);
            ++k;
        }
        if (rc) break;
    }
    st.k = k; st.tl = tl; st.maxx = maxx; st.col0 = col0; st.lo = lo; st.hi = hi;
    return rc;
}

// batches of at least this many accesses are built by the host workers, in chunks of at least
// BUILD_CHUNK accesses
constexpr uint64_t BUILD_PARALLEL_MIN = 1u << 16, BUILD_CHUNK = 1u << 14;

// host records (AoS) of n accesses into out[0 .. cap): rects in the domain
// frame, access ids from first_id, launch-local task numbers from task0;
// *ntask = tasks used.  A large batch is built in chunks on several threads:
// a first pass counts each chunk's records and task changes and takes its
// widest rect and column-0 flag (all independent of what came before), then
// each chunk is built from the state its predecessors leave.  A chunk that
// fails in the first pass fails in the second as well (the refusals only get
// likelier with more state behind them), and the batch's error is the one of
// its earliest failing access, as a single pass would find.
int build_records(xkrt_hip_dep_domain *d, const xkrt_hip_access_t *acc, uint64_t n, uint32_t first_id, uint32_t task0,
                  uint64_t cap, xkrt_hip_rect_t *out, uint64_t *nrect, uint32_t *ylo, uint32_t *yhi, uint32_t *ntask,
                  const uint32_t *ids = nullptr)
{
    // caller ids: increasing across the domain's batches (the lowest id of a task = its lowest slot)
    if (ids)
        for (uint64_t i = 0; i < n; ++i)
            if ((i ? ids[i] <= ids[i - 1] : (d->any_id && ids[0] <= d->last_id))) return -EINVAL;
    const uint64_t LDs = d->LDs;
    // frame: fixed by the first access ever seen (or given at creation)
    if (!d->framed && n) {
        uint64_t ymin = UINT64_MAX;
        for (uint64_t i = 0; i < n; ++i) ymin = std::min(ymin, acc[i].addr / LDs);
        d->ybase = ymin > (1ull << 30) ? ymin - (1ull << 30) : 0;
        d->framed = true;
    }
    BuildState st;
    st.tl = task0;
    st.maxx = d->max_xlen;
    st.col0 = d->col0;
    HostWorkers &hw = HostWorkers::get();
    const uint64_t nch = n >= BUILD_PARALLEL_MIN && hw.threads() > 1
                             ? std::min<uint64_t>(4ull * hw.threads(), n / BUILD_CHUNK) : 1;
    if (nch <= 1) {
        const int rc = build_range(d, acc, 0, n, first_id, cap, out, ids, st);
        if (rc) return rc;
    } else {
        const uint64_t per = (n + nch - 1) / nch;
        std::vector<BuildState> part(nch), fin(nch);
        std::vector<int> rc1(nch, 0), rc2(nch, 0);
        hw.run((uint32_t)nch, [&](uint32_t c) {
            const uint64_t a = c * per, b = std::min(n, a + per);
            rc1[c] = build_range(d, acc, a, b, first_id, 0, nullptr, ids, part[c]);
        });
        BuildState at = st;
        for (uint64_t c = 0; c < nch; ++c) {
            fin[c] = at;
            at.k += part[c].k;
            at.tl += part[c].tl;
            at.maxx = std::max(at.maxx, part[c].maxx);
            at.col0 = at.col0 || part[c].col0;
        }
        hw.run((uint32_t)nch, [&](uint32_t c) {
            const uint64_t a = c * per, b = std::min(n, a + per);
            rc2[c] = build_range(d, acc, a, b, first_id, cap, out, ids, fin[c]);
        });
        for (uint64_t c = 0; c < nch; ++c)
            if (rc2[c]) return rc2[c];
        st = fin[nch - 1];
        for (uint64_t c = 0; c + 1 < nch; ++c) {
            st.lo = std::min(st.lo, fin[c].lo);
            st.hi = std::max(st.hi, fin[c].hi);
        }
    }
    d->max_xlen = st.maxx;
    d->col0 = st.col0;
    if (ids && n) {
        d->any_id = true;
        d->last_id = ids[n - 1];
    }
    *nrect = st.k;
    if (ylo) *ylo = n ? (uint32_t)st.lo : 0;
    if (yhi) *yhi = n ? (uint32_t)st.hi : 0;
    if (ntask) *ntask = n ? st.tl - task0 + 1 : 0;
    return 0;
}

// append a batch to the pool's open launch (host staging)
int stage(xkrt_hip_pool *p, xkrt_hip_dep_domain *d, const xkrt_hip_access_t *acc, uint64_t n, uint64_t *first_id,
          const uint32_t *ids = nullptr)
{
    if (p->poisoned) return -EINVAL;
    if (d->err) return -ECANCELED;  // a failed domain takes no batch until it is reset
    if (first_id) *first_id = d->next_id;
    if (n == 0) return 0;
    if (n > 0xFFFFFFFFull || d->next_id + n > 0xFFFFFFFFull) return -ERANGE;
    if (p->open_nrec + n > p->rec_cap) return p->open_nrec ? -ENOSPC : -EINVAL;
    // the staging buffer must be free: wait for the previous upload
    const uint32_t hb = p->cur_host;
    if (p->staged_pending[hb]) {
        HIPCHK(hipEventSynchronize(p->staged[hb]));
        p->staged_pending[hb] = false;
    }
    if (p->buf_ticket[hb] > p->done) {  // a small launch still reads this buffer in place
        const int rw = retire(p, p->buf_ticket[hb]);
        if (rw) return rw;
    }
    uint64_t nrect = 0;
    uint32_t ylo = 0, yhi = 0, ntask = 0;
    const uint64_t saved_max = d->max_xlen, room = p->rec_cap - p->open_nrec;
    const bool was_framed = d->framed, saved_col0 = d->col0, saved_any = d->any_id;
    const uint32_t saved_last = d->last_id;
    int rc = build_records(d, acc, n, (uint32_t)d->next_id, p->open_tasks, room,
                           (xkrt_hip_rect_t *)p->rec_host + p->open_nrec, &nrect, &ylo, &yhi, &ntask, ids);
    // the batch's rects (1 to 3 per access) must fit the open launch whole
    if (!rc && nrect > room) rc = p->open_nrec ? -ENOSPC : -EINVAL;
    if (!rc) rc = ensure_grid(d, ylo, yhi);
    if (rc) {
        d->max_xlen = saved_max;
        d->col0 = saved_col0;
        d->any_id = saved_any;
        d->last_id = saved_last;
        if (!was_framed) d->framed = false;
        return rc;
    }
    p->open_nrec += (uint32_t)nrect;
    p->open_tasks += ntask;
    p->open_acc += n;
    p->open_batches++;
    d->next_id += n;
    d->n_acc += n;
    d->n_rect += nrect;
    return 0;
}

int launch_open(xkrt_hip_pool *p)
{
    if (p->open_nrec == 0) return 0;
    const uint32_t n = p->open_nrec;
    const uint32_t hb = p->cur_host;
    *(uint32_t *)p->rec_host_hdr[hb] = n;
    // a small launch reads its records in place, from pinned host memory: no upload
    const bool small = n <= p->cfg.small_launch_rects;
    if (!small) {
        HIPCHK(hipMemcpyAsync(p->rec_dev_hdr, p->rec_host_hdr[hb], 32ull * (n + 1), hipMemcpyHostToDevice,
                              p->stream));
        HIPCHK(hipEventRecord(p->staged[hb], p->stream));
        p->staged_pending[hb] = true;
    } else {
        p->buf_ticket[hb] = p->n_launch + 1;  // the launch enqueued below
    }
    p->cur_host = hb ^ 1u;
    p->rec_host = p->rec_host_hdr[p->cur_host] + 2;
    p->open_nrec = 0;
    p->open_tasks = 0;
    p->open_acc = 0;
    p->open_batches = 0;
    const int rc = small ? enqueue_small(p, p->rec_host_dev[hb], n) : enqueue(p, p->rec_dev, n, true);
    if (rc) p->poisoned = true;
    return rc;
}

// Arms the gate: enqueues a small launch over the staging buffer the next
// stage writes into, held by k_gate until gate_ring (or gate_cancel).  Nothing
// else may be enqueued on the stream while it is armed.
int gate_arm(xkrt_hip_pool *p)
{
    if (p->gate_armed) return 0;
    if (p->open_nrec) return -EBUSY;
    const uint32_t hb = p->cur_host;
    if (p->staged_pending[hb]) {
        HIPCHK(hipEventSynchronize(p->staged[hb]));
        p->staged_pending[hb] = false;
    }
    if (p->buf_ticket[hb] > p->done) {
        const int rw = retire(p, p->buf_ticket[hb]);
        if (rw) return rw;
    }
    int rc = push_domains(p);
    if (rc) return rc;
    volatile uint32_t *h = (volatile uint32_t *)p->rec_host_hdr[hb];
    p->gate_prev_ticket[hb] = p->buf_ticket[hb];  // that launch completed: its gate result is final
    p->gate_prev_result[hb] = h[2];
    const uint32_t seq = ++p->gate_seq;
    h[0] = 0;
    h[1] = seq - 1;
    h[2] = GATE_WAIT;
    std::atomic_thread_fence(std::memory_order_seq_cst);
    hipLaunchKernelGGL(k_gate, dim3(1), dim3(64), 0, p->stream, (uint32_t *)p->rec_host_dev[hb], seq, p->gate_limit,
                       p->gate_count);
    KArgs a = make_args(p, p->rec_host_dev[hb] + 2, NONE);
    a.nrec_ptr = p->gate_count;
    // grids for any count: k_count strides over the records
    launch_kernels(p, a, p->cfg.small_launch_rects, true, nullptr);
    rc = ticket(p);
    if (rc) return rc;
    p->gate_armed = true;
    p->gate_hb = hb;
    p->gate_ticket = p->n_launch;
    return 0;
}

// arms the gate when the staging buffer it would take is free, without
// waiting: with two launches in flight that buffer may still be read by the
// older one; the engine tries again later
int gate_try_arm(xkrt_hip_pool *p)
{
    if (p->gate_armed || p->open_nrec) return 0;
    const uint32_t hb = p->cur_host;
    if (p->staged_pending[hb]) {
        const hipError_t q = hipEventQuery(p->staged[hb]);
        if (q == hipErrorNotReady) return 0;
        if (q != hipSuccess) return hip_err(q);
        p->staged_pending[hb] = false;
    }
    if (p->buf_ticket[hb] > p->done) {
        const int rc = retire(p, 0);
        if (rc) return rc;
        if (p->buf_ticket[hb] > p->done) return 0;
    }
    return gate_arm(p);
}

// the gate result (GATE_*) of completed launch `ticket`, gated on buffer hb
uint32_t gate_result(const xkrt_hip_pool *p, uint32_t hb, uint64_t ticket)
{
    if (p->gate_prev_ticket[hb] == ticket) return p->gate_prev_result[hb];
    return ((volatile uint32_t *)p->rec_host_hdr[hb])[2];
}

// opens the armed gate on the staged records (count 0: an empty launch)
void gate_open(xkrt_hip_pool *p, uint32_t n)
{
    const uint32_t hb = p->gate_hb;
    volatile uint32_t *h = (volatile uint32_t *)p->rec_host_hdr[hb];
    h[0] = n;
    std::atomic_thread_fence(std::memory_order_seq_cst);  // records and count before the sequence number
    h[1] = p->gate_seq;
    p->buf_ticket[hb] = p->gate_ticket;
    p->gate_armed = false;
}

// resolves the staged records through the armed gate; *ticket = the launch
int gate_ring(xkrt_hip_pool *p, uint64_t *ticket_out)
{
    const uint32_t n = p->open_nrec;
    gate_open(p, n);
    *ticket_out = p->gate_ticket;
    p->cur_host = p->gate_hb ^ 1u;
    p->rec_host = p->rec_host_hdr[p->cur_host] + 2;
    p->open_nrec = 0;
    p->open_tasks = 0;
    p->open_acc = 0;
    p->open_batches = 0;
    return 0;
}

// lets the armed gate's launch run empty and waits until the gate has seen
// it, so the staging buffer's header is the host's again
int gate_cancel(xkrt_hip_pool *p)
{
    if (!p->gate_armed) return 0;
    const uint32_t hb = p->gate_hb;
    gate_open(p, 0);
    volatile uint32_t *h = (volatile uint32_t *)p->rec_host_hdr[hb];
    const auto t0 = std::chrono::steady_clock::now();
    while (h[2] == GATE_WAIT) {
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(5)) return -EIO;
        std::this_thread::yield();
    }
    if (!p->open_nrec) {  // nothing staged in this buffer: the next stage takes the other one
        p->cur_host = hb ^ 1u;
        p->rec_host = p->rec_host_hdr[p->cur_host] + 2;
    }
    return 0;
}

} // namespace

extern "C" {

int xkrt_hip_init(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device) return -ENODEV;
    return hip_err(hipSetDevice(device));
}

// ---- pools

int xkrt_hip_pool_create(const xkrt_hip_pool_config_t *cfg, void *stream, xkrt_hip_pool_t *out)
{
    if (!out) return -EINVAL;
    *out = nullptr;
    return create_pool(cfg, stream, out);
}

int xkrt_hip_pool_destroy(xkrt_hip_pool_t p)
{
    if (!p) return -EINVAL;
    if (p->server) return -EBUSY;  // destroy the serving engine first
    for (uint32_t i = 0; i < MAX_DOMS; ++i)
        if (p->doms[i]) remove_domain(p->doms[i]);
    destroy_pool(p);
    return 0;
}

int xkrt_hip_pool_domain_create(xkrt_hip_pool_t p, uint64_t ld, uint64_t sizeof_type, uint32_t bin_x_bytes,
                                uint32_t bin_y_cols, uint64_t y_lo, uint64_t y_hi, xkrt_hip_dep_domain_t *out)
{
    if (!p || !out || y_hi < y_lo || y_hi - y_lo >= (1ull << 32)) return -EINVAL;
    *out = nullptr;
    if (p->server) return -EBUSY;
    xkrt_hip_dep_domain *d = nullptr;
    int rc = add_domain(p, ld, sizeof_type, bin_x_bytes ? bin_x_bytes : 1024u, bin_y_cols ? bin_y_cols : 128u, &d);
    if (rc) return rc;
    if (y_lo == y_hi) {  // no range given: framed by its first batch, the grid grows with the batches
        *out = d;
        return 0;
    }
    d->ybase = y_lo;
    d->framed = true;
    rc = ensure_grid(d, 0, (uint32_t)(y_hi - y_lo));
    if (rc) {
        remove_domain(d);
        return rc;  // -ENOMEM: the pool's budget cannot admit the domain now
    }
    d->fixed_frame = true;
    *out = d;
    return 0;
}

int xkrt_hip_pool_seg_domain_create(xkrt_hip_pool_t p, uint64_t col_bytes, uint32_t bin_x_bytes, uint32_t bin_y_cols,
                                    uint64_t y_lo, uint64_t y_hi, xkrt_hip_dep_domain_t *out)
{
    const int rc = xkrt_hip_pool_domain_create(p, col_bytes, 1, bin_x_bytes, bin_y_cols, y_lo, y_hi, out);
    if (!rc) (*out)->segments = true;
    return rc;
}

int xkrt_hip_pool_stage(xkrt_hip_pool_t p, xkrt_hip_dep_domain_t d, const xkrt_hip_access_t *acc, uint64_t n,
                        uint64_t *first_id)
{
    if (!p || !d || d->pool != p || (!acc && n)) return -EINVAL;
    if (p->server) return -EBUSY;
    return stage(p, d, acc, n, first_id);
}

int xkrt_hip_pool_stage_group(xkrt_hip_pool_t p, uint32_t nd, const xkrt_hip_dep_domain_t *d,
                              const xkrt_hip_access_t *const *acc, const uint32_t *const *ids, const uint64_t *n,
                              uint64_t *first_id)
{
    if (!p || (nd && (!d || !acc || !n))) return -EINVAL;
    if (p->server) return -EBUSY;
    for (uint32_t k = 0; k < nd; ++k)
        if (!d[k] || d[k]->pool != p || (!acc[k] && n[k])) return -EINVAL;
    // what staging changes, restored when a later batch of the group is refused
    struct Saved {
        uint64_t next_id, n_acc, n_rect, max_xlen, ybase;
        bool col0, framed, any_id;
        uint32_t last_id;
    };
    std::vector<Saved> saved(nd);
    for (uint32_t k = 0; k < nd; ++k) {
        const xkrt_hip_dep_domain *q = d[k];
        saved[k] = {q->next_id, q->n_acc, q->n_rect, q->max_xlen, q->ybase, q->col0, q->framed, q->any_id, q->last_id};
    }
    const uint32_t o_nrec = p->open_nrec, o_tasks = p->open_tasks;
    const uint64_t o_acc = p->open_acc, o_batches = p->open_batches;
    for (uint32_t k = 0; k < nd; ++k) {
        const int rc = stage(p, d[k], acc[k], n[k], first_id ? first_id + k : nullptr, ids ? ids[k] : nullptr);
        if (rc) {
            for (uint32_t j = nd; j-- > 0;) {  // newest first: a domain named twice gets its oldest state back
                xkrt_hip_dep_domain *q = d[j];
                const Saved &v = saved[j];
                q->next_id = v.next_id; q->n_acc = v.n_acc; q->n_rect = v.n_rect;
                q->max_xlen = v.max_xlen; q->col0 = v.col0;
                q->any_id = v.any_id; q->last_id = v.last_id;
                // a grid allocated meanwhile holds no fragment: it stays, with the frame it was sized in
                if (!q->nby) { q->framed = v.framed; q->ybase = v.ybase; }
            }
            p->open_nrec = o_nrec; p->open_tasks = o_tasks;
            p->open_acc = o_acc; p->open_batches = o_batches;
            return rc;
        }
    }
    return 0;
}

int xkrt_hip_pool_launch(xkrt_hip_pool_t p, uint64_t *ticket)
{
    if (!p) return -EINVAL;
    if (p->server) return -EBUSY;
    if (p->poisoned) return -EINVAL;
    const int rc = launch_open(p);
    if (ticket) *ticket = p->n_launch;
    return rc;
}

int xkrt_hip_pool_poll(xkrt_hip_pool_t p, uint64_t *done)
{
    if (!p || !done) return -EINVAL;
    if (p->server) return -EBUSY;
    const int rc = retire(p, 0);
    *done = p->done;
    return rc;
}

int xkrt_hip_pool_wait(xkrt_hip_pool_t p, uint64_t ticket)
{
    if (!p) return -EINVAL;
    if (p->server) return -EBUSY;
    return retire(p, ticket);
}

int xkrt_hip_pool_open(xkrt_hip_pool_t p, uint64_t *nrect, uint64_t *naccess, uint64_t *nbatches)
{
    if (!p) return -EINVAL;
    if (nrect) *nrect = p->open_nrec;
    if (naccess) *naccess = p->open_acc;
    if (nbatches) *nbatches = p->open_batches;
    return 0;
}

int xkrt_hip_pool_sync(xkrt_hip_pool_t p, uint64_t *nedges)
{
    if (!p) return -EINVAL;
    if (p->server) return -EBUSY;
    const int rc = read_counters(p);
    if (nedges) *nedges = p->ctr_host->edge_total;
    return rc;
}

int xkrt_hip_pool_edges_copy(xkrt_hip_pool_t p, uint32_t *edges, uint32_t *doms, uint64_t cap, uint64_t *n)
{
    if (!p || !n || (cap && (!edges || !doms))) return -EINVAL;
    if (p->server) return -EBUSY;
    int rc = read_counters(p);
    const uint64_t ne = p->ctr_host->edge_total;
    *n = ne;
    if (rc) return rc;
    const uint64_t m = std::min(ne, cap);
    if (m) {
        if (p->edges_h) {  // already in host memory (the stream is synchronised)
            memcpy(edges, p->edges_h, 8 * m);
            memcpy(doms, p->edom_h, 4 * m);
        } else {
            HIPCHK(hipMemcpyAsync(edges, p->edges, 8 * m, hipMemcpyDeviceToHost, p->stream));
            HIPCHK(hipMemcpyAsync(doms, p->edom, 4 * m, hipMemcpyDeviceToHost, p->stream));
            HIPCHK(hipStreamSynchronize(p->stream));
        }
        // a failed domain's edges are not its reference edges: they are dropped
        bool any = false;
        for (uint32_t i = 0; i < MAX_DOMS; ++i) any = any || (p->doms[i] && p->doms[i]->err);
        if (any) {
            uint64_t k = 0;
            for (uint64_t i = 0; i < m; ++i) {
                const xkrt_hip_dep_domain *d = p->doms[doms[i] & (MAX_DOMS - 1)];
                if (d && d->err) continue;
                edges[2 * k] = edges[2 * i];
                edges[2 * k + 1] = edges[2 * i + 1];
                doms[k++] = doms[i];
            }
            *n = k + (ne - m);
        }
    }
    hipLaunchKernelGGL(k_consume, dim3(1), dim3(64), 0, p->stream, p->ctr);
    return hip_err(hipGetLastError());
}

int xkrt_hip_pool_edges_device(xkrt_hip_pool_t p, const uint32_t **edges, const uint32_t **doms,
                               const uint64_t **count)
{
    if (!p || !edges || !doms || !count) return -EINVAL;
    *edges = p->edges;
    *doms = p->edom;
    *count = (const uint64_t *)&p->ctr->edge_total;
    return 0;
}

int xkrt_hip_pool_edges_consume(xkrt_hip_pool_t p)
{
    if (!p) return -EINVAL;
    if (p->server) return -EBUSY;
    hipLaunchKernelGGL(k_consume, dim3(1), dim3(64), 0, p->stream, p->ctr);
    return hip_err(hipGetLastError());
}

int xkrt_hip_pool_stats(xkrt_hip_pool_t p, xkrt_hip_pool_stats_t *out)
{
    if (!p || !out) return -EINVAL;
    memset(out, 0, sizeof(*out));
    uint64_t freeb = 0;
    for (auto &r : p->free_ranges) freeb += r.second;
    out->domains = p->ndoms;
    out->bins_capacity = p->cap_bins;
    out->bins_used = p->cap_bins - freeb;
    out->device_bytes = (uint64_t)p->cap_bins * bin_bytes(p) + p->bytes_scratch;
    out->bytes_per_bin = bin_bytes(p);
    out->launches = p->n_launch;
    out->graph = (p->graph_valid[0] || p->graph_valid[1] || p->graph_valid[2]) ? 1 : 0;
    out->poisoned = p->poisoned ? 1 : 0;
    return 0;
}

int xkrt_hip_pool_timing(xkrt_hip_pool_t p, int enable)
{
    if (!p) return -EINVAL;
    p->timing = enable != 0;
    return 0;
}

int xkrt_hip_pool_launch_times(xkrt_hip_pool_t p, float *ms, uint64_t cap, uint64_t *n, int clear)
{
    if (!p || !n || (cap && !ms)) return -EINVAL;
    HIPCHK(hipStreamSynchronize(p->stream));
    harvest_events(p);
    *n = p->launch_ms.size();
    for (uint64_t i = 0; i < std::min<uint64_t>(cap, p->launch_ms.size()); ++i) ms[i] = p->launch_ms[i];
    if (clear) p->launch_ms.clear();
    return 0;
}

void *xkrt_hip_pool_stream(xkrt_hip_pool_t p) { return p ? (void *)p->stream : nullptr; }

int xkrt_hip_dep_index(xkrt_hip_dep_domain_t d, uint32_t *idx)
{
    if (!d || !idx) return -EINVAL;
    *idx = d->idx;
    return 0;
}

// ---- single domains (a private pool each)

namespace {

int create_single(uint64_t ld, uint64_t sizeof_type, const xkrt_hip_dep_config_t *cfg, void *stream, bool segments,
                  xkrt_hip_dep_domain_t *out)
{
    if (!out || ld == 0 || sizeof_type == 0) return -EINVAL;
    *out = nullptr;
    if (ld * sizeof_type > 0xFFFFFFFFull) return -ERANGE;
    xkrt_hip_dep_config_t c{};
    if (cfg) c = *cfg;
    const uint32_t BX = c.bin_x_bytes ? c.bin_x_bytes : 1024u, BY = c.bin_y_cols ? c.bin_y_cols : 128u;
    if (!is_pow2(BX) || !is_pow2(BY) || c.kernel > 2) return -EINVAL;
    xkrt_hip_pool_config_t pc{};
    pc.kernel = c.kernel;
    pc.max_rects = c.max_rects;
    pc.max_edges = c.max_edges;
    pc.budget_bytes = c.budget_bytes;
    xkrt_hip_pool *p = nullptr;
    int rc = create_pool(&pc, stream, &p);
    if (rc) return rc;
    xkrt_hip_dep_domain *d = nullptr;
    rc = add_domain(p, ld, sizeof_type, BX, BY, &d);
    if (!rc) {
        d->private_pool = true;
        d->segments = segments;
        if (c.y_hint_lo < c.y_hint_hi) {
            d->ybase = c.y_hint_lo;
            d->framed = true;
            const uint64_t span = c.y_hint_hi - c.y_hint_lo;
            rc = span >= (1ull << 32) ? -ERANGE : ensure_grid(d, 0, (uint32_t)span);
        }
    }
    if (!rc) rc = hip_err(hipStreamSynchronize(p->stream));
    if (rc) {
        xkrt_hip_pool_destroy(p);
        return rc;
    }
    *out = d;
    return 0;
}

} // namespace

int xkrt_hip_dep_domain_create(uint64_t ld, uint64_t sizeof_type, const xkrt_hip_dep_config_t *cfg, void *stream,
                               xkrt_hip_dep_domain_t *out)
{
    return create_single(ld, sizeof_type, cfg, stream, false, out);
}

int xkrt_hip_seg_domain_create(uint64_t col_bytes, const xkrt_hip_dep_config_t *cfg, void *stream,
                               xkrt_hip_dep_domain_t *out)
{
    return create_single(col_bytes, 1, cfg, stream, true, out);
}

int xkrt_hip_dep_replay_segments(xkrt_hip_dep_domain_t d, const xkrt_hip_access_t *acc, uint64_t n)
{
    if (!d || (!acc && n) || d->segments) return -EINVAL;
    // task.cc:177-198 puts the uncompleted segment accesses into a new BLAS tree, with no
    // LINK.  The PUT tests nodes against the segment's own rects in the ld = SIZE_MAX
    // frame (access.hpp:495-513), which meet only nodes starting at column 0: a segment
    // below ld * sizeof_type is refused, any other marks nothing.  Its rects are still
    // inserted, splitting nodes, and a later wrapped tile must see their dim-0 extent.
    const uint64_t LDs = d->LDs;
    uint64_t maxx = d->max_xlen;
    for (uint64_t i = 0; i < n; ++i) {
        const xkrt_hip_access_t &q = acc[i];
        if (q.kind != XKRT_HIP_KIND_SEGMENT || q.m == 0) return -EINVAL;
        if (q.addr < LDs || d->col0) return -ENOTSUP;
        const uint64_t X0 = q.addr % LDs, x1 = std::min(X0 + q.m, LDs), rest = q.m - (x1 - X0);
        maxx = std::max(maxx, x1 - X0);
        if (rest >= LDs) maxx = LDs;  // a whole-column middle rect
        else if (rest) maxx = std::max(maxx, rest);
    }
    d->max_xlen = maxx;
    return 0;
}

int xkrt_hip_dep_domain_destroy(xkrt_hip_dep_domain_t d)
{
    if (!d) return -EINVAL;
    xkrt_hip_pool *p = d->pool;
    if (d->private_pool) return xkrt_hip_pool_destroy(p);
    (void)hipStreamSynchronize(p->stream);  // no launch may still read the domain's bins
    remove_domain(d);
    return 0;
}

int xkrt_hip_dep_frame(xkrt_hip_dep_domain_t d, uint64_t *ybase)
{
    if (!d || !ybase) return -EINVAL;
    *ybase = d->ybase;
    return d->framed ? 0 : -EINVAL;
}

int xkrt_hip_dep_build_records(xkrt_hip_dep_domain_t d, const xkrt_hip_access_t *acc, uint64_t n, uint32_t first_id,
                               uint64_t cap, xkrt_hip_rect_t *out, uint64_t *nrect, uint32_t *ylo, uint32_t *yhi)
{
    if (!d || (!acc && n) || !nrect || (cap && !out)) return -EINVAL;
    return build_records(d, acc, n, first_id, 0, cap, out, nrect, ylo, yhi, nullptr);
}

int xkrt_hip_dep_submit_device(xkrt_hip_dep_domain_t d, const xkrt_hip_rect_t *rec, uint64_t nrect, uint64_t naccess,
                               uint32_t ylo, uint32_t yhi)
{
    if (!d || !rec || ylo > yhi) return -EINVAL;
    xkrt_hip_pool *p = d->pool;
    if (p->poisoned) return -EINVAL;
    if (d->err) return domain_error(d);
    if (nrect == 0) return 0;
    if (nrect > p->cfg.max_rects) return -EINVAL;
    if (!d->framed) d->framed = true;  // device records are already in the frame (ybase 0 unless hinted)
    int rc = ensure_grid(d, ylo, yhi);
    if (rc) return rc;
    rc = enqueue(p, (const uint4 *)rec, (uint32_t)nrect, false);
    if (rc) { p->poisoned = true; return rc; }
    d->n_acc += naccess;
    d->n_rect += nrect;
    return 0;
}

int xkrt_hip_dep_submit(xkrt_hip_dep_domain_t d, const xkrt_hip_access_t *acc, uint64_t n, uint64_t *first_id)
{
    if (!d || (!acc && n)) return -EINVAL;
    xkrt_hip_pool *p = d->pool;
    if (p->poisoned) return -EINVAL;
    if (d->err) return domain_error(d);
    int rc = stage(p, d, acc, n, first_id);
    if (rc == -ENOSPC) {  // the open launch is full: launch it, then stage again
        rc = launch_open(p);
        if (!rc) rc = stage(p, d, acc, n, first_id);
    }
    if (rc) return rc;
    return launch_open(p);
}

int xkrt_hip_dep_sync(xkrt_hip_dep_domain_t d, uint64_t *nedges)
{
    if (!d) return -EINVAL;
    xkrt_hip_pool *p = d->pool;
    if (p->server) return -EBUSY;
    const int rc = read_counters(p);
    if (nedges) *nedges = p->ctr_host->edge_total;
    return rc ? rc : domain_error(d);
}

int xkrt_hip_dep_edges_device(xkrt_hip_dep_domain_t d, const uint32_t **edges, uint64_t *n)
{
    if (!d || !edges || !n) return -EINVAL;
    xkrt_hip_pool *p = d->pool;
    int rc = read_counters(p);
    *edges = p->edges;
    *n = p->ctr_host->edge_total;
    return rc ? rc : domain_error(d);
}

int xkrt_hip_dep_edges_consume(xkrt_hip_dep_domain_t d)
{
    if (!d) return -EINVAL;
    hipLaunchKernelGGL(k_consume, dim3(1), dim3(64), 0, d->pool->stream, d->pool->ctr);
    return hip_err(hipGetLastError());
}

int xkrt_hip_dep_edges_copy(xkrt_hip_dep_domain_t d, uint32_t *host_edges, uint64_t cap, uint64_t *n)
{
    if (!d || !n || (cap && !host_edges)) return -EINVAL;
    xkrt_hip_pool *p = d->pool;
    if (!d->private_pool) return -EINVAL;  // a shared pool mixes domains: xkrt_hip_pool_edges_copy
    int rc = read_counters(p);
    const uint64_t ne = p->ctr_host->edge_total;
    *n = ne;
    if (!rc) rc = domain_error(d);
    if (rc) return rc;
    const uint64_t m = std::min(ne, cap);
    if (m && p->edges_h) {
        memcpy(host_edges, p->edges_h, 8 * m);
    } else if (m) {
        HIPCHK(hipMemcpyAsync(host_edges, p->edges, 8 * m, hipMemcpyDeviceToHost, p->stream));
        HIPCHK(hipStreamSynchronize(p->stream));
    }
    return xkrt_hip_dep_edges_consume(d);
}

int xkrt_hip_dep_stats(xkrt_hip_dep_domain_t d, xkrt_hip_dep_stats_t *out)
{
    if (!d || !out) return -EINVAL;
    xkrt_hip_pool *p = d->pool;
    (void)read_counters(p);  // the error bits are reported in out->err
    const DevCounters &c = *p->ctr_host;
    std::vector<uint4> blk(4 * BLKSLOTS);
    HIPCHK(hipMemcpy(blk.data(), p->blk, 16ull * 4 * BLKSLOTS, hipMemcpyDeviceToHost));
    uint32_t maxf = 0, maxg = 0;
    uint64_t visits = 0;
    for (const uint4 &b : blk) {
        maxf = std::max(maxf, b.x);
        maxg = std::max(maxg, b.y);
        visits += (uint64_t)b.w << 32 | b.z;
    }
    out->accesses = d->n_acc;
    out->rects = d->n_rect;
    out->edges = c.edge_total;
    out->candidates = c.cand_total;
    out->launches = p->n_launch;
    out->bins = (uint64_t)d->nbx * d->nby;
    out->max_frags = maxf;
    out->max_reads = maxg;
    out->frag_visits = visits;
    out->device_bytes = (uint64_t)d->nbx * d->nby * bin_bytes(p) + (d->private_pool ? p->bytes_scratch : 0);
    out->err = c.err | d->err;
    return 0;
}

void *xkrt_hip_dep_stream(xkrt_hip_dep_domain_t d) { return d ? (void *)d->pool->stream : nullptr; }

int xkrt_hip_dep_retire(xkrt_hip_dep_domain_t d, uint64_t horizon)
{
    if (!d) return -EINVAL;
    xkrt_hip_pool *p = d->pool;
    if (p->server) return -EBUSY;
    if (p->poisoned) return -EINVAL;
    if (d->err) return -ECANCELED;
    const uint64_t top = std::max<uint64_t>(d->next_id, d->any_id ? (uint64_t)d->last_id + 1 : 0);
    if (horizon > top) return -EINVAL;
    if (horizon == 0) return 0;
    if (p->open_nrec) return -EBUSY;  // staged records carry the current ids
    int rc = read_counters(p);        // the stream is idle after this
    if (rc) return rc;
    if (p->ctr_host->edge_total) return -EBUSY;  // edges in the buffer carry the current ids
    const uint32_t nb = d->nbx * d->nby;
    if (nb) {
        const uint32_t grid = (uint32_t)std::min<uint64_t>(((uint64_t)nb + 3) / 4, 65536);
        hipLaunchKernelGGL(k_retire, dim3(grid), dim3(256), 0, p->stream, p->bin_nf, p->bin_nrs, p->bin_one, p->frag,
                           p->rs, p->fcap, p->rcap, d->bin0, nb, (uint32_t)horizon);
        rc = hip_err(hipGetLastError());
        if (rc) return rc;
    }
    d->next_id = d->next_id > horizon ? d->next_id - horizon : 0;
    if (d->any_id) {
        if (d->last_id >= horizon) d->last_id -= (uint32_t)horizon;
        else d->any_id = false;  // every caller id retired: any id may come next
    }
    d->id_base += horizon;
    return 0;
}

int xkrt_hip_dep_id_base(xkrt_hip_dep_domain_t d, uint64_t *id_base)
{
    if (!d || !id_base) return -EINVAL;
    *id_base = d->id_base;
    return 0;
}

int xkrt_hip_dep_reset(xkrt_hip_dep_domain_t d)
{
    if (!d) return -EINVAL;
    xkrt_hip_pool *p = d->pool;
    const uint32_t nb = d->nbx * d->nby;
    DevCounters *c = d->private_pool ? p->ctr : nullptr;
    if (nb) {
        hipLaunchKernelGGL(k_reset, dim3(std::min<uint32_t>((nb + 255) / 256, 4096u)), dim3(256), 0, p->stream,
                           p->bin_nf, p->bin_nrs, p->bin_cnt, p->bin_one, d->bin0, nb, c, p->blk);
    } else if (c) {
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, p->stream, p->bin_nf, p->bin_nrs, p->bin_cnt, p->bin_one,
                           0u, 0u, c, p->blk);
    }
    HIPCHK(hipGetLastError());
    if (d->private_pool) p->poisoned = false;
    const int rc = clear_domain_error(p, d->idx);
    if (rc) return rc;
    d->err = 0;
    d->next_id = 0;
    d->max_xlen = 0;
    d->col0 = false;
    return 0;
}

int xkrt_hip_copy(void *dst, const void *src, uint64_t bytes, void *stream)
{
    if (!dst || !src || (bytes & 15)) return -EINVAL;
    const uint64_t n16 = bytes / 16;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n16 + 255) / 256, 65536);
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (uint4 *)dst, (const uint4 *)src, n16);
    return hip_err(hipGetLastError());
}

int xkrt_hip_dep_timing(xkrt_hip_dep_domain_t d, int enable)
{
    if (!d) return -EINVAL;
    return xkrt_hip_pool_timing(d->pool, enable);
}

int xkrt_hip_dep_kernel_times(xkrt_hip_dep_domain_t d, double *ms, uint64_t *launches, int clear)
{
    if (!d || !ms || !launches) return -EINVAL;
    xkrt_hip_pool *p = d->pool;
    HIPCHK(hipStreamSynchronize(p->stream));
    harvest_events(p);
    for (int k = 0; k < 6; ++k) ms[k] = p->kernel_ms[k];
    *launches = p->timed_launches;
    if (clear) {
        for (int k = 0; k < 6; ++k) p->kernel_ms[k] = 0;
        p->timed_launches = 0;
    }
    return 0;
}

int xkrt_hip_dep_launch_times(xkrt_hip_dep_domain_t d, float *ms, uint64_t cap, uint64_t *n, int clear)
{
    if (!d) return -EINVAL;
    return xkrt_hip_pool_launch_times(d->pool, ms, cap, n, clear);
}

} // extern "C"

// ---------------------------------------------------------------- serving engine
//
// One host thread drives a pool for many clients.  The API threads only
// append to an inbox under a mutex; every pool call on the pool's stream
// happens on the engine thread, which loops over: take the inbox; retire
// completed launches (what each launch's k_final wrote to pinned host memory
// gives its edge count and error bits without a stream sync) and hand them to
// the finisher thread, which copies their edges out per client and stamps
// their batches ready; stage every arrived batch of admitted clients into one
// launch while fewer than `depth` launches are in flight, admitting a client's
// domain at its first batch; arm the next launch; reset the edge buffer when
// it is half full, the pipeline is empty and every edge is out; release closed
// clients.

namespace {

struct SrvClient;

struct SrvRequest {
    SrvClient *c = nullptr;
    std::vector<xkrt_hip_access_t> acc;
    double t_submit = 0, t_ready = 0;
    std::atomic<int> status{1};  // 1 pending, 0 resolved, < 0 refused (negative errno)
    int outcome = 0;             // engine thread: the launch's verdict, published once the edges are out
};

struct SrvClient {
    uint64_t ld = 0, s = 0, ylo = 0, yhi = 0;
    uint32_t bx = 0, by = 0;
    xkrt_hip_dep_domain *dom = nullptr;
    std::atomic<int> status{0};  // < 0: a batch was refused; the client's later batches are refused too
    bool closing = false;   // its close arrived (after all of its batches)
    bool closed_api = false;  // close was called (API side, under the engine mutex)
    bool left = false;
    uint64_t round = 0;     // staging round in which it may not stage further batches
    uint64_t first_id = UINT64_MAX;  // domain access id of its first access
    uint64_t nrecv = 0;
    std::atomic<uint64_t> ndone{0};  // batches stamped (by the finisher, or refused by the engine)
    uint64_t need = 0;      // pool edge count to harvest before it may leave
    std::vector<uint32_t> edges;  // (pred, succ) in its own access ids; guarded by the server's emu
};

struct SrvFlight {
    uint64_t ticket;
    std::vector<SrvRequest *> reqs;
    int gate_hb;  // staging buffer of a gated launch, -1 otherwise
};

double srv_us(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
}

} // namespace

struct xkrt_hip_server {
    xkrt_hip_pool *p = nullptr;
    xkrt_hip_server_config_t cfg{};
    std::thread th;
    std::mutex mu;
    std::atomic<bool> stop{false};
    std::chrono::steady_clock::time_point t0;
    // shared, under mu
    std::vector<std::unique_ptr<SrvClient>> clients;
    std::vector<std::unique_ptr<SrvRequest>> reqs;
    std::vector<std::pair<SrvRequest *, SrvClient *>> inbox;  // a batch, or a close (null batch)
    std::atomic<int> err{0};
    std::atomic<uint64_t> n_final{0}, n_closed{0}, n_left{0};
    // engine thread only
    std::deque<SrvRequest *> pending;
    std::deque<SrvFlight> flights;
    std::vector<SrvClient *> live;   // received a batch or a close, not left
    SrvClient *owner[MAX_DOMS] = {};
    uint64_t round = 0;
    uint64_t edges_seen = 0;  // pool edge count after the last retired launch
    std::atomic<uint64_t> harvested{0};  // edges copied out (by the finisher)
    // The finisher thread: the retired launches, each with the pool's edge count after it, go to
    // it in order; it copies their edges into the clients' buffers and stamps their batches ready,
    // while the engine thread stages and arms the next launches.
    std::thread fin_th;
    std::mutex emu;  // the clients' edge buffers and the owner table (apart from mu: submits never wait for a copy)
    std::mutex fin_mu;
    std::condition_variable fin_cv;
    std::deque<std::pair<SrvFlight, uint64_t>> fin_q;  // under fin_mu
    bool fin_stop = false;                             // under fin_mu: finish the queue, then end
    std::atomic<uint64_t> fin_pending{0};              // launches queued or being finished
    bool hold = false;        // no launches until the pipeline drains (edge buffer reset)
    bool arm_due = false;     // a launch went out: queue the next one behind it (prearm)
    hipStream_t hs = nullptr;
    uint32_t *he = nullptr, *hd = nullptr;  // pinned harvest buffers
    uint64_t hcap = 0;
    // statistics, in the order of xkrt_hip_server_stats_t
    enum { S_LAUNCHES, S_BATCHES, S_ADMITTED, S_LEFT, S_WAITED, S_HARVESTS, S_CONSUMES, S_GATED, S_N };
    std::atomic<uint64_t> st[S_N] = {};
    void count(int k, uint64_t v = 1) { st[k].fetch_add(v, std::memory_order_relaxed); }
};

namespace {

void srv_finish(xkrt_hip_server *s, SrvRequest *r, int status, double t)
{
    r->t_ready = t;
    r->c->ndone++;
    r->status.store(status, std::memory_order_release);
    s->n_final.fetch_add(1, std::memory_order_release);
}

int srv_admit(xkrt_hip_server *s, SrvClient *c)
{
    xkrt_hip_pool *p = s->p;
    xkrt_hip_dep_domain *d = nullptr;
    int rc = add_domain(p, c->ld, c->s, c->bx, c->by, &d);
    if (rc == -ENOSPC) return -ENOMEM;  // every domain slot is taken: wait as for memory
    if (rc) return rc;
    d->ybase = c->ylo;
    d->framed = true;
    rc = ensure_grid(d, 0, (uint32_t)(c->yhi - c->ylo));
    if (rc) {
        remove_domain(d);
        return rc;
    }
    d->fixed_frame = true;
    c->dom = d;
    {
        std::lock_guard<std::mutex> g(s->emu);  // the finisher reads the table
        s->owner[d->idx] = c;
    }
    s->count(xkrt_hip_server::S_ADMITTED);
    return 0;
}

int srv_harvest(xkrt_hip_server *s, uint64_t upto);

int srv_retire(xkrt_hip_server *s, bool &work)
{
    if (s->flights.empty()) return 0;
    xkrt_hip_pool *p = s->p;
    int rc = retire(p, 0);
    if (rc) return rc;
    while (!s->flights.empty() && s->flights.front().ticket <= p->done) {
        SrvFlight &f = s->flights.front();
        // a gate that gave up before the engine opened it: the batches were not resolved
        if (f.gate_hb >= 0 && gate_result(p, (uint32_t)f.gate_hb, f.ticket) != GATE_OPEN) return -EIO;
        // the launch's k_final wrote its outcome before the launch completed
        LaunchSnap *o = &p->snap[f.ticket % NSNAP];
        if (__atomic_load_n(&o->launch, __ATOMIC_ACQUIRE) != f.ticket) return -EIO;
        const uint32_t err = __atomic_load_n(&o->err, __ATOMIC_RELAXED);
        const uint64_t total = __atomic_load_n(&o->edge_total, __ATOMIC_RELAXED);
        if (err) {
            p->poisoned = true;
            rc = err_to_errno(err);
            return rc ? rc : -EIO;
        }
        s->edges_seen = total;
        // past half the edge buffer: no new launches until it is copied out and reset
        if (s->edges_seen > p->edge_cap / 2) s->hold = true;
        // A client whose domain failed in this launch (a bin outgrew its capacity class, or a
        // record left its frame) fails alone: its batches of this launch and every later one
        // are refused with its error, and its edges are dropped; the other clients of the
        // launch are resolved.
        auto failed = [&](const SrvClient *c) -> uint32_t {
            if (!c->dom) return 0u;
            const uint32_t idx = c->dom->idx;
            return (__atomic_load_n(&o->derr[idx / 4], __ATOMIC_RELAXED) >> (8 * (idx % 4))) & 0xffu;
        };
        for (SrvRequest *r : f.reqs) {
            SrvClient *c = r->c;
            c->need = total;
            const uint32_t bits = failed(c);
            if (bits && !c->status) {
                c->dom->err |= bits;
                c->status = err_to_errno(bits);
            }
            r->outcome = bits ? c->status.load() : 0;  // published by the finisher
        }
        s->fin_pending.fetch_add(1, std::memory_order_acq_rel);
        {
            std::lock_guard<std::mutex> g(s->fin_mu);
            s->fin_q.emplace_back(std::move(f), total);
        }
        s->fin_cv.notify_one();
        s->flights.pop_front();
        work = true;
    }
    return 0;
}

// A batch is ready once its edges are in its client's host buffer.  The finisher
// thread takes the retired launches in order: their edges are copied out, then
// their batches are stamped -- while the engine thread stages and arms the next
// launch.  After an engine error it stamps what it still holds with the error.
void srv_finisher(xkrt_hip_server *s)
{
    (void)hipSetDevice(s->p->device);
    while (true) {
        // poll for a while before sleeping: a wake-up through the condition variable costs tens
        // of microseconds, which a batch would wait on the way to ready
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t k = 0; s->fin_pending.load(std::memory_order_acquire) == 0; ++k) {
            if ((k & 63u) == 63u && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) break;
            std::this_thread::yield();
        }
        std::pair<SrvFlight, uint64_t> job;
        {
            std::unique_lock<std::mutex> g(s->fin_mu);
            s->fin_cv.wait(g, [&] { return s->fin_stop || !s->fin_q.empty(); });
            if (s->fin_q.empty()) return;
            job = std::move(s->fin_q.front());
            s->fin_q.pop_front();
        }
        int rc = s->err.load(std::memory_order_acquire);
        if (!rc) {
            rc = srv_harvest(s, job.second);
            if (rc) s->err.store(rc, std::memory_order_release);
        }
        const double t = srv_us(s->t0);
        for (SrvRequest *r : job.first.reqs) srv_finish(s, r, rc ? rc : r->outcome, t);
        s->count(xkrt_hip_server::S_BATCHES, job.first.reqs.size());
        s->fin_pending.fetch_sub(1, std::memory_order_acq_rel);
    }
}

// ends the finisher once it has stamped every launch handed to it
void srv_finisher_join(xkrt_hip_server *s)
{
    {
        std::lock_guard<std::mutex> g(s->fin_mu);
        s->fin_stop = true;
    }
    s->fin_cv.notify_all();
    if (s->fin_th.joinable()) s->fin_th.join();
}

int srv_launch(xkrt_hip_server *s, bool &work)
{
    xkrt_hip_pool *p = s->p;
    if (s->hold || s->pending.empty() || s->flights.size() >= s->cfg.depth) return 0;
    const uint64_t round = ++s->round;
    const uint64_t cap = s->cfg.max_rects ? s->cfg.max_rects : p->rec_cap;
    // admissions first: they enqueue the new bins' resets, which must run before
    // the launch that uses them (so an armed gate is cancelled first)
    for (SrvRequest *r : s->pending) {
        SrvClient *c = r->c;
        if (c->dom || c->status || c->round == round) continue;
        if (p->gate_armed) {
            const int rc = gate_cancel(p);
            if (rc) return rc;
        }
        const int rc = srv_admit(s, c);
        if (rc == -ENOMEM) {
            s->count(xkrt_hip_server::S_WAITED);
            c->round = round;
            continue;
        }
        if (rc) c->status = rc;
        work = true;
    }
    // A gated launch reads its records in place and so stays a small launch.  When more is
    // queued than one small launch takes, the gate is cancelled and the queue goes out in one
    // large launch: small gated launches one after another let a backlog under load grow.
    if (p->gate_armed) {
        uint64_t est = 0, nb = 0;
        for (SrvRequest *r : s->pending) {
            if (!r->c->dom || r->c->status) continue;
            if (s->cfg.max_batches && nb++ >= s->cfg.max_batches) break;  // what this launch may take
            est += 3 * (uint64_t)r->acc.size();  // at most 3 rects per access
            if (est > p->cfg.small_launch_rects) break;
        }
        if (est > p->cfg.small_launch_rects) {
            const int rc = gate_cancel(p);
            if (rc) return rc;
        }
    }
    std::vector<SrvRequest *> batch;
    std::deque<SrvRequest *> keep;
    while (!s->pending.empty()) {
        SrvRequest *r = s->pending.front();
        SrvClient *c = r->c;
        if (c->status) {  // an earlier batch of this client was refused
            s->pending.pop_front();
            srv_finish(s, r, c->status, srv_us(s->t0));
            work = true;
            continue;
        }
        if (c->round == round || !c->dom) {  // waits for admission; keeps the client's batches in order
            c->round = round;
            keep.push_back(r);
            s->pending.pop_front();
            continue;
        }
        if ((s->cfg.max_batches && batch.size() >= s->cfg.max_batches) || p->open_nrec >= cap) break;
        // a gated launch reads its records in place: it stays a small launch (at most 3 rects per access)
        if (p->gate_armed && !batch.empty() && p->open_nrec + 3 * r->acc.size() > p->cfg.small_launch_rects) break;
        uint64_t first = 0;
        const int rc = stage(p, c->dom, r->acc.data(), r->acc.size(), &first);
        if (p->poisoned) return -EIO;
        if (rc == -ENOSPC) break;  // the launch is full
        if (rc) {
            c->status = rc;
            continue;
        }
        if (c->first_id == UINT64_MAX) c->first_id = first;
        batch.push_back(r);
        s->pending.pop_front();
    }
    while (!keep.empty()) {
        s->pending.push_front(keep.back());
        keep.pop_back();
    }
    if (batch.empty()) return 0;
    uint64_t tk = 0;
    int ghb = -1;
    if (p->gate_armed && p->open_nrec <= std::max<uint64_t>(p->cfg.small_launch_rects, 1)) {
        ghb = (int)p->gate_hb;
        gate_ring(p, &tk);  // the launch is already queued: opening the gate starts it
        s->count(xkrt_hip_server::S_GATED);
    } else {
        if (p->gate_armed) {
            const int rc = gate_cancel(p);
            if (rc) return rc;
        }
        const int rc = launch_open(p);
        if (rc) return rc;
        tk = p->n_launch;
    }
    for (SrvRequest *r : batch) std::vector<xkrt_hip_access_t>().swap(r->acc);
    s->flights.push_back({tk, std::move(batch), ghb});
    s->count(xkrt_hip_server::S_LAUNCHES);
    work = true;
    s->arm_due = s->cfg.prearm;
    return 0;
}

// Queues the next launch behind the one just started, while the device works (with two
// launches in flight, once the staging buffer it takes is free).  Arming enqueues the whole
// kernel chain, tens of microseconds of host time, while the finisher stamps the retired
// batches.
int srv_arm(xkrt_hip_server *s)
{
    if (!s->arm_due) return 0;
    s->arm_due = false;
    if (s->hold) return 0;  // the edge buffer is reset first
    return s->cfg.depth > 1 ? gate_try_arm(s->p) : gate_arm(s->p);
}

// The edges of retired launches, up to the pool's edge count `upto`, into their
// clients' host buffers (finisher thread).  Retired launches never change these
// edges again; later launches append after them.
int srv_harvest(xkrt_hip_server *s, uint64_t upto)
{
    xkrt_hip_pool *p = s->p;
    while (s->harvested.load(std::memory_order_acquire) < upto) {
        const uint64_t a = s->harvested.load(std::memory_order_relaxed);
        uint64_t n = upto - a;
        const uint32_t *ee, *ed;
        if (p->edges_h) {  // k_final wrote them to host memory before its launch completed
            ee = p->edges_h + 2 * a;
            ed = p->edom_h + a;
        } else {
            // at most hcap edges per copy: the pinned buffers are allocated once, at creation
            // (freeing pinned memory may synchronise the device)
            n = std::min(n, s->hcap);
            HIPCHK(hipMemcpyAsync(s->he, p->edges + 2 * a, 8 * n, hipMemcpyDeviceToHost, s->hs));
            HIPCHK(hipMemcpyAsync(s->hd, p->edom + a, 4 * n, hipMemcpyDeviceToHost, s->hs));
            HIPCHK(hipStreamSynchronize(s->hs));
            ee = s->he;
            ed = s->hd;
        }
        {
            std::lock_guard<std::mutex> g(s->emu);
            for (uint64_t i = 0; i < n; ++i) {
                SrvClient *c = s->owner[ed[i] & (MAX_DOMS - 1)];
                if (!c) return -EIO;
                if (c->status) continue;  // a failed client's edges are not its reference edges
                const uint32_t f = (uint32_t)c->first_id;
                c->edges.push_back(ee[2 * i] - f);
                c->edges.push_back(ee[2 * i + 1] - f);
            }
        }
        s->harvested.store(a + n, std::memory_order_release);
        s->count(xkrt_hip_server::S_HARVESTS);
    }
    return 0;
}

int srv_housekeep(xkrt_hip_server *s, bool &work, double idle_us)
{
    xkrt_hip_pool *p = s->p;
    bool leaving = false;
    for (SrvClient *c : s->live)
        if (c->closing && c->ndone == c->nrecv) leaving = true;
    // reset the half-used edge buffer once the pipeline has drained and the finisher has
    // copied every edge out and is idle
    if (s->hold && s->flights.empty() && s->fin_pending.load(std::memory_order_acquire) == 0 &&
        s->harvested.load(std::memory_order_acquire) == s->edges_seen) {
        const int rc = gate_cancel(p);  // the reset must not wait behind a gated launch
        if (rc) return rc;
        hipLaunchKernelGGL(k_consume, dim3(1), dim3(64), 0, p->stream, p->ctr);
        HIPCHK(hipGetLastError());
        s->edges_seen = 0;
        s->harvested.store(0, std::memory_order_release);
        for (SrvClient *c : s->live) c->need = 0;
        s->hold = false;
        s->count(xkrt_hip_server::S_CONSUMES);
        work = true;
    }
    // with two launches in flight the next gate is armed once a staging buffer is free
    if (s->cfg.prearm && s->cfg.depth > 1 && !p->gate_armed && !s->hold) {
        const int rc = gate_try_arm(p);
        if (rc) return rc;
    }
    // a gate armed for long without work runs its launch empty (it would give up after 1 s)
    if (p->gate_armed && s->flights.empty() && s->pending.empty() && idle_us > 200000.0) {
        const int rc = gate_cancel(p);
        if (rc) return rc;
        work = true;
    }
    if (!leaving) return 0;
    for (size_t i = 0; i < s->live.size();) {
        SrvClient *c = s->live[i];
        if (c->closing && c->ndone == c->nrecv && c->need <= s->harvested) {
            if (c->dom) {
                {
                    std::lock_guard<std::mutex> g(s->emu);  // the finisher reads the table
                    s->owner[c->dom->idx] = nullptr;
                }
                remove_domain(c->dom);  // stream order keeps later launches off its old bins
                c->dom = nullptr;
            }
            c->left = true;
            s->count(xkrt_hip_server::S_LEFT);
            s->n_left.fetch_add(1, std::memory_order_release);
            s->live[i] = s->live.back();
            s->live.pop_back();
            work = true;
            continue;
        }
        ++i;
    }
    return 0;
}

void srv_loop(xkrt_hip_server *s)
{
    (void)hipSetDevice(s->p->device);
    s->fin_th = std::thread(srv_finisher, s);
    std::vector<std::pair<SrvRequest *, SrvClient *>> in;
    double last_work = srv_us(s->t0);
    while (true) {
        bool work = false;
        {
            std::lock_guard<std::mutex> g(s->mu);
            in.swap(s->inbox);
        }
        for (auto &[r, c] : in) {
            if (c->nrecv == 0 && !c->closing) s->live.push_back(c);
            if (r) {
                s->pending.push_back(r);
                c->nrecv++;
            } else {
                c->closing = true;
            }
        }
        work = !in.empty();
        in.clear();
        int rc = srv_retire(s, work);
        if (!rc) rc = srv_launch(s, work);
        if (!rc) rc = srv_arm(s);
        const double now = srv_us(s->t0);
        if (!rc) rc = srv_housekeep(s, work, now - last_work);
        if (!rc) rc = s->err.load(std::memory_order_acquire);  // the finisher's
        if (rc) {
            (void)gate_cancel(s->p);
            s->err.store(rc, std::memory_order_release);
            // refuse everything not resolved, so that drains end (the finisher stamps what it
            // holds with the error)
            srv_finisher_join(s);
            const double t = srv_us(s->t0);
            for (SrvRequest *r : s->pending) srv_finish(s, r, rc, t);
            for (auto &f : s->flights)
                for (SrvRequest *r : f.reqs) srv_finish(s, r, rc, t);
            s->pending.clear();
            s->flights.clear();
            break;
        }
        if (work) {
            last_work = now;
            continue;
        }
        if (s->stop.load(std::memory_order_acquire) && s->flights.empty()) {
            const int rc2 = gate_cancel(s->p);
            if (rc2) s->err.store(rc2, std::memory_order_release);
            srv_finisher_join(s);
            break;
        }
        // poll: an arrival is seen within microseconds; after 2 ms without work, back off
        if (s->flights.empty() && now - last_work > 2000.0) std::this_thread::sleep_for(std::chrono::microseconds(50));
        else std::this_thread::yield();
    }
    // late submissions after a failure are refused by the API; fail what is still queued
    std::lock_guard<std::mutex> g(s->mu);
    const int e = s->err.load();
    for (auto &[r, c] : s->inbox)
        if (r) {
            r->c->ndone++;
            r->status.store(e ? e : -ECANCELED, std::memory_order_release);
            s->n_final.fetch_add(1, std::memory_order_release);
        }
    s->inbox.clear();
}

} // namespace

extern "C" {

int xkrt_hip_server_create(xkrt_hip_pool_t p, const xkrt_hip_server_config_t *cfg, xkrt_hip_server_t *out)
{
    if (!p || !out) return -EINVAL;
    *out = nullptr;
    if (p->server || p->open_nrec) return -EBUSY;
    if (p->poisoned) return -EINVAL;
    xkrt_hip_server_config_t c{};
    if (cfg) c = *cfg;
    if (c.depth == 0) c.depth = 2;
    if (c.depth > 16) return -EINVAL;
    // a gated launch waits in the stream: only with one launch in flight, over small launches
    if (c.prearm && (c.depth > 2 || p->cfg.small_launch_rects == 0)) return -EINVAL;
    int rc = read_counters(p);  // edges already in the buffer are not the engine's
    if (rc) return rc;
    if (c.reserve_bytes) {
        uint64_t bins = c.reserve_bytes / bin_bytes(p);
        if (p->cfg.budget_bytes)
            bins = std::min<uint64_t>(bins, (p->cfg.budget_bytes - std::min(p->cfg.budget_bytes, p->bytes_scratch)) /
                                                bin_bytes(p));
        bins = std::min<uint64_t>(bins, 0xFFFFFFull);
        if (bins > p->cap_bins && (rc = grow_bins(p, (uint32_t)bins))) return rc;
    }
    auto *s = new xkrt_hip_server();
    s->p = p;
    s->cfg = c;
    s->edges_seen = s->harvested = p->ctr_host->edge_total;
    s->hcap = std::min<uint64_t>(p->edge_cap, 1ull << 20);
    if (hipStreamCreateWithFlags(&s->hs, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc(&s->he, 8 * s->hcap, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&s->hd, 4 * s->hcap, hipHostMallocDefault) != hipSuccess) {
        if (s->he) (void)hipHostFree(s->he);
        if (s->hd) (void)hipHostFree(s->hd);
        delete s;
        return -ENOMEM;
    }
    p->server = s;
    s->t0 = std::chrono::steady_clock::now();
    s->th = std::thread(srv_loop, s);
    *out = s;
    return 0;
}

int xkrt_hip_server_destroy(xkrt_hip_server_t s)
{
    if (!s) return -EINVAL;
    s->stop.store(true, std::memory_order_release);
    if (s->th.joinable()) s->th.join();
    xkrt_hip_pool *p = s->p;
    (void)hipStreamSynchronize(p->stream);
    for (auto &c : s->clients)
        if (c->dom) {
            remove_domain(c->dom);
            c->dom = nullptr;
        }
    if (s->he) (void)hipHostFree(s->he);
    if (s->hd) (void)hipHostFree(s->hd);
    if (s->hs) (void)hipStreamDestroy(s->hs);
    p->server = nullptr;
    const int rc = s->err.load();
    delete s;
    return rc;
}

int xkrt_hip_server_client(xkrt_hip_server_t s, uint64_t ld, uint64_t sizeof_type, uint32_t bin_x_bytes,
                           uint32_t bin_y_cols, uint64_t y_lo, uint64_t y_hi, uint32_t *client)
{
    if (!s || !client || ld == 0 || sizeof_type == 0 || y_hi <= y_lo || y_hi - y_lo >= (1ull << 32)) return -EINVAL;
    bin_x_bytes = bin_x_bytes ? bin_x_bytes : 1024u;
    bin_y_cols = bin_y_cols ? bin_y_cols : 128u;
    if (!is_pow2(bin_x_bytes) || !is_pow2(bin_y_cols) || ld * sizeof_type > 0xFFFFFFFFull) return -EINVAL;
    auto c = std::make_unique<SrvClient>();
    c->ld = ld; c->s = sizeof_type; c->ylo = y_lo; c->yhi = y_hi;
    c->bx = bin_x_bytes; c->by = bin_y_cols;
    std::lock_guard<std::mutex> g(s->mu);
    *client = (uint32_t)s->clients.size();
    s->clients.push_back(std::move(c));
    return 0;
}

int xkrt_hip_server_submit(xkrt_hip_server_t s, uint32_t client, const xkrt_hip_access_t *acc, uint64_t n,
                           uint64_t *request)
{
    if (!s || (!acc && n) || n == 0) return -EINVAL;
    const double t = srv_us(s->t0);
    if (const int e = s->err.load(std::memory_order_acquire)) return e;
    auto r = std::make_unique<SrvRequest>();
    r->acc.assign(acc, acc + n);
    r->t_submit = t;
    std::lock_guard<std::mutex> g(s->mu);
    // checked under the mutex: the engine sets the error before it refuses the inbox under it
    if (const int e = s->err.load(std::memory_order_acquire)) return e;
    if (client >= s->clients.size()) return -EINVAL;
    SrvClient *c = s->clients[client].get();
    if (c->closed_api) return -EINVAL;
    r->c = c;
    if (request) *request = s->reqs.size();
    s->inbox.push_back({r.get(), c});
    s->reqs.push_back(std::move(r));
    return 0;
}

int xkrt_hip_server_close(xkrt_hip_server_t s, uint32_t client)
{
    if (!s) return -EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    if (client >= s->clients.size() || s->clients[client]->closed_api) return -EINVAL;
    s->clients[client]->closed_api = true;
    s->inbox.push_back({nullptr, s->clients[client].get()});
    s->n_closed.fetch_add(1, std::memory_order_release);
    return 0;
}

int xkrt_hip_server_drain(xkrt_hip_server_t s, uint64_t timeout_us)
{
    if (!s) return -EINVAL;
    const auto t0 = std::chrono::steady_clock::now();
    while (true) {
        uint64_t nreq;
        {
            std::lock_guard<std::mutex> g(s->mu);
            nreq = s->reqs.size();
        }
        const int e = s->err.load(std::memory_order_acquire);
        if (s->n_final.load(std::memory_order_acquire) == nreq &&
            (e || s->n_left.load(std::memory_order_acquire) == s->n_closed.load(std::memory_order_acquire)))
            return e;
        if (timeout_us && srv_us(t0) > (double)timeout_us) return -ETIMEDOUT;
        std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

int xkrt_hip_server_times(xkrt_hip_server_t s, uint64_t first, uint64_t n, double *submit_us, double *ready_us,
                          int32_t *status)
{
    if (!s) return -EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    if (first > s->reqs.size() || n > s->reqs.size() - first) return -ERANGE;
    for (uint64_t i = 0; i < n; ++i) {
        const SrvRequest *r = s->reqs[first + i].get();
        const int st = r->status.load(std::memory_order_acquire);
        if (submit_us) submit_us[i] = r->t_submit;
        if (ready_us) ready_us[i] = st == 1 ? 0.0 : r->t_ready;
        if (status) status[i] = st;
    }
    return 0;
}

int xkrt_hip_server_edges(xkrt_hip_server_t s, uint32_t client, uint32_t *edges, uint64_t cap, uint64_t *n)
{
    if (!s || !n || (cap && !edges)) return -EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    if (client >= s->clients.size()) return -EINVAL;
    std::lock_guard<std::mutex> ge(s->emu);  // order: mu, then emu
    const auto &v = s->clients[client]->edges;
    *n = v.size() / 2;
    const uint64_t m = std::min<uint64_t>(*n, cap);
    if (m) memcpy(edges, v.data(), 8 * m);
    return 0;
}

int xkrt_hip_server_stats(xkrt_hip_server_t s, xkrt_hip_server_stats_t *out)
{
    if (!s || !out) return -EINVAL;
    uint64_t *f = &out->launches;
    for (int k = 0; k < xkrt_hip_server::S_N; ++k) f[k] = s->st[k].load(std::memory_order_relaxed);
    out->error = s->err.load();
    return 0;
}

} // extern "C"
