// MI355X (gfx950, wave64) coherency planner for one (ld, sizeof_type) memory
// tree with a host and one GPU: for every access, the valid-region split
// (the tree's fragments inside the access, with their coherency bits before
// it) and the fetch list, with the reference's semantics
// (include/xkrt/memory/access/blas/memory-tree.hpp:1067-1215, 1293-1874,
// 955-1061; synchronous completion: the fetches of an access complete before
// the next access; no eviction).
//
// The tree is restated per point:
//   * geometry: the fragments are the global dim-0 breakpoints (every x
//     endpoint inserted so far) times, per slab, the dim-1 breakpoints of the
//     rects covering that slab.  TX[u] and TY[u][y] hold the first access
//     that introduced breakpoint u (x unit = sizeof_type bytes) and y (at
//     x unit u), so access a sees exactly the breakpoints with a stamp <= a,
//     and the fragments of every access of a batch are enumerated in
//     parallel: one work item per (slab, block of rows), so a large access
//     spreads over many waves (c_slabs, c_icount, c_irows);
//   * allocation views: a GPU access a reuses the chunk of the earliest
//     earlier GPU access whose rect contains a's rect (that access allocated
//     it and appended it to every point of its rect; the reference takes the
//     first chunk of the first partite's view list present in every partite,
//     1376-1425); otherwise a allocates the next chunk (c_contain, c_chunks);
//   * state: per point, coherency bits (1 host, 2 GPU), the view list (chunk
//     ids in allocation order, at most 8: the reference aborts on a ninth,
//     1362-1365), the per-view coherency bits and, for fetches in flight, the
//     device and per-view fetching bits.  c_state replays a tile's accesses in
//     order on 32 x 32 points, one workgroup per tile, and fills the payload
//     of every fragment whose first point lies in the tile;
//   * fetches in flight (async_completion): a read of a view being fetched
//     awaits it and a read while another view is being fetched registers a
//     forward (1516-1527, 1567-1600); c_complete applies completions point by
//     point with atomics (their effects commute), and the host fires the
//     registrations of each completed fetch.
//
// HBM layout per tree: TX (nxu + 1 u32), TY (nxu x (ny + 1) u32, y fastest),
// point state tile-major (u32 word + 8 u32 views per point), the GPU access
// history (rect, bytes, earliest container, chunk), chunk sizes.  Per plan
// call: access arrays (SoA), fragment rows (SoA x0, x1, y0, y1, payload,
// source chunk, access), fetch rows (SoA), tile incidence lists.

#include <hip/hip_runtime.h>

#include <errno.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <mutex>
#include <vector>

#include "../../include/xkrt_hip.h"

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint32_t VMAX = 8;                 // memory-tree.hpp:86 allocation views per replica
constexpr uint32_t TW = 32, TH = 32;         // state tile: x units by columns
constexpr uint32_t TP = TW * TH;
constexpr uint32_t TTHREADS = 256;
constexpr uint32_t PPT = TP / TTHREADS;      // points per thread
constexpr uint32_t TLIST = 4096;             // accesses of one tile per plan call (LDS list)
constexpr uint32_t ITEM_ROWS = 1024;          // rows of a fragment work item (at least)
constexpr uint32_t SCAN_BLOCK = 1024, SCAN_PER = 4, SCAN_CHUNK = SCAN_BLOCK * SCAN_PER;

enum : uint32_t { F_GPU = 1u, F_INS = 2u, F_R = 4u, F_W = 8u };
enum : uint32_t { CERR_TILE = 1u };

struct CCounters {
    unsigned int nchunks;   // chunks allocated so far
    unsigned int fail;      // lowest access id that would add a ninth view (NONE = none)
    unsigned int err;       // CERR_* bits
    unsigned int nact;      // active tiles this call
    unsigned int rows;      // fragment rows this call
    unsigned int fetches;   // fetch rows this call
    unsigned int viewerr;   // lowest access id whose chunk was missing from a point (NONE = none)
    unsigned int regs;      // registration rows this call (fetches in flight)
    unsigned int items;     // fragment work items this call (slab x row block)
};

constexpr uint32_t F_ALLOC = 1u << 31;  // descriptor flag: the access allocates its chunk

// an access as the state replay reads it: rect, flags, chunk, index, unit offset
struct __align__(16) TDesc {
    uint32_t x0, x1, y0, y1, fl, chunk, i, xo;
};
static_assert(sizeof(TDesc) == 32, "descriptors are two 16-byte words");

struct CArgs {
    // batch (index i in [0, n))
    uint32_t n, gid0;
    uint32_t async;          // 1: fetches stay in flight (fetching bits) until c_complete
    uint32_t *rreg;          // per fragment row: 1 when the access awaits or registers a forward there
    uint32_t *rrpos;         // exclusive scan of rreg
    uint32_t *reg;           // registration rows, 8 u32 each (x0, x1, y0, y1, kind, chunk, dst chunk, access)
    const uint32_t *bx0, *bx1, *by0, *by1, *bflags, *bhist, *bxoff;
    uint32_t *rowbase;
    // tree
    uint32_t nxu, ny, ny1, s;
    uint32_t *TX, *TY;
    uint32_t ntx, nty;
    uint32_t *pst, *pviews;
    // GPU history
    const uint32_t *hx0, *hx1, *hy0, *hy1;
    const unsigned long long *hbytes;
    uint32_t *hbeta, *hchunk;
    unsigned long long *chunk_size;
    // per unit of the batch's accesses (xtot, slab starts only where noted):
    // fragment rows of the slab (start, count), its items (all units), their
    // offset, the slab's end unit and access
    uint32_t *xcnt, *xrow, *xitems, *xioff, *xend, *xacc;
    uint32_t xtot, irows;    // units; rows per item
    // work items: slab start unit, fragment count, first row
    uint32_t *imap, *icnt, *irow;
    uint32_t nitems, nrows;
    // fragment rows
    uint32_t *rx0, *rx1, *ry0, *ry1, *rpay, *rsrc, *racc;
    uint32_t *rflag, *rfpos;
    // fetch rows
    uint32_t *fx0, *fx1, *fy0, *fy1, *fsd, *fsc, *fdd, *fdc;
    uint32_t *foff;
    // tiles
    uint32_t *tcnt, *tcur, *tstart, *tinc, *tact;
    TDesc *tdesc;            // per access of the batch (inserted ones)
    CCounters *ctr;
};

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }
__device__ __forceinline__ uint32_t wave_bcast(uint32_t v, int src) { return __shfl(v, src, 64); }

// ---------------------------------------------------------------- views

// earliest earlier GPU access of the tree whose rect contains this one's
__global__ void c_contain(CArgs a, uint32_t h0, uint32_t nh)
{
    const uint32_t lane = lane_id();
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t k = wid; k < nh; k += nw) {
        const uint32_t h = h0 + k;
        const uint32_t x0 = a.hx0[h], x1 = a.hx1[h], y0 = a.hy0[h], y1 = a.hy1[h];
        uint32_t found = NONE;
        for (uint32_t c = 0; c < h && found == NONE; c += 64) {
            const uint32_t j = c + lane;
            const bool in = j < h && a.hx0[j] <= x0 && x1 <= a.hx1[j] && a.hy0[j] <= y0 && y1 <= a.hy1[j];
            const uint64_t m = __ballot(in);
            if (m) found = c + (uint32_t)__builtin_ctzll(m);
        }
        if (lane == 0) a.hbeta[h] = found;
    }
}

// chunk ids: allocating accesses in order get the next ids; the others take
// their earliest container's chunk (that container allocated)
__global__ void __launch_bounds__(1024) c_chunks(CArgs a, uint32_t h0, uint32_t nh)
{
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t base;
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    if (threadIdx.x == 0) base = a.ctr->nchunks;
    __syncthreads();
    for (uint32_t c = 0; c < nh; c += 1024) {
        const uint32_t k = c + threadIdx.x;
        const bool al = k < nh && a.hbeta[h0 + k] == NONE;
        const uint64_t m = __ballot(al);
        if (lane == 0) wsum[w] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t off = 0, tot = 0;
        for (uint32_t q = 0; q < 16; ++q) {
            off += q < w ? wsum[q] : 0u;
            tot += wsum[q];
        }
        if (al) {
            const uint32_t ch = base + off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            a.hchunk[h0 + k] = ch;
            a.chunk_size[ch] = a.hbytes[h0 + k];
        }
        __syncthreads();
        if (threadIdx.x == 0) base += tot;
        __syncthreads();
    }
    __threadfence();
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nh; k += 1024) {
        const uint32_t b = a.hbeta[h0 + k];
        if (b != NONE) a.hchunk[h0 + k] = a.hchunk[b];
    }
    if (threadIdx.x == 0) a.ctr->nchunks = base;
}

// ---------------------------------------------------------------- geometry

__global__ void c_breaks(CArgs a)
{
    const uint32_t lane = lane_id();
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t i = wid; i < a.n; i += nw) {
        if (!(a.bflags[i] & F_INS)) continue;
        const uint32_t aid = a.gid0 + i;
        const uint32_t x0 = a.bx0[i], x1 = a.bx1[i], y0 = a.by0[i], y1 = a.by1[i];
        if (lane == 0) {
            atomicMin(&a.TX[x0], aid);
            atomicMin(&a.TX[x1], aid);
        }
        for (uint32_t u = x0 + lane; u < x1; u += 64) {
            atomicMin(&a.TY[(size_t)u * a.ny1 + y0], aid);
            atomicMin(&a.TY[(size_t)u * a.ny1 + y1], aid);
        }
    }
}

// slab starts of each access (the x units where a column of its fragments
// begins: x0 and every dim-0 breakpoint seen by the access) with their ends;
// per unit of the access, the row blocks of a.irows rows to enumerate there
// (0 off a slab start)
__global__ void c_slabs(CArgs a)
{
    const uint32_t lane = lane_id();
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t i = wid; i < a.n; i += nw) {
        if (!(a.bflags[i] & F_INS)) continue;
        const uint32_t aid = a.gid0 + i;
        const uint32_t x0 = a.bx0[i], x1 = a.bx1[i], xo = a.bxoff[i];
        const uint32_t nyb = (a.by1[i] - a.by0[i] + a.irows - 1) / a.irows;
        uint32_t pend = NONE;  // the last slab start seen, its end not known yet
        for (uint32_t uc = x0; uc < x1; uc += 64) {
            const uint32_t u = uc + lane;
            const bool st = u < x1 && (u == x0 || a.TX[u] <= aid);
            const uint64_t m = __ballot(st);
            if (u < x1) a.xitems[xo + (u - x0)] = st ? nyb : 0u;
            if (!m) continue;
            if (pend != NONE && lane == 0) a.xend[xo + (pend - x0)] = uc + (uint32_t)__builtin_ctzll(m);
            const uint64_t above = lane == 63 ? 0ull : (m & (~0ull << (lane + 1)));
            if (st) {
                a.xacc[xo + (u - x0)] = i;
                if (above) a.xend[xo + (u - x0)] = uc + (uint32_t)__builtin_ctzll(above);
            }
            pend = uc + 63u - (uint32_t)__builtin_clzll(m);
        }
        if (pend != NONE && lane == 0) a.xend[xo + (pend - x0)] = x1;
    }
}

// work items: item k is row block k - xioff[g] of the slab starting at unit g
__global__ void c_imap(CArgs a)
{
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < a.xtot; g += gridDim.x * blockDim.x) {
        const uint32_t c = a.xitems[g];
        if (!c) continue;
        const uint32_t b = a.xioff[g];
        for (uint32_t j = 0; j < c; ++j) a.imap[b + j] = g;
    }
}

struct Item {
    uint32_t g, i, aid, us, y0, ya, yb;
};

__device__ __forceinline__ Item item_of(const CArgs &a, uint32_t k)
{
    Item t;
    t.g = a.imap[k];
    const uint32_t j = k - a.xioff[t.g];
    t.i = a.xacc[t.g];
    t.aid = a.gid0 + t.i;
    t.us = a.bx0[t.i] + (t.g - a.bxoff[t.i]);
    t.y0 = a.by0[t.i];
    t.ya = t.y0 + j * a.irows;
    t.yb = min(a.by1[t.i], t.ya + a.irows);
    return t;
}

// fragments of each item: the rows y of its block where a fragment starts
// (y0, or a dim-1 breakpoint of the slab seen by the access)
__global__ void c_icount(CArgs a)
{
    const uint32_t lane = lane_id();
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t k = wid; k < a.nitems; k += nw) {
        const Item t = item_of(a, k);
        const uint32_t *row = a.TY + (size_t)t.us * a.ny1;
        uint32_t np = 0;
        // four 64-row chunks per step: their loads are in flight together
        for (uint32_t yc = t.ya; yc < t.yb; yc += 256) {
            uint32_t v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t y = yc + 64 * q + lane;
                v[q] = (y < t.yb && y != t.y0) ? row[y] : 0u;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t y = yc + 64 * q + lane;
                np += (uint32_t)__popcll(__ballot(y < t.yb && (y == t.y0 || v[q] <= t.aid)));
            }
        }
        if (lane == 0) a.icnt[k] = np;
    }
}

// fragment rows of each item from row irow[k] (their dim-1 ends: c_yfix); a
// slab's first item records the slab's rows; the row base of every access
__global__ void c_irows(CArgs a)
{
    const uint32_t lane = lane_id();
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t k = wid; k < a.nitems; k += nw) {
        const Item t = item_of(a, k);
        const uint32_t ue = a.xend[t.g];
        const uint32_t *row = a.TY + (size_t)t.us * a.ny1;
        uint32_t r = a.irow[k];
        if (t.ya == t.y0 && lane == 0) {
            const uint32_t e = k + a.xitems[t.g];
            a.xrow[t.g] = r;
            a.xcnt[t.g] = (e < a.nitems ? a.irow[e] : a.nrows) - r;
        }
        for (uint32_t yc = t.ya; yc < t.yb; yc += 256) {
            uint32_t v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t y = yc + 64 * q + lane;
                v[q] = (y < t.yb && y != t.y0) ? row[y] : 0u;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t y = yc + 64 * q + lane;
                const bool ps = y < t.yb && (y == t.y0 || v[q] <= t.aid);
                const uint64_t m = __ballot(ps);
                if (ps) {
                    const uint32_t idx = r + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                    a.rx0[idx] = t.us * a.s;
                    a.rx1[idx] = ue * a.s;
                    a.ry0[idx] = y;
                    a.racc[idx] = t.i;
                }
                r += (uint32_t)__popcll(m);
            }
        }
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
        const uint32_t xo = a.bxoff[i];
        const uint32_t it = xo < a.xtot ? a.xioff[xo] : a.nitems;  // items of the accesses before
        a.rowbase[i] = it < a.nitems ? a.irow[it] : a.nrows;
    }
}

// dim-1 end of each row: the next row's start in the same slab, else the
// access's y1 (the rows of a slab are consecutive, by ascending y)
__global__ void c_yfix(CArgs a)
{
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < a.nrows; r += gridDim.x * blockDim.x) {
        const uint32_t i = a.racc[r];
        const bool same = r + 1 < a.nrows && a.racc[r + 1] == i && a.rx0[r + 1] == a.rx0[r];
        a.ry1[r] = same ? a.ry0[r + 1] : a.by1[i];
    }
}

// ---------------------------------------------------------------- tiles

__device__ __forceinline__ void tile_range(const CArgs &a, uint32_t i, uint32_t &tx0, uint32_t &tx1, uint32_t &ty0,
                                           uint32_t &ty1)
{
    tx0 = a.bx0[i] / TW; tx1 = (a.bx1[i] - 1) / TW;
    ty0 = a.by0[i] / TH; ty1 = (a.by1[i] - 1) / TH;
}

__global__ void c_tcount(CArgs a)
{
    const uint32_t lane = lane_id();
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t i = wid; i < a.n; i += nw) {
        if (!(a.bflags[i] & F_INS)) continue;
        uint32_t tx0, tx1, ty0, ty1;
        tile_range(a, i, tx0, tx1, ty0, ty1);
        if (lane == 0) {
            TDesc d;
            d.x0 = a.bx0[i]; d.x1 = a.bx1[i]; d.y0 = a.by0[i]; d.y1 = a.by1[i];
            d.fl = a.bflags[i];
            d.chunk = NONE;
            if (d.fl & F_GPU) {
                const uint32_t hh = a.bhist[i];
                d.chunk = a.hchunk[hh];
                if (a.hbeta[hh] == NONE) d.fl |= F_ALLOC;
            }
            d.i = i;
            d.xo = a.bxoff[i];
            a.tdesc[i] = d;
        }
        const uint32_t w = tx1 - tx0 + 1, nt = w * (ty1 - ty0 + 1);
        for (uint32_t k = lane; k < nt; k += 64) atomicAdd(&a.tcnt[(ty0 + k / w) * a.ntx + tx0 + k % w], 1u);
    }
}

__global__ void c_tfill(CArgs a)
{
    const uint32_t lane = lane_id();
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t i = wid; i < a.n; i += nw) {
        if (!(a.bflags[i] & F_INS)) continue;
        uint32_t tx0, tx1, ty0, ty1;
        tile_range(a, i, tx0, tx1, ty0, ty1);
        const uint32_t w = tx1 - tx0 + 1, nt = w * (ty1 - ty0 + 1);
        for (uint32_t k = lane; k < nt; k += 64) {
            const uint32_t t = (ty0 + k / w) * a.ntx + tx0 + k % w;
            a.tinc[a.tstart[t] + atomicAdd(&a.tcur[t], 1u)] = i;
        }
    }
}

__global__ void c_tact(CArgs a, uint32_t ntiles)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = t < ntiles && a.tcnt[t] > 0;
    const uint64_t m = __ballot(on);
    if (!m) return;
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(&a.ctr->nact, (uint32_t)__popcll(m));
    base = wave_bcast(base, 0);
    if (on) a.tact[base + (uint32_t)__popcll(m & ((1ull << lane_id()) - 1ull))] = t;
}

// ---------------------------------------------------------------- state

// Per-point state word: bits 0-1 coherency (1 host, 2 GPU), 2-9 coherent
// views, 10-13 number of views, 14-15 fetching devices, 16-23 fetching views.
constexpr uint32_t ST_DF = 14, ST_VF = 16;
constexpr uint32_t DMAX = 256;  // access descriptors of a tile held in LDS at a time
constexpr uint32_t LCAP = 512;  // fragment-row lookups of a tile deferred to its end

__device__ void load_descs(const CArgs &a, const uint32_t *L, TDesc *D, uint32_t q0, uint32_t m)
{
    for (uint32_t q = threadIdx.x; q < m; q += TTHREADS) D[q] = a.tdesc[L[q0 + q]];
}

// the fragment row starting at (u, y) of the slab at unit xo: its payload
__device__ void fill_row(const CArgs &a, uint32_t xo, uint32_t y, uint32_t pay, uint32_t schunk)
{
    // the slab's last row starting at or before y; seven probes in flight per step
    uint32_t lo = a.xrow[xo], hi = lo + a.xcnt[xo];
    while (hi - lo > 1) {
        const uint32_t step = (hi - lo + 7) / 8;
        uint32_t v[7];
#pragma unroll
        for (uint32_t j = 0; j < 7; ++j) {
            const uint32_t idx = lo + step * (j + 1);
            v[j] = idx < hi ? a.ry0[idx] : NONE;
        }
        uint32_t nlo = lo;
#pragma unroll
        for (uint32_t j = 0; j < 7; ++j)
            if (lo + step * (j + 1) < hi && v[j] <= y) nlo = lo + step * (j + 1);
        hi = min(hi, nlo + step);
        lo = nlo;
    }
    a.rpay[lo] = pay;
    a.rsrc[lo] = schunk;
    a.rflag[lo] = (pay >> 8) & 1u;
    if (a.async) a.rreg[lo] = ((pay >> 12) & 3u) ? 1u : 0u;
}

// Replay the tile's accesses in order on its 32 x 32 points (4 per thread,
// views in registers through constant-index selects).  Only the points some
// access of the tile covers are read and written (their state, the views
// they hold, the breakpoint stamps at the point); the fragment rows whose
// first point is in the tile are looked up once the replay is done.
__global__ void __launch_bounds__(TTHREADS) __attribute__((amdgpu_waves_per_eu(4))) c_state(CArgs a)
{
    __shared__ uint32_t L[TLIST];
    __shared__ TDesc D[DMAX];
    __shared__ uint4 LK[LCAP];  // deferred lookups: slab unit, y, payload, source chunk
    __shared__ uint32_t nlk;
    const uint32_t nact = a.ctr->nact;
    for (uint32_t b = blockIdx.x; b < nact; b += gridDim.x) {
        const uint32_t tile = a.tact[b];
        const uint32_t cnt = a.tcnt[tile], start = a.tstart[tile];
        __syncthreads();
        if (cnt > TLIST) {
            if (threadIdx.x == 0) atomicOr(&a.ctr->err, CERR_TILE);
            continue;
        }
        // the points' state words and breakpoint stamps: loads in flight while the list is sorted
        const uint32_t tx0 = (tile % a.ntx) * TW, ty0 = (tile / a.ntx) * TH;
        uint32_t st[PPT], vw[PPT][VMAX], px[PPT], py[PPT], tx[PPT], ty[PPT], nv0[PPT];
        bool cov[PPT];
#pragma unroll
        for (uint32_t k = 0; k < PPT; ++k) {
            const uint32_t p = threadIdx.x + TTHREADS * k;
            px[k] = tx0 + p % TW;
            py[k] = ty0 + p / TW;
            cov[k] = px[k] < a.nxu && py[k] < a.ny;
            st[k] = 0u; tx[k] = NONE; ty[k] = NONE;
            if (cov[k]) {
                st[k] = a.pst[(size_t)tile * TP + p];
                tx[k] = a.TX[px[k]];
                ty[k] = a.TY[(size_t)px[k] * a.ny1 + py[k]];
            }
        }
        if (threadIdx.x == 0) nlk = 0;
        const uint32_t nch = (cnt + DMAX - 1) / DMAX;
        if (cnt <= TTHREADS) {
            // rank sort: access q's descriptor goes to its rank among the tile's access indices
            const uint32_t key = threadIdx.x < cnt ? a.tinc[start + threadIdx.x] : NONE;
            uint4 d0 = make_uint4(0, 0, 0, 0), d1 = d0;
            if (threadIdx.x < cnt) {
                const uint4 *g = reinterpret_cast<const uint4 *>(a.tdesc + key);
                d0 = g[0];
                d1 = g[1];
                L[threadIdx.x] = key;
            }
            __syncthreads();
            if (threadIdx.x < cnt) {
                uint32_t r = 0;
                for (uint32_t q = 0; q < cnt; ++q) r += L[q] < key ? 1u : 0u;
                uint4 *o = reinterpret_cast<uint4 *>(D + r);
                o[0] = d0;
                o[1] = d1;
            }
            __syncthreads();
        } else {
            uint32_t n2 = 1;
            while (n2 < cnt) n2 <<= 1;
            for (uint32_t q = threadIdx.x; q < n2; q += TTHREADS) L[q] = q < cnt ? a.tinc[start + q] : NONE;
            __syncthreads();
            for (uint32_t k = 2; k <= n2; k <<= 1)
                for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                    for (uint32_t q = threadIdx.x; q < n2; q += TTHREADS) {
                        const uint32_t p = q ^ j;
                        if (p > q) {
                            const uint32_t x = L[q], y = L[p];
                            if ((x > y) == ((q & k) == 0)) { L[q] = y; L[p] = x; }
                        }
                    }
                    __syncthreads();
                }
            load_descs(a, L, D, 0, DMAX);
            __syncthreads();
        }
#pragma unroll
        for (uint32_t k = 0; k < PPT; ++k) {
            if (nch == 1 && cov[k]) {  // covered by an access of the tile (all valid points past one chunk)
                bool in = false;
                for (uint32_t q = 0; q < cnt && !in; ++q)
                    in = D[q].x0 <= px[k] && px[k] < D[q].x1 && D[q].y0 <= py[k] && py[k] < D[q].y1;
                cov[k] = in;
            }
            const uint32_t p = threadIdx.x + TTHREADS * k;
            nv0[k] = cov[k] ? (st[k] >> 10) & 0xfu : 0u;
#pragma unroll
            for (uint32_t v = 0; v < VMAX; ++v)
                vw[k][v] = v < nv0[k] ? a.pviews[((size_t)tile * VMAX + v) * TP + p] : NONE;
        }

        for (uint32_t c = 0; c < nch; ++c) {
            if (c > 0) {
                __syncthreads();
                load_descs(a, L, D, c * DMAX, min(cnt - c * DMAX, DMAX));
                __syncthreads();
            }
            const uint32_t m = min(cnt - c * DMAX, DMAX);
            for (uint32_t q = 0; q < m; ++q) {
                const TDesc d = D[q];
                const uint32_t aid = a.gid0 + d.i, fl = d.fl, chunk = d.chunk;
                const bool alloc = (fl & F_ALLOC) != 0u;
#pragma unroll
                for (uint32_t k = 0; k < PPT; ++k) {
                    if (!cov[k] || px[k] < d.x0 || px[k] >= d.x1 || py[k] < d.y0 || py[k] >= d.y1) continue;
                    uint32_t coh = st[k] & 3u, vc = (st[k] >> 2) & 0xffu, nv = (st[k] >> 10) & 0xfu;
                    uint32_t df = (st[k] >> ST_DF) & 3u, vf = (st[k] >> ST_VF) & 0xffu;
                    const uint32_t before = coh | (df << 2);
                    bool need = false;
                    uint32_t act = 0;  // 2: awaits a fetch in flight, 3: registers a forward
                    uint32_t sdev = 0, schunk = NONE;
                    if (fl & F_GPU) {
                        uint32_t v = NONE;
                        if (alloc) {
                            if (nv >= VMAX) {
                                atomicMin(&a.ctr->fail, aid);
                                continue;
                            }
#pragma unroll
                            for (uint32_t t = 0; t < VMAX; ++t)
                                if (t == nv) vw[k][t] = chunk;
                            v = nv++;
                        } else {
#pragma unroll
                            for (uint32_t t = 0; t < VMAX; ++t)
                                if (t < nv && vw[k][t] == chunk && v == NONE) v = t;
                            if (v == NONE) {  // only after an abort has left this point's views short
                                atomicMin(&a.ctr->viewerr, aid);
                                continue;
                            }
                        }
                        const uint32_t bit = 1u << v;
                        if ((fl & F_R) && !(vc & bit)) {
                            // fetch_access_setup_copies (1471-1631)
                            if (vf & bit) {
                                act = 2;  // this view is being fetched: await it (1516-1527)
                                schunk = chunk;
                            } else {
                                if (coh & 2u) {  // from the first coherent view of the GPU
                                    need = true;
                                    const uint32_t sv = (uint32_t)__builtin_ctz(vc);
                                    sdev = 1;
#pragma unroll
                                    for (uint32_t t = 0; t < VMAX; ++t)
                                        if (t == sv) schunk = vw[k][t];
                                } else if (df & 2u) {  // forward from the first view being fetched (1567-1600)
                                    act = 3;
                                    const uint32_t sv = (uint32_t)__builtin_ctz(vf);
#pragma unroll
                                    for (uint32_t t = 0; t < VMAX; ++t)
                                        if (t == sv) schunk = vw[k][t];
                                } else {  // from host memory
                                    need = true;
                                }
                                vf |= bit;
                                df |= 2u;
                            }
                        }
                        if (fl & F_W) {  // fetch_access_set_coherent (1633-1667)
                            coh = 2u;
                            vc = bit;
                        }
                        if (need && !a.async) {  // synchronous completion (2067-2127)
                            vc |= bit;
                            coh |= 2u;
                            vf &= ~bit;
                            if (!vf) df &= ~2u;
                        }
                    } else if (coh != 0u && !(coh & 1u) && !(df & 1u)) {
                        // fetch_list_to_host_setup_partition (1110-1181): from the GPU view of the largest chunk
                        need = true;
                        sdev = 1;
                        unsigned long long best = 0;
#pragma unroll
                        for (uint32_t t = 0; t < VMAX; ++t) {
                            if (t < nv) {
                                const unsigned long long sz = a.chunk_size[vw[k][t]];
                                if (t == 0 || best < sz) {
                                    best = sz;
                                    schunk = vw[k][t];
                                }
                            }
                        }
                        if (a.async) df |= 1u;  // fetching to the host (1126-1130)
                        else coh |= 1u;          // synchronous completion
                    }
                    st[k] = coh | (vc << 2) | (nv << 10) | (df << ST_DF) | (vf << ST_VF);
                    // the first point of one of the access's fragments: its row gets the payload
                    const uint32_t u = px[k], y = py[k];
                    if ((u == d.x0 || tx[k] <= aid) && (y == d.y0 || ty[k] <= aid)) {
                        const uint32_t pay = before | ((need ? 1u : 0u) << 8) | (sdev << 9) | (act << 12);
                        const uint32_t xo = d.xo + (u - d.x0);
                        const uint32_t slot = atomicAdd(&nlk, 1u);
                        if (slot < LCAP) LK[slot] = make_uint4(xo, y, pay, schunk);
                        else fill_row(a, xo, y, pay, schunk);
                    }
                }
            }
        }
        __syncthreads();
        const uint32_t nl = min(nlk, LCAP);
        for (uint32_t e = threadIdx.x; e < nl; e += TTHREADS) {
            const uint4 l = LK[e];
            fill_row(a, l.x, l.y, l.z, l.w);
        }
#pragma unroll
        for (uint32_t k = 0; k < PPT; ++k) {
            if (!cov[k]) continue;
            const uint32_t p = threadIdx.x + TTHREADS * k;
            a.pst[(size_t)tile * TP + p] = st[k];
            const uint32_t nv = (st[k] >> 10) & 0xfu;  // views only append: store the new ones
#pragma unroll
            for (uint32_t v = 0; v < VMAX; ++v)
                if (v >= nv0[k] && v < nv) a.pviews[((size_t)tile * VMAX + v) * TP + p] = vw[k][v];
        }
        if (threadIdx.x == 0) {
            a.tcnt[tile] = 0;
            a.tcur[tile] = 0;
        }
    }
}

// ---------------------------------------------------------------- fetch rows

__global__ void c_fetch(CArgs a, uint32_t nrows)
{
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += gridDim.x * blockDim.x) {
        if (!a.rflag[r]) continue;
        const uint32_t f = a.rfpos[r], i = a.racc[r];
        a.fx0[f] = a.rx0[r]; a.fx1[f] = a.rx1[r]; a.fy0[f] = a.ry0[r]; a.fy1[f] = a.ry1[r];
        a.fsd[f] = (a.rpay[r] >> 9) & 1u;
        a.fsc[f] = a.rsrc[r];
        const bool gpu = a.bflags[i] & F_GPU;
        a.fdd[f] = gpu ? 1u : 0u;
        a.fdc[f] = gpu ? a.hchunk[a.bhist[i]] : NONE;
    }
}

__global__ void c_foff(CArgs a, uint32_t nrows, uint32_t nfetch)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n) a.foff[i] = a.rowbase[i] < nrows ? a.rfpos[a.rowbase[i]] : nfetch;
}

// registration rows of this plan call, compacted in fragment-row order
__global__ void c_regs(CArgs a, uint32_t nrows)
{
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += gridDim.x * blockDim.x) {
        if (!a.rreg[r]) continue;
        const uint32_t i = a.racc[r];
        uint32_t *o = a.reg + 8ull * a.rrpos[r];
        o[0] = a.rx0[r]; o[1] = a.rx1[r]; o[2] = a.ry0[r]; o[3] = a.ry1[r];
        o[4] = (a.rpay[r] >> 12) & 3u;
        o[5] = a.rsrc[r];
        o[6] = ((a.rpay[r] >> 12) & 3u) == 3u ? a.hchunk[a.bhist[i]] : NONE;
        o[7] = i;
    }
}

// Completions (fetch_callback, memory-tree.hpp:955-1061, SEARCH_FETCHED
// 2067-2127), one block per completed fetch: q = 6 u32 per fetch (x0, x1 in
// sizeof_type units, y0, y1, dst device, dst chunk).  On every point of the
// rect the view holding the chunk becomes coherent and stops fetching, the
// device becomes coherent, and stops fetching once none of its views is.
// Completions of one call may share points: the updates are atomic and
// commute, and the one that clears a point's last fetching view clears the
// device bit.
__global__ void __launch_bounds__(256) c_complete(CArgs a, const uint32_t *q, uint32_t n)
{
    for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
        const uint32_t u0 = q[6 * k], u1 = q[6 * k + 1], y0 = q[6 * k + 2], y1 = q[6 * k + 3];
        const uint32_t dd = q[6 * k + 4], dc = q[6 * k + 5];
        const uint32_t w = u1 - u0;
        const uint64_t npts = (uint64_t)w * (y1 - y0);
        for (uint64_t p = threadIdx.x; p < npts; p += blockDim.x) {
            const uint32_t u = u0 + (uint32_t)(p % w), y = y0 + (uint32_t)(p / w);
            const uint32_t tile = (y / TH) * a.ntx + u / TW, off = (y % TH) * TW + (u % TW);
            uint32_t *word = a.pst + (size_t)tile * TP + off;
            if (dd == 1) {
                const uint32_t nv = (*word >> 10) & 0xfu;  // the views do not change while completing
                uint32_t bit = 0;
                for (uint32_t v = 0; v < nv && !bit; ++v)
                    if (a.pviews[((size_t)tile * VMAX + v) * TP + off] == dc) bit = 1u << v;
                atomicOr(word, (bit << 2) | 2u);
                const uint32_t old = atomicAnd(word, ~(bit << ST_VF));
                if ((((old >> ST_VF) & 0xffu) & ~bit) == 0) atomicAnd(word, ~(2u << ST_DF));
            } else {
                atomicOr(word, 1u);
                atomicAnd(word, ~(1u << ST_DF));
            }
        }
    }
}

// ---------------------------------------------------------------- fetch execution

// One 2D copy per fetch (the reference issues hipMemcpyParam2DAsync per fetch,
// src/driver/driver_hip.cc:661-751): `height` columns of `width` bytes from
// src (pitch spitch) to dst (pitch dpitch).  Host memory is reached through
// its device mapping; every fetch of a batch goes in one launch, a block per
// fetch at a time, 16-byte moves when all of a fetch's addresses allow them.
struct CopyJob {
    unsigned long long src, dst;
    uint32_t spitch, dpitch, width, height;
};
static_assert(sizeof(CopyJob) == 32, "copy jobs are 32 B");

__global__ void __launch_bounds__(256) c_exec(const CopyJob *jobs, uint32_t njobs)
{
    for (uint32_t j = blockIdx.x; j < njobs; j += gridDim.x) {
        const CopyJob c = jobs[j];
        const bool v16 = ((c.src | c.dst | c.spitch | c.dpitch | c.width) & 15ull) == 0;
        const uint32_t unit = v16 ? 16u : 1u;
        const uint32_t per_col = c.width / unit;
        const uint64_t total = (uint64_t)per_col * c.height;
        for (uint64_t t = threadIdx.x; t < total; t += blockDim.x) {
            const uint32_t col = (uint32_t)(t / per_col), k = (uint32_t)(t % per_col);
            const unsigned long long so = c.src + (unsigned long long)col * c.spitch + (unsigned long long)k * unit;
            const unsigned long long d = c.dst + (unsigned long long)col * c.dpitch + (unsigned long long)k * unit;
            if (v16) *reinterpret_cast<uint4 *>(d) = *reinterpret_cast<const uint4 *>(so);
            else *reinterpret_cast<unsigned char *>(d) = *reinterpret_cast<const unsigned char *>(so);
        }
    }
}

// ---------------------------------------------------------------- who_owns

// memory-tree.hpp:1903-1933 (SEARCH_OWNERS 2130-2139): bytes per coherent
// device over the query rect, then the devices owning the most bytes
__global__ void __launch_bounds__(256) c_owners(CArgs a, const uint32_t *qx0, const uint32_t *qx1,
                                                const uint32_t *qy0, const uint32_t *qy1, uint32_t nq,
                                                uint32_t *owners)
{
    __shared__ unsigned long long part[2][4];
    for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
        const uint32_t x0 = qx0[q], x1 = qx1[q], y0 = qy0[q], y1 = qy1[q];
        const uint32_t w = x1 - x0;
        const uint64_t npts = (uint64_t)w * (y1 - y0);
        unsigned long long h = 0, g = 0;
        for (uint64_t p = threadIdx.x; p < npts; p += blockDim.x) {
            const uint32_t u = x0 + (uint32_t)(p % w), y = y0 + (uint32_t)(p / w);
            const uint32_t tile = (y / TH) * a.ntx + u / TW;
            const uint32_t c = a.pst[(size_t)tile * TP + (y % TH) * TW + (u % TW)] & 3u;
            h += c & 1u;
            g += (c >> 1) & 1u;
        }
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) {
            h += __shfl_down(h, d, 64);
            g += __shfl_down(g, d, 64);
        }
        if (lane_id() == 0) {
            part[0][threadIdx.x >> 6] = h;
            part[1][threadIdx.x >> 6] = g;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned long long bh = (part[0][0] + part[0][1] + part[0][2] + part[0][3]) * a.s;
            const unsigned long long bg = (part[1][0] + part[1][1] + part[1][2] + part[1][3]) * a.s;
            uint32_t o = 0;
            if (bh > 0 || bg > 0) o = bh > bg ? 1u : (bg > bh ? 2u : 3u);
            owners[q] = o;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- who_owns without allocation
//
// xkrt_hip_mcc_who_owns_static stages its queries (the caller's access
// records), their rects and the answers in one static device array instead of
// per-call allocations.  Each GPU holds its own copy of the array until the
// process ends, and every call on a GPU reuses it, so calls that overlap on
// one GPU race.  In host code the array's name is not a device address: its
// address on each device comes from hipGetSymbolAddress, once per device, when
// the first tree is created there (scratch_setup), and only that address is
// used.  Every word a call reads is written by the same call first (the
// copied queries, then their rects, then one answer per query), so nothing
// has to be cleared between calls.
constexpr uint64_t WHO_OWNS_STATIC_MAX = XKRT_HIP_WHO_OWNS_STATIC_MAX;  // queries per call
constexpr size_t OWN_ACC_BYTES = WHO_OWNS_STATIC_MAX * sizeof(xkrt_hip_mcc_access_t);
constexpr size_t OWN_Q_BYTES = WHO_OWNS_STATIC_MAX * 16;  // x0, x1, y0, y1 (u32, SoA)
constexpr size_t OWN_OUT_BYTES = WHO_OWNS_STATIC_MAX * 4;
static_assert(OWN_ACC_BYTES % 256 == 0 && OWN_Q_BYTES % 256 == 0, "scratch sections stay 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[OWN_ACC_BYTES + OWN_Q_BYTES + OWN_OUT_BYTES];

// the queries' rects in the tree frame (x in sizeof_type units, y relative to
// the frame's first column), from access records already checked on the host
__global__ void __launch_bounds__(256) c_owners_prep(const xkrt_hip_mcc_access_t *acc, uint32_t n, uint64_t LDs,
                                                     uint32_t s, uint64_t ylo, uint32_t *q)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const xkrt_hip_mcc_access_t a = acc[i];
        const uint64_t xb = a.addr % LDs, col = a.addr / LDs;
        q[i] = (uint32_t)(xb / s);
        q[n + i] = (uint32_t)(xb / s + a.m);
        q[2 * n + i] = (uint32_t)(col - ylo);
        q[3 * n + i] = (uint32_t)(col - ylo + a.n);
    }
}

// ---------------------------------------------------------------- scan

__global__ void __launch_bounds__(SCAN_BLOCK) s_reduce(const uint32_t *in, uint32_t n, uint32_t *bsum)
{
    __shared__ uint32_t ws[SCAN_BLOCK / 64];
    const uint32_t base = blockIdx.x * SCAN_CHUNK;
    uint32_t v = 0;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_PER; ++k) {
        const uint32_t i = base + k * SCAN_BLOCK + threadIdx.x;
        v += i < n ? in[i] : 0u;
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
    if (lane_id() == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < SCAN_BLOCK / 64; ++w) t += ws[w];
        bsum[blockIdx.x] = t;
    }
}

// exclusive scan of the block sums in place; total to *total
__global__ void __launch_bounds__(1024) s_top(uint32_t *bsum, uint32_t nb, uint32_t *total)
{
    __shared__ uint32_t ws[16];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t c = 0; c < nb; c += 1024) {
        const uint32_t i = c + threadIdx.x;
        const uint32_t v = i < nb ? bsum[i] : 0u;
        uint32_t x = v;
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(x, d, 64);
            if (lane_id() >= d) x += y;
        }
        if (lane_id() == 63) ws[threadIdx.x >> 6] = x;
        __syncthreads();
        uint32_t off = carry;
        for (uint32_t w = 0; w < (threadIdx.x >> 6); ++w) off += ws[w];
        if (i < nb) bsum[i] = off + x - v;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (uint32_t w = 0; w < 16; ++w) t += ws[w];
            carry += t;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(SCAN_BLOCK) s_down(const uint32_t *in, uint32_t n, const uint32_t *bsum,
                                                     uint32_t *out)
{
    __shared__ uint32_t ws[SCAN_BLOCK / 64];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = bsum[blockIdx.x];
    __syncthreads();
    const uint32_t base = blockIdx.x * SCAN_CHUNK;
    for (uint32_t k = 0; k < SCAN_PER; ++k) {
        const uint32_t i = base + k * SCAN_BLOCK + threadIdx.x;
        const uint32_t v = i < n ? in[i] : 0u;
        uint32_t x = v;
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(x, d, 64);
            if (lane_id() >= d) x += y;
        }
        if (lane_id() == 63) ws[threadIdx.x >> 6] = x;
        __syncthreads();
        uint32_t off = carry;
        for (uint32_t w = 0; w < (threadIdx.x >> 6); ++w) off += ws[w];
        if (i < n) out[i] = off + x - v;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (uint32_t w = 0; w < SCAN_BLOCK / 64; ++w) t += ws[w];
            carry += t;
        }
        __syncthreads();
    }
}

int cerr(hipError_t e)
{
    if (e == hipSuccess) return 0;
    if (e == hipErrorOutOfMemory) return -ENOMEM;
    return -ENODEV;
}

#define CCHK(x) do { hipError_t _e = (x); if (_e != hipSuccess) return cerr(_e); } while (0)

template <typename T>
int grow(T **p, uint64_t *cap, uint64_t need, bool keep, hipStream_t s)
{
    if (need <= *cap) return 0;
    uint64_t nc = std::max<uint64_t>(need, *cap + *cap / 2);
    nc = std::max<uint64_t>(nc, 1024);
    T *q = nullptr;
    CCHK(hipMalloc(&q, nc * sizeof(T)));
    if (keep && *p && *cap) CCHK(hipMemcpyAsync(q, *p, *cap * sizeof(T), hipMemcpyDeviceToDevice, s));
    if (*p) {
        CCHK(hipStreamSynchronize(s));
        (void)hipFree(*p);
    }
    *p = q;
    *cap = nc;
    return 0;
}

uint32_t waves_grid(uint64_t n) { return (uint32_t)std::min<uint64_t>(std::max<uint64_t>((n + 3) / 4, 1), 65536); }

// per device: the address of `scratch` on that device (0 until looked up)
constexpr int MAX_DEVICES = 64;
void *g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// looks up the address of `scratch` on `device`, which must be current; once per device
int scratch_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return -ENODEV;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch[device]) return 0;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return -ENODEV;
    g_scratch[device] = p;
    return 0;
}

void *scratch_of(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    return g_scratch[device];
}

// the checks of who_owns on one query, and its rect in the tree frame
int owner_query(const xkrt_hip_mcc *h, const xkrt_hip_mcc_access_t &a);

} // namespace

// fetches in flight and registrations (trees with asynchronous completion)
struct MccFetch {
    uint32_t u0, u1, y0, y1;  // rect: x in sizeof_type units, y in frame columns
    uint32_t dd, dc;          // destination device and chunk (NONE: host memory)
    bool inflight;
    uint32_t sd = 0, sc = NONE;  // source device and chunk
};
struct MccReg {
    uint64_t acc;             // access index in the tree's order
    uint32_t x0, x1, y0, y1;  // x in bytes, y in frame columns
    uint32_t kind;            // 2 await, 3 forward
    uint32_t chunk, dst;      // the chunk whose fetch fires it; a forward's destination chunk
};

struct xkrt_hip_mcc {
    int device = 0;
    bool async = false;
    uint64_t ld = 0, s = 0, LDs = 0;
    xkrt_hip_mcc_config_t cfg{};
    uint32_t nxu = 0, ny = 0, ntx = 0, nty = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool gpu_seen = false;
    bool poisoned = false;
    uint64_t naccess = 0;   // accesses planned so far (global ids)
    uint32_t nhist = 0;     // GPU accesses so far
    uint64_t device_bytes = 0;
    // tree
    uint32_t *TX = nullptr, *TY = nullptr, *pst = nullptr, *pviews = nullptr;
    uint32_t *tcnt = nullptr, *tcur = nullptr, *tstart = nullptr, *tact = nullptr;
    uint32_t *tbsum = nullptr;
    uint64_t tbsum_cap = 0;
    CCounters *ctr = nullptr, *ctr_host = nullptr;
    // GPU history
    uint32_t *hx0 = nullptr, *hx1 = nullptr, *hy0 = nullptr, *hy1 = nullptr, *hbeta = nullptr, *hchunk = nullptr;
    unsigned long long *hbytes = nullptr;
    uint64_t hcap = 0, hcap1 = 0, hcap2 = 0, hcap3 = 0, hcap4 = 0, hcap5 = 0, hcap6 = 0;
    unsigned long long *chunk_size = nullptr;
    uint64_t chunk_cap = 0;
    // batch
    uint32_t *bx0 = nullptr, *bx1 = nullptr, *by0 = nullptr, *by1 = nullptr, *bflags = nullptr, *bhist = nullptr,
             *bxoff = nullptr, *rowbase = nullptr, *foff = nullptr;
    uint64_t bcap[9] = {};
    uint32_t *xcnt = nullptr, *xrow = nullptr, *xitems = nullptr, *xioff = nullptr, *xend = nullptr, *xacc = nullptr;
    uint64_t xcap[6] = {};
    uint32_t *imap = nullptr, *icnt = nullptr, *irow = nullptr;
    uint64_t icap[3] = {};
    uint32_t b_xtot = 0, b_irows = 1, b_items = 0, b_rows = 0;  // the plan call in progress
    uint32_t state_grid = 2048;  // c_state workgroups: as many as the device holds at once
    uint32_t *tinc = nullptr;
    uint64_t tinc_cap = 0;
    TDesc *tdesc = nullptr;
    uint64_t tdesc_cap = 0;
    uint32_t *bsum = nullptr;
    uint64_t bsum_cap = 0;
    // rows
    uint32_t *rx0 = nullptr, *rx1 = nullptr, *ry0 = nullptr, *ry1 = nullptr, *rpay = nullptr, *rsrc = nullptr,
             *racc = nullptr, *rflag = nullptr, *rfpos = nullptr;
    uint64_t rcap[9] = {};
    uint32_t *fx0 = nullptr, *fx1 = nullptr, *fy0 = nullptr, *fy1 = nullptr, *fsd = nullptr, *fsc = nullptr,
             *fdd = nullptr, *fdc = nullptr;
    uint64_t fcap[8] = {};
    // fetches in flight (async): registration flags and rows, completion queue
    uint32_t *rreg = nullptr, *rrpos = nullptr, *reg = nullptr, *cq = nullptr;
    uint64_t rregcap[2] = {}, regcap = 0, cqcap = 0;
    std::vector<MccFetch> fetches;      // every fetch issued, by id
    std::vector<MccReg> regs;           // every registration, by id
    std::vector<uint32_t> live;         // registrations not fired yet, ascending
    uint64_t inflight = 0;
    uint64_t last_fetch_base = 0;
    std::vector<int64_t> last_reg_off, last_regs;                 // per access of the last plan call
    std::vector<int64_t> c_rel_off, c_rel, c_fwd_off, c_fwd;      // the last completion call
    // chunk geometry: the allocating access's rect (x in units, y in frame columns), by chunk id
    std::vector<std::array<uint32_t, 4>> chunk_rect;
    std::vector<int64_t> last_chunks;  // per access of the last plan call: its GPU chunk, -1 on the host
    // fetch execution: the host matrix (device mapping), the chunk arena, placements, job buffers
    unsigned char *host_dev = nullptr;
    unsigned char *arena = nullptr;
    uint64_t arena_bytes = 0, arena_used = 0;
    std::vector<uint64_t> chunk_off;   // NONE64: not placed yet
    CopyJob *jobs[2] = {};             // pinned, mapped; used in turn
    uint64_t jobs_cap[2] = {};
    hipEvent_t jobs_ev[2] = {};
    bool jobs_pending[2] = {};
    uint32_t jobs_cur = 0;
    // results of the last plan call
    uint32_t last_n = 0, last_rows = 0, last_fetch = 0;
    int64_t last_fail = -1;
    std::vector<uint32_t> h_rowbase, h_foff;
    // timing of the last plan call (HIP events around its kernels)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.f;
};

namespace {

// exclusive scan of n u32 into out; the total lands in *total (device)
int scan(xkrt_hip_mcc *h, const uint32_t *in, uint32_t n, uint32_t *out, uint32_t *total)
{
    const uint32_t nb = std::max<uint32_t>(1, (n + SCAN_CHUNK - 1) / SCAN_CHUNK);
    int rc = grow(&h->bsum, &h->bsum_cap, nb, false, h->stream);
    if (rc) return rc;
    hipLaunchKernelGGL(s_reduce, dim3(nb), dim3(SCAN_BLOCK), 0, h->stream, in, n, h->bsum);
    hipLaunchKernelGGL(s_top, dim3(1), dim3(1024), 0, h->stream, h->bsum, nb, total);
    hipLaunchKernelGGL(s_down, dim3(nb), dim3(SCAN_BLOCK), 0, h->stream, in, n, h->bsum, out);
    return cerr(hipGetLastError());
}

CArgs make_cargs(xkrt_hip_mcc *h, uint32_t n)
{
    CArgs a{};
    a.n = n;
    a.gid0 = (uint32_t)h->naccess;
    a.bx0 = h->bx0; a.bx1 = h->bx1; a.by0 = h->by0; a.by1 = h->by1; a.bflags = h->bflags; a.bhist = h->bhist;
    a.bxoff = h->bxoff; a.rowbase = h->rowbase;
    a.nxu = h->nxu; a.ny = h->ny; a.ny1 = h->ny + 1; a.s = (uint32_t)h->s;
    a.TX = h->TX; a.TY = h->TY; a.ntx = h->ntx; a.nty = h->nty; a.pst = h->pst; a.pviews = h->pviews;
    a.hx0 = h->hx0; a.hx1 = h->hx1; a.hy0 = h->hy0; a.hy1 = h->hy1; a.hbytes = h->hbytes;
    a.hbeta = h->hbeta; a.hchunk = h->hchunk; a.chunk_size = h->chunk_size;
    a.xcnt = h->xcnt; a.xrow = h->xrow; a.xitems = h->xitems; a.xioff = h->xioff; a.xend = h->xend;
    a.xacc = h->xacc;
    a.xtot = h->b_xtot; a.irows = h->b_irows; a.nitems = h->b_items; a.nrows = h->b_rows;
    a.imap = h->imap; a.icnt = h->icnt; a.irow = h->irow;
    a.rx0 = h->rx0; a.rx1 = h->rx1; a.ry0 = h->ry0; a.ry1 = h->ry1; a.rpay = h->rpay; a.rsrc = h->rsrc;
    a.racc = h->racc; a.rflag = h->rflag; a.rfpos = h->rfpos;
    a.fx0 = h->fx0; a.fx1 = h->fx1; a.fy0 = h->fy0; a.fy1 = h->fy1; a.fsd = h->fsd; a.fsc = h->fsc;
    a.fdd = h->fdd; a.fdc = h->fdc; a.foff = h->foff;
    a.tcnt = h->tcnt; a.tcur = h->tcur; a.tstart = h->tstart; a.tinc = h->tinc; a.tact = h->tact;
    a.tdesc = h->tdesc;
    a.ctr = h->ctr;
    a.async = h->async ? 1u : 0u;
    a.rreg = h->rreg; a.rrpos = h->rrpos; a.reg = h->reg;
    return a;
}

void free_all(xkrt_hip_mcc *h)
{
    void *ptrs[] = {h->TX, h->TY, h->pst, h->pviews, h->tcnt, h->tcur, h->tstart, h->tact, h->tbsum, h->ctr,
                    h->hx0, h->hx1, h->hy0, h->hy1, h->hbeta, h->hchunk, h->hbytes, h->chunk_size,
                    h->bx0, h->bx1, h->by0, h->by1, h->bflags, h->bhist, h->bxoff, h->rowbase, h->foff,
                    h->xcnt, h->xrow, h->xitems, h->xioff, h->xend, h->xacc, h->imap, h->icnt, h->irow, h->tinc, h->tdesc, h->bsum, h->rx0, h->rx1, h->ry0, h->ry1, h->rpay, h->rsrc, h->racc,
                    h->rflag, h->rfpos, h->fx0, h->fx1, h->fy0, h->fy1, h->fsd, h->fsc, h->fdd, h->fdc,
                    h->rreg, h->rrpos, h->reg, h->cq};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (h->ctr_host) (void)hipHostFree(h->ctr_host);
    for (int b = 0; b < 2; ++b) {
        if (h->jobs_ev[b]) {
            if (h->jobs_pending[b]) (void)hipEventSynchronize(h->jobs_ev[b]);
            (void)hipEventDestroy(h->jobs_ev[b]);
        }
        if (h->jobs[b]) (void)hipHostFree(h->jobs[b]);
    }
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
}

} // namespace

extern "C" {

int xkrt_hip_mcc_create(uint64_t ld, uint64_t sizeof_type, const xkrt_hip_mcc_config_t *cfg, void *stream,
                        xkrt_hip_mcc_t *out)
{
    if (!out || !cfg || ld == 0 || sizeof_type == 0) return -EINVAL;
    *out = nullptr;
    if (cfg->y_hi <= cfg->y_lo || cfg->y_hi - cfg->y_lo > (1ull << 24) || ld > (1ull << 24)) return -EINVAL;
    if (ld * sizeof_type > 0xFFFFFFFFull) return -ERANGE;
    xkrt_hip_mcc *h = new xkrt_hip_mcc();
    h->ld = ld; h->s = sizeof_type; h->LDs = ld * sizeof_type;
    h->cfg = *cfg;
    h->async = cfg->async_completion != 0;
    h->nxu = (uint32_t)ld;
    h->ny = (uint32_t)(cfg->y_hi - cfg->y_lo);
    h->ntx = (h->nxu + TW - 1) / TW;
    h->nty = (h->ny + TH - 1) / TH;
    const uint64_t ntiles = (uint64_t)h->ntx * h->nty;
    const uint64_t tbytes = 4ull * (h->nxu + 1) + 4ull * h->nxu * (h->ny + 1ull) + ntiles * TP * 4ull * (1 + VMAX) +
                            ntiles * 16ull;
    if (cfg->budget_bytes && tbytes > cfg->budget_bytes) {
        delete h;
        return -ENOMEM;
    }
    int rc = 0;
    auto chk = [&](hipError_t e) { if (!rc) rc = cerr(e); };
    chk(hipGetDevice(&h->device));
    if (!rc) rc = scratch_setup(h->device);  // the static staging of who_owns_static on this device
    if (stream) {
        h->stream = (hipStream_t)stream;
    } else {
        chk(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        h->own_stream = true;
    }
    chk(hipMalloc(&h->TX, 4ull * (h->nxu + 1)));
    chk(hipMalloc(&h->TY, 4ull * h->nxu * (h->ny + 1ull)));
    chk(hipMalloc(&h->pst, 4ull * ntiles * TP));
    chk(hipMalloc(&h->pviews, 4ull * ntiles * TP * VMAX));
    chk(hipMalloc(&h->tcnt, 4ull * ntiles));
    chk(hipMalloc(&h->tcur, 4ull * ntiles));
    chk(hipMalloc(&h->tstart, 4ull * ntiles));
    chk(hipMalloc(&h->tact, 4ull * ntiles));
    chk(hipMalloc(&h->ctr, sizeof(CCounters)));
    chk(hipHostMalloc(&h->ctr_host, sizeof(CCounters), hipHostMallocDefault));
    chk(hipEventCreate(&h->ev0));
    chk(hipEventCreate(&h->ev1));
    if (!rc) {
        chk(hipMemsetAsync(h->TX, 0xFF, 4ull * (h->nxu + 1), h->stream));
        chk(hipMemsetAsync(h->TY, 0xFF, 4ull * h->nxu * (h->ny + 1ull), h->stream));
        chk(hipMemsetAsync(h->pst, 0, 4ull * ntiles * TP, h->stream));
        chk(hipMemsetAsync(h->pviews, 0xFF, 4ull * ntiles * TP * VMAX, h->stream));
        chk(hipMemsetAsync(h->tcnt, 0, 4ull * ntiles, h->stream));
        chk(hipMemsetAsync(h->tcur, 0, 4ull * ntiles, h->stream));
        CCounters c0{};
        c0.fail = NONE;
        c0.viewerr = NONE;
        chk(hipMemcpyAsync(h->ctr, &c0, sizeof(c0), hipMemcpyHostToDevice, h->stream));
        chk(hipStreamSynchronize(h->stream));
    }
    if (rc) {
        free_all(h);
        if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
        delete h;
        return rc;
    }
    {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, c_state, TTHREADS, 0) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) == hipSuccess &&
            per_cu > 0 && cus > 0)
            h->state_grid = (uint32_t)(per_cu * cus);
    }
    h->device_bytes = tbytes;
    *out = h;
    return 0;
}

int xkrt_hip_mcc_destroy(xkrt_hip_mcc_t h)
{
    if (!h) return -EINVAL;
    (void)hipStreamSynchronize(h->stream);
    free_all(h);
    if (h->own_stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

namespace {

// after an asynchronous plan call: its fetch rows (up to the abort) are in flight
// with ids from last_fetch_base, its registrations join the live list
int record_async(xkrt_hip_mcc *h, uint64_t n, uint64_t keep, uint32_t nreg)
{
    const uint32_t nf = h->h_foff[keep];
    if (nf) {
        std::vector<uint32_t> v[8];
        uint32_t *src[8] = {h->fx0, h->fx1, h->fy0, h->fy1, h->fdd, h->fdc, h->fsd, h->fsc};
        for (int k = 0; k < 8; ++k) {
            v[k].resize(nf);
            CCHK(hipMemcpy(v[k].data(), src[k], 4ull * nf, hipMemcpyDeviceToHost));
        }
        for (uint32_t r = 0; r < nf; ++r) {
            MccFetch f{v[0][r] / (uint32_t)h->s, v[1][r] / (uint32_t)h->s, v[2][r], v[3][r], v[4][r], v[5][r], true};
            f.sd = v[6][r];
            f.sc = v[7][r];
            h->fetches.push_back(f);
        }
        h->inflight += nf;
    }
    h->last_reg_off.assign(n + 1, 0);
    h->last_regs.clear();
    std::vector<uint32_t> rows(8ull * nreg);
    if (nreg) CCHK(hipMemcpy(rows.data(), h->reg, 32ull * nreg, hipMemcpyDeviceToHost));
    std::vector<uint64_t> per(n + 1, 0);
    for (uint32_t r = 0; r < nreg; ++r) {
        const uint32_t *o = rows.data() + 8ull * r;
        if (o[7] >= keep) continue;  // after the abort: not planned
        const uint32_t id = (uint32_t)h->regs.size();
        h->regs.push_back({h->naccess + o[7], o[0], o[1], o[2], o[3], o[4], o[5], o[6]});
        h->live.push_back(id);
        per[o[7] + 1]++;
        const int64_t row[8] = {o[0], o[1], (int64_t)(o[2] + h->cfg.y_lo), (int64_t)(o[3] + h->cfg.y_lo), o[4],
                                o[5] == NONE ? -1 : (int64_t)o[5], o[6] == NONE ? -1 : (int64_t)o[6], id};
        h->last_regs.insert(h->last_regs.end(), row, row + 8);
    }
    for (uint64_t i = 0; i < n; ++i) h->last_reg_off[i + 1] = h->last_reg_off[i] + (int64_t)per[i + 1];
    return 0;
}

} // namespace

int xkrt_hip_mcc_plan(xkrt_hip_mcc_t h, const xkrt_hip_mcc_access_t *acc, uint64_t n, uint64_t *nparts,
                      uint64_t *nfetches, int64_t *fail_at)
{
    if (!h || (!acc && n)) return -EINVAL;
    if (h->poisoned) return -ECANCELED;
    if (fail_at) *fail_at = -1;
    if (nparts) *nparts = 0;
    if (nfetches) *nfetches = 0;
    if (n == 0) return 0;
    if (h->naccess + n >= NONE || n > (1ull << 30)) return -ERANGE;

    // host side: rects in the tree frame, flags, history entries
    std::vector<uint32_t> x0(n), x1(n), y0(n), y1(n), fl(n), hist(n, NONE), xoff(n);
    std::vector<uint32_t> gx0, gx1, gy0, gy1;
    std::vector<unsigned long long> gbytes;
    uint64_t xtot = 0, ttot = 0;
    bool seen = h->gpu_seen;
    for (uint64_t i = 0; i < n; ++i) {
        const xkrt_hip_mcc_access_t &q = acc[i];
        if (q.m == 0 || q.n == 0 || q.dev > 1) return -EINVAL;
        const uint64_t xb = q.addr % h->LDs, col = q.addr / h->LDs;
        if (xb % h->s) return -EINVAL;
        if (xb + q.m * h->s > h->LDs) return -ENOTSUP;  // column-wrapping tile: two rects
        if (col < h->cfg.y_lo || col + q.n > h->cfg.y_hi) return -ERANGE;
        x0[i] = (uint32_t)(xb / h->s);
        x1[i] = (uint32_t)(xb / h->s + q.m);
        y0[i] = (uint32_t)(col - h->cfg.y_lo);
        y1[i] = (uint32_t)(col - h->cfg.y_lo + q.n);
        uint32_t f = 0;
        if (q.dev == 1) {
            f |= F_GPU | F_INS;
            seen = true;
            hist[i] = h->nhist + (uint32_t)gx0.size();
            gx0.push_back(x0[i]); gx1.push_back(x1[i]); gy0.push_back(y0[i]); gy1.push_back(y1[i]);
            gbytes.push_back(q.m * q.n * h->s);
        } else if (seen) {
            f |= F_INS;  // fetch_list_to_host: nothing while the tree is empty (1191-1192)
        }
        if (q.mode & XKRT_HIP_MODE_R) f |= F_R;
        if (q.mode & XKRT_HIP_MODE_W) f |= F_W;
        fl[i] = f;
        xoff[i] = (uint32_t)xtot;
        if (f & F_INS) {
            xtot += x1[i] - x0[i];
            ttot += (uint64_t)((x1[i] - 1) / TW - x0[i] / TW + 1) * ((y1[i] - 1) / TH - y0[i] / TH + 1);
        }
    }
    if (xtot >= NONE || ttot >= NONE) return -ERANGE;
    // rows per work item: ITEM_ROWS, doubled while the items could overflow
    // their u32 count (one item per slab at most 1 << 24 rows: xtot items)
    uint32_t irows = ITEM_ROWS;
    for (;;) {
        uint64_t bound = 0;
        for (uint64_t i = 0; i < n; ++i)
            if (fl[i] & F_INS) bound += (uint64_t)(x1[i] - x0[i]) * ((y1[i] - y0[i] + irows - 1) / irows);
        if (bound < NONE || irows >= (1u << 24)) break;
        irows <<= 1;
    }
    const uint32_t ng = (uint32_t)gx0.size();
    hipStream_t s = h->stream;
    int rc = 0;
    // capacities
    uint64_t hneed = (uint64_t)h->nhist + ng;
    uint32_t **bufs[] = {&h->bx0, &h->bx1, &h->by0, &h->by1, &h->bflags, &h->bhist, &h->bxoff, &h->rowbase,
                         &h->foff};
    for (int k = 0; k < 9 && !rc; ++k) rc = grow(bufs[k], &h->bcap[k], n, false, s);
    if (!rc) rc = grow(&h->hx0, &h->hcap, hneed, true, s);
    if (!rc) rc = grow(&h->hx1, &h->hcap1, hneed, true, s);
    if (!rc) rc = grow(&h->hy0, &h->hcap2, hneed, true, s);
    if (!rc) rc = grow(&h->hy1, &h->hcap3, hneed, true, s);
    if (!rc) rc = grow(&h->hbeta, &h->hcap4, hneed, true, s);
    if (!rc) rc = grow(&h->hchunk, &h->hcap5, hneed, true, s);
    if (!rc) rc = grow(&h->hbytes, &h->hcap6, hneed, true, s);
    if (!rc) rc = grow(&h->chunk_size, &h->chunk_cap, hneed, true, s);
    uint32_t **xb[] = {&h->xcnt, &h->xrow, &h->xitems, &h->xioff, &h->xend, &h->xacc};
    for (int k = 0; k < 6 && !rc; ++k) rc = grow(xb[k], &h->xcap[k], std::max<uint64_t>(xtot, 1), false, s);
    if (!rc) rc = grow(&h->tinc, &h->tinc_cap, std::max<uint64_t>(ttot, 1), false, s);
    if (!rc) rc = grow(&h->tdesc, &h->tdesc_cap, n, false, s);
    if (!rc) rc = grow(&h->tbsum, &h->tbsum_cap, 1, false, s);
    if (rc) return rc;
    // uploads
    auto up = [&](uint32_t *dst, const void *src, uint64_t bytes) {
        if (!rc && bytes) rc = cerr(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    };
    up(h->bx0, x0.data(), 4 * n); up(h->bx1, x1.data(), 4 * n); up(h->by0, y0.data(), 4 * n);
    up(h->by1, y1.data(), 4 * n); up(h->bflags, fl.data(), 4 * n); up(h->bhist, hist.data(), 4 * n);
    up(h->bxoff, xoff.data(), 4 * n);
    up(h->hx0 + h->nhist, gx0.data(), 4ull * ng); up(h->hx1 + h->nhist, gx1.data(), 4ull * ng);
    up(h->hy0 + h->nhist, gy0.data(), 4ull * ng); up(h->hy1 + h->nhist, gy1.data(), 4ull * ng);
    if (!rc && ng) rc = cerr(hipMemcpyAsync(h->hbytes + h->nhist, gbytes.data(), 8ull * ng, hipMemcpyHostToDevice, s));
    if (rc) return rc;

    CCHK(hipEventRecord(h->ev0, s));
    CArgs a = make_cargs(h, (uint32_t)n);
    const uint32_t gw = waves_grid(n);
    if (ng) {
        hipLaunchKernelGGL(c_contain, dim3(waves_grid(ng)), dim3(256), 0, s, a, h->nhist, ng);
        hipLaunchKernelGGL(c_chunks, dim3(1), dim3(1024), 0, s, a, h->nhist, ng);
    }
    h->b_xtot = (uint32_t)xtot;
    h->b_irows = irows;
    h->b_items = h->b_rows = 0;
    a = make_cargs(h, (uint32_t)n);
    hipLaunchKernelGGL(c_breaks, dim3(gw), dim3(256), 0, s, a);
    // fragments: slab starts, then work items of (slab, row block), counted
    // and scanned into row offsets
    hipLaunchKernelGGL(c_slabs, dim3(gw), dim3(256), 0, s, a);
    rc = scan(h, h->xitems, (uint32_t)xtot, h->xioff, &h->ctr->items);
    if (rc) return rc;
    CCHK(hipMemcpyAsync(h->ctr_host, h->ctr, sizeof(CCounters), hipMemcpyDeviceToHost, s));
    CCHK(hipStreamSynchronize(s));
    const uint32_t nitems = xtot ? h->ctr_host->items : 0;
    uint32_t **ib[] = {&h->imap, &h->icnt, &h->irow};
    for (int k = 0; k < 3 && !rc; ++k) rc = grow(ib[k], &h->icap[k], std::max<uint32_t>(nitems, 1), false, s);
    if (rc) return rc;
    h->b_items = nitems;
    a = make_cargs(h, (uint32_t)n);
    if (xtot) hipLaunchKernelGGL(c_imap, dim3(std::min<uint64_t>((xtot + 255) / 256, 8192)), dim3(256), 0, s, a);
    if (nitems) hipLaunchKernelGGL(c_icount, dim3(waves_grid(nitems)), dim3(256), 0, s, a);
    rc = scan(h, h->icnt, nitems, h->irow, &h->ctr->rows);
    if (rc) return rc;
    CCHK(hipMemcpyAsync(h->ctr_host, h->ctr, sizeof(CCounters), hipMemcpyDeviceToHost, s));
    CCHK(hipStreamSynchronize(s));
    const uint32_t nrows = h->ctr_host->rows;
    h->b_rows = nrows;
    uint32_t **rb[] = {&h->rx0, &h->rx1, &h->ry0, &h->ry1, &h->rpay, &h->rsrc, &h->racc, &h->rflag, &h->rfpos};
    for (int k = 0; k < 9 && !rc; ++k) rc = grow(rb[k], &h->rcap[k], std::max<uint32_t>(nrows, 1), false, s);
    if (rc) return rc;
    if (h->async) {
        if (!rc) rc = grow(&h->rreg, &h->rregcap[0], std::max<uint32_t>(nrows, 1), false, s);
        if (!rc) rc = grow(&h->rrpos, &h->rregcap[1], std::max<uint32_t>(nrows, 1), false, s);
        if (rc) return rc;
        if (nrows) CCHK(hipMemsetAsync(h->rreg, 0, 4ull * nrows, s));
    }
    a = make_cargs(h, (uint32_t)n);
    if (nrows) CCHK(hipMemsetAsync(h->rflag, 0, 4ull * nrows, s));
    hipLaunchKernelGGL(c_irows, dim3(std::max(waves_grid(nitems), gw)), dim3(256), 0, s, a);
    if (nrows) hipLaunchKernelGGL(c_yfix, dim3(std::min<uint32_t>((nrows + 255) / 256, 8192)), dim3(256), 0, s, a);
    // tiles
    const uint32_t ntiles = h->ntx * h->nty;
    hipLaunchKernelGGL(c_tcount, dim3(gw), dim3(256), 0, s, a);
    rc = scan(h, h->tcnt, ntiles, h->tstart, h->tbsum);
    if (rc) return rc;
    CCHK(hipMemsetAsync(&h->ctr->nact, 0, 4, s));
    hipLaunchKernelGGL(c_tfill, dim3(gw), dim3(256), 0, s, a);
    hipLaunchKernelGGL(c_tact, dim3((ntiles + 255) / 256), dim3(256), 0, s, a, ntiles);
    hipLaunchKernelGGL(c_state, dim3(std::min<uint32_t>(ntiles, h->state_grid)), dim3(TTHREADS), 0, s, a);
    // fetch rows
    rc = scan(h, h->rflag, std::max<uint32_t>(nrows, 1), h->rfpos, &h->ctr->fetches);
    if (rc) return rc;
    if (h->async) rc = scan(h, h->rreg, std::max<uint32_t>(nrows, 1), h->rrpos, &h->ctr->regs);
    if (rc) return rc;
    CCHK(hipMemcpyAsync(h->ctr_host, h->ctr, sizeof(CCounters), hipMemcpyDeviceToHost, s));
    CCHK(hipStreamSynchronize(s));
    const uint32_t nf = nrows ? h->ctr_host->fetches : 0;
    uint32_t **fb[] = {&h->fx0, &h->fx1, &h->fy0, &h->fy1, &h->fsd, &h->fsc, &h->fdd, &h->fdc};
    for (int k = 0; k < 8 && !rc; ++k) rc = grow(fb[k], &h->fcap[k], std::max<uint32_t>(nf, 1), false, s);
    if (rc) return rc;
    const uint32_t nreg = h->async && nrows ? h->ctr_host->regs : 0;
    if (nreg) rc = grow(&h->reg, &h->regcap, 8ull * nreg, false, s);
    if (rc) return rc;
    a = make_cargs(h, (uint32_t)n);
    if (nrows) hipLaunchKernelGGL(c_fetch, dim3(std::min<uint32_t>((nrows + 255) / 256, 8192)), dim3(256), 0, s, a, nrows);
    if (nreg) hipLaunchKernelGGL(c_regs, dim3(std::min<uint32_t>((nrows + 255) / 256, 8192)), dim3(256), 0, s, a, nrows);
    hipLaunchKernelGGL(c_foff, dim3(((uint32_t)n + 255) / 256), dim3(256), 0, s, a, nrows, nf);
    CCHK(hipEventRecord(h->ev1, s));
    CCHK(hipGetLastError());
    CCHK(hipMemcpyAsync(h->ctr_host, h->ctr, sizeof(CCounters), hipMemcpyDeviceToHost, s));
    CCHK(hipStreamSynchronize(s));
    (void)hipEventElapsedTime(&h->last_ms, h->ev0, h->ev1);
    const CCounters c = *h->ctr_host;
    if (c.viewerr != NONE && c.viewerr < c.fail) {  // a reused chunk missing before any abort: internal error
        h->poisoned = true;
        return -EIO;
    }
    if (c.err) {
        h->poisoned = true;
        return c.err & CERR_TILE ? -ENOSPC : -EIO;
    }
    // results: offsets per access, cut at the access the reference would abort on
    h->h_rowbase.resize(n + 1);
    h->h_foff.resize(n + 1);
    CCHK(hipMemcpy(h->h_rowbase.data(), h->rowbase, 4 * n, hipMemcpyDeviceToHost));
    CCHK(hipMemcpy(h->h_foff.data(), h->foff, 4 * n, hipMemcpyDeviceToHost));
    h->h_rowbase[n] = nrows;
    h->h_foff[n] = nf;
    int64_t fail = -1;
    uint64_t keep = n;
    if (c.fail != NONE) {
        fail = (int64_t)c.fail - (int64_t)h->naccess;
        keep = (uint64_t)fail;
        for (uint64_t i = keep; i <= n; ++i) {
            h->h_rowbase[i] = h->h_rowbase[keep];
            h->h_foff[i] = h->h_foff[keep];
        }
        h->poisoned = true;  // the reference aborts (memory-tree.hpp:1362-1365)
    }
    h->last_n = (uint32_t)n;
    h->last_rows = h->h_rowbase[n];
    h->last_fetch = h->h_foff[n];
    h->last_fail = fail;
    h->last_fetch_base = h->fetches.size();
    h->last_chunks.assign(n, -1);
    if (ng) {  // geometry of the chunks this call allocated (its allocating accesses' rects)
        std::vector<uint32_t> beta(ng), chk(ng);
        CCHK(hipMemcpy(beta.data(), h->hbeta + h->nhist, 4ull * ng, hipMemcpyDeviceToHost));
        CCHK(hipMemcpy(chk.data(), h->hchunk + h->nhist, 4ull * ng, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < ng; ++k) {
            if (beta[k] != NONE) continue;
            if (h->chunk_rect.size() <= chk[k]) h->chunk_rect.resize(chk[k] + 1, {0, 0, 0, 0});
            h->chunk_rect[chk[k]] = {gx0[k], gx1[k], gy0[k], gy1[k]};
        }
        for (uint64_t i = 0; i < n; ++i)
            if (hist[i] != NONE) h->last_chunks[i] = chk[hist[i] - h->nhist];
    }
    if (h->async) {
        // the fetch rows go in flight; the registrations wait for their fetch
        rc = record_async(h, n, keep, nreg);
        if (rc) return rc;
    }
    h->naccess += n;
    h->nhist += ng;
    h->gpu_seen = seen;
    if (nparts) *nparts = h->last_rows;
    if (nfetches) *nfetches = h->last_fetch;
    if (fail_at) *fail_at = fail;
    return 0;
}

int xkrt_hip_mcc_parts_copy(xkrt_hip_mcc_t h, uint64_t *part_off, int64_t *parts, uint64_t cap)
{
    if (!h) return -EINVAL;
    const uint32_t nr = h->last_rows;
    if (part_off)
        for (uint32_t i = 0; i <= h->last_n; ++i) part_off[i] = h->h_rowbase.empty() ? 0 : h->h_rowbase[i];
    if (!parts || nr == 0) return 0;
    if (cap < nr) return -ENOSPC;
    std::vector<uint32_t> v[5];
    uint32_t *src[5] = {h->rx0, h->rx1, h->ry0, h->ry1, h->rpay};
    for (int k = 0; k < 5; ++k) {
        v[k].resize(nr);
        CCHK(hipMemcpy(v[k].data(), src[k], 4ull * nr, hipMemcpyDeviceToHost));
    }
    for (uint32_t r = 0; r < nr; ++r) {
        int64_t *o = parts + 6ull * r;
        o[0] = v[0][r]; o[1] = v[1][r];
        o[2] = (int64_t)(v[2][r] + h->cfg.y_lo); o[3] = (int64_t)(v[3][r] + h->cfg.y_lo);
        o[4] = v[4][r] & 3u;
        o[5] = (v[4][r] >> 2) & 3u;  // fetching bits before the access (0 with synchronous completion)
    }
    return 0;
}

int xkrt_hip_mcc_fetches_copy(xkrt_hip_mcc_t h, uint64_t *fetch_off, int64_t *fetches, uint64_t cap)
{
    if (!h) return -EINVAL;
    const uint32_t nf = h->last_fetch;
    if (fetch_off)
        for (uint32_t i = 0; i <= h->last_n; ++i) fetch_off[i] = h->h_foff.empty() ? 0 : h->h_foff[i];
    if (!fetches || nf == 0) return 0;
    if (cap < nf) return -ENOSPC;
    std::vector<uint32_t> v[8];
    uint32_t *src[8] = {h->fx0, h->fx1, h->fy0, h->fy1, h->fsd, h->fsc, h->fdd, h->fdc};
    for (int k = 0; k < 8; ++k) {
        v[k].resize(nf);
        CCHK(hipMemcpy(v[k].data(), src[k], 4ull * nf, hipMemcpyDeviceToHost));
    }
    for (uint32_t r = 0; r < nf; ++r) {
        int64_t *o = fetches + 8ull * r;
        o[0] = v[0][r]; o[1] = v[1][r];
        o[2] = (int64_t)(v[2][r] + h->cfg.y_lo); o[3] = (int64_t)(v[3][r] + h->cfg.y_lo);
        o[4] = v[4][r];
        o[5] = v[5][r] == NONE ? -1 : (int64_t)v[5][r];
        o[6] = v[6][r];
        o[7] = v[7][r] == NONE ? -1 : (int64_t)v[7][r];
    }
    return 0;
}

namespace {

int owner_query(const xkrt_hip_mcc *h, const xkrt_hip_mcc_access_t &a)
{
    const uint64_t xb = a.addr % h->LDs, col = a.addr / h->LDs;
    if (a.m == 0 || a.n == 0 || xb % h->s) return -EINVAL;
    if (xb + a.m * h->s > h->LDs) return -ENOTSUP;
    if (col < h->cfg.y_lo || col + a.n > h->cfg.y_hi) return -ERANGE;
    return 0;
}

} // namespace

int xkrt_hip_mcc_awaits_copy(xkrt_hip_mcc_t h, uint64_t *awaits_off, int64_t *rows, uint64_t cap)
{
    if (!h) return -EINVAL;
    if (!h->async) return -EINVAL;
    const uint64_t nr = h->last_regs.size() / 8;
    if (awaits_off)
        for (uint32_t i = 0; i <= h->last_n; ++i)
            awaits_off[i] = h->last_reg_off.size() > i ? (uint64_t)h->last_reg_off[i] : 0;
    if (!rows || nr == 0) return 0;
    if (cap < nr) return -ENOSPC;
    memcpy(rows, h->last_regs.data(), 8 * sizeof(int64_t) * nr);
    return 0;
}

int xkrt_hip_mcc_complete(xkrt_hip_mcc_t h, const uint64_t *fetch_ids, uint64_t n, uint64_t *nreleased,
                          uint64_t *nforwards)
{
    if (!h || (!fetch_ids && n)) return -EINVAL;
    if (!h->async) return -EINVAL;
    if (h->poisoned) return -ECANCELED;
    if (nreleased) *nreleased = 0;
    if (nforwards) *nforwards = 0;
    std::vector<uint32_t> q(6 * n);
    for (uint64_t k = 0; k < n; ++k) {
        const uint64_t id = fetch_ids[k];
        if (id >= h->fetches.size() || !h->fetches[id].inflight) return -EINVAL;
        for (uint64_t j = 0; j < k; ++j)
            if (fetch_ids[j] == id) return -EINVAL;
        const MccFetch &f = h->fetches[id];
        const uint32_t row[6] = {f.u0, f.u1, f.y0, f.y1, f.dd, f.dc};
        memcpy(q.data() + 6 * k, row, sizeof(row));
    }
    if (n) {
        int rc = grow(&h->cq, &h->cqcap, 6 * n, false, h->stream);
        if (rc) return rc;
        CCHK(hipMemcpyAsync(h->cq, q.data(), 24ull * n, hipMemcpyHostToDevice, h->stream));
        CArgs a = make_cargs(h, 0);
        hipLaunchKernelGGL(c_complete, dim3((uint32_t)std::min<uint64_t>(n, 4096)), dim3(256), 0, h->stream, a, h->cq,
                           (uint32_t)n);
        CCHK(hipGetLastError());
        CCHK(hipStreamSynchronize(h->stream));
    }
    // the registrations on each fetched view fire, in completion order then registration order
    h->c_rel_off.assign(n + 1, 0);
    h->c_fwd_off.assign(n + 1, 0);
    h->c_rel.clear();
    h->c_fwd.clear();
    for (uint64_t k = 0; k < n; ++k) {
        MccFetch &f = h->fetches[fetch_ids[k]];
        f.inflight = false;
        h->inflight--;
        if (f.dd == 1) {
            const uint32_t fx0 = f.u0 * (uint32_t)h->s, fx1 = f.u1 * (uint32_t)h->s;
            size_t w = 0;
            for (size_t i = 0; i < h->live.size(); ++i) {
                const uint32_t id = h->live[i];
                const MccReg &r = h->regs[id];
                const bool hit = r.chunk == f.dc && r.x0 < fx1 && fx0 < r.x1 && r.y0 < f.y1 && f.y0 < r.y1;
                if (!hit) {
                    h->live[w++] = id;
                    continue;
                }
                const int64_t y0 = (int64_t)(r.y0 + h->cfg.y_lo), y1 = (int64_t)(r.y1 + h->cfg.y_lo);
                if (r.kind == 2) {
                    const int64_t row[7] = {(int64_t)r.acc, r.x0, r.x1, y0, y1, r.chunk, id};
                    h->c_rel.insert(h->c_rel.end(), row, row + 7);
                } else {
                    // the forward copy: from the view just fetched to the forward's view
                    const uint64_t fid = h->fetches.size();
                    const int64_t row[9] = {(int64_t)fid, r.x0, r.x1, y0, y1, 1, r.chunk, 1, r.dst};
                    h->c_fwd.insert(h->c_fwd.end(), row, row + 9);
                    MccFetch nf{r.x0 / (uint32_t)h->s, r.x1 / (uint32_t)h->s, r.y0, r.y1, 1u, r.dst, true};
                    nf.sd = 1;
                    nf.sc = r.chunk;
                    h->fetches.push_back(nf);
                    h->inflight++;
                }
            }
            h->live.resize(w);
        }
        h->c_rel_off[k + 1] = (int64_t)(h->c_rel.size() / 7);
        h->c_fwd_off[k + 1] = (int64_t)(h->c_fwd.size() / 9);
    }
    if (nreleased) *nreleased = h->c_rel.size() / 7;
    if (nforwards) *nforwards = h->c_fwd.size() / 9;
    return 0;
}

int xkrt_hip_mcc_completed_copy(xkrt_hip_mcc_t h, uint64_t *rel_off, int64_t *released, uint64_t rel_cap,
                                uint64_t *fwd_off, int64_t *forwards, uint64_t fwd_cap)
{
    if (!h) return -EINVAL;
    const uint64_t nc = h->c_rel_off.empty() ? 0 : h->c_rel_off.size() - 1;
    if (rel_off)
        for (uint64_t k = 0; k <= nc; ++k) rel_off[k] = nc ? (uint64_t)h->c_rel_off[k] : 0;
    if (fwd_off)
        for (uint64_t k = 0; k <= nc; ++k) fwd_off[k] = nc ? (uint64_t)h->c_fwd_off[k] : 0;
    const uint64_t nr = h->c_rel.size() / 7, nf = h->c_fwd.size() / 9;
    if (released && nr) {
        if (rel_cap < nr) return -ENOSPC;
        memcpy(released, h->c_rel.data(), 7 * sizeof(int64_t) * nr);
    }
    if (forwards && nf) {
        if (fwd_cap < nf) return -ENOSPC;
        memcpy(forwards, h->c_fwd.data(), 9 * sizeof(int64_t) * nf);
    }
    return 0;
}

namespace {

constexpr uint64_t NONE64 = ~0ull;

// device address of (x bytes, y frame column) in chunk c, placing the chunk in the arena on first use
int chunk_addr(xkrt_hip_mcc *h, uint32_t c, uint32_t x, uint32_t y, unsigned long long *addr, uint32_t *pitch)
{
    if (c >= h->chunk_rect.size()) return -EIO;
    const std::array<uint32_t, 4> &r = h->chunk_rect[c];
    const uint64_t w = (uint64_t)(r[1] - r[0]) * h->s, bytes = w * (r[3] - r[2]);
    if (h->chunk_off.size() <= c) h->chunk_off.resize(c + 1, NONE64);
    if (h->chunk_off[c] == NONE64) {
        const uint64_t off = (h->arena_used + 255) & ~255ull;
        if (off + bytes > h->arena_bytes) return -ENOMEM;
        h->chunk_off[c] = off;
        h->arena_used = off + bytes;
    }
    if (x < r[0] * h->s || y < r[2]) return -EIO;
    *addr = (unsigned long long)(uintptr_t)(h->arena + h->chunk_off[c]) + (unsigned long long)(y - r[2]) * w +
            (x - r[0] * h->s);
    *pitch = (uint32_t)w;
    return 0;
}

// the end of a fetch's rect must stay inside its chunk (a fetch that does not fit is refused)
bool rect_in_chunk(const xkrt_hip_mcc *h, uint32_t c, const MccFetch &f)
{
    if (c >= h->chunk_rect.size()) return false;
    const std::array<uint32_t, 4> &r = h->chunk_rect[c];
    return r[0] <= f.u0 && f.u1 <= r[1] && r[2] <= f.y0 && f.y1 <= r[3];
}

} // namespace

int xkrt_hip_mcc_exec_bind(xkrt_hip_mcc_t h, void *host_base, void *arena, uint64_t arena_bytes)
{
    if (!h || !host_base || (!arena && arena_bytes)) return -EINVAL;
    void *dev = nullptr;
    if (hipHostGetDevicePointer(&dev, host_base, 0) != hipSuccess || !dev) return -EINVAL;  // must be pinned
    h->host_dev = (unsigned char *)dev;
    h->arena = (unsigned char *)arena;
    h->arena_bytes = arena_bytes;
    h->arena_used = 0;
    h->chunk_off.clear();
    return 0;
}

int xkrt_hip_mcc_exec(xkrt_hip_mcc_t h, const uint64_t *fetch_ids, uint64_t n, void *stream)
{
    if (!h || (!fetch_ids && n)) return -EINVAL;
    if (!h->async || !h->host_dev) return -EINVAL;
    if (n == 0) return 0;
    // the job buffer used two calls ago must have been read
    const uint32_t b = h->jobs_cur;
    if (h->jobs_pending[b]) {
        CCHK(hipEventSynchronize(h->jobs_ev[b]));
        h->jobs_pending[b] = false;
    }
    if (h->jobs_cap[b] < n) {
        if (h->jobs[b]) CCHK(hipHostFree(h->jobs[b]));
        h->jobs[b] = nullptr;
        h->jobs_cap[b] = 0;
        CCHK(hipHostMalloc(&h->jobs[b], sizeof(CopyJob) * n, hipHostMallocCoherent));
        h->jobs_cap[b] = n;
    }
    if (!h->jobs_ev[b]) CCHK(hipEventCreateWithFlags(&h->jobs_ev[b], hipEventDisableTiming));
    for (uint64_t k = 0; k < n; ++k) {
        const uint64_t id = fetch_ids[k];
        if (id >= h->fetches.size() || !h->fetches[id].inflight) return -EINVAL;
        const MccFetch &f = h->fetches[id];
        const uint32_t x0 = f.u0 * (uint32_t)h->s;
        CopyJob &j = h->jobs[b][k];
        j.width = (f.u1 - f.u0) * (uint32_t)h->s;
        j.height = f.y1 - f.y0;
        auto end = [&](uint32_t dev, uint32_t chunk, unsigned long long *addr, uint32_t *pitch) -> int {
            if (dev == 0) {
                *addr = (unsigned long long)(uintptr_t)(h->host_dev + (uint64_t)f.y0 * h->LDs + x0);
                *pitch = (uint32_t)h->LDs;
                return 0;
            }
            if (!rect_in_chunk(h, chunk, f)) return -EIO;
            return chunk_addr(h, chunk, x0, f.y0, addr, pitch);
        };
        int rc = end(f.sd, f.sc, &j.src, &j.spitch);
        if (!rc) rc = end(f.dd, f.dc, &j.dst, &j.dpitch);
        if (rc) return rc;
    }
    CopyJob *dj = nullptr;
    CCHK(hipHostGetDevicePointer((void **)&dj, h->jobs[b], 0));
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(c_exec, dim3((uint32_t)std::min<uint64_t>(n, 8192)), dim3(256), 0, st, dj, (uint32_t)n);
    CCHK(hipGetLastError());
    CCHK(hipEventRecord(h->jobs_ev[b], st));
    h->jobs_pending[b] = true;
    h->jobs_cur = b ^ 1u;
    return 0;
}

int xkrt_hip_mcc_access_chunks(xkrt_hip_mcc_t h, int64_t *chunks)
{
    if (!h || !chunks) return -EINVAL;
    for (size_t i = 0; i < h->last_chunks.size(); ++i) chunks[i] = h->last_chunks[i];
    return 0;
}

int xkrt_hip_mcc_chunk_ptr(xkrt_hip_mcc_t h, uint32_t chunk, void **ptr, uint64_t *pitch, int64_t *rect)
{
    if (!h || !ptr || !h->arena) return -EINVAL;
    if (chunk >= h->chunk_rect.size()) return -EINVAL;
    const std::array<uint32_t, 4> &r = h->chunk_rect[chunk];
    unsigned long long a = 0;
    uint32_t p = 0;
    const int rc = chunk_addr(h, chunk, r[0] * (uint32_t)h->s, r[2], &a, &p);
    if (rc) return rc;
    *ptr = (void *)(uintptr_t)a;
    if (pitch) *pitch = p;
    if (rect) {
        rect[0] = (int64_t)r[0] * (int64_t)h->s; rect[1] = (int64_t)r[1] * (int64_t)h->s;
        rect[2] = (int64_t)(r[2] + h->cfg.y_lo); rect[3] = (int64_t)(r[3] + h->cfg.y_lo);
    }
    return 0;
}

int xkrt_hip_mcc_who_owns_static(xkrt_hip_mcc_t h, const xkrt_hip_mcc_access_t *acc, uint64_t n, uint32_t *owners)
{
    if (!h || (!acc && n) || (!owners && n)) return -EINVAL;
    if (h->poisoned) return -ECANCELED;  // the reference has aborted
    if (n == 0) return 0;
    if (n > WHO_OWNS_STATIC_MAX) return -E2BIG;
    for (uint64_t i = 0; i < n; ++i) {
        const int rc = owner_query(h, acc[i]);
        if (rc) return rc;
    }
    unsigned char *base = (unsigned char *)scratch_of(h->device);
    if (!base) return -ENODEV;
    xkrt_hip_mcc_access_t *dacc = (xkrt_hip_mcc_access_t *)base;
    uint32_t *dq = (uint32_t *)(base + OWN_ACC_BYTES);
    uint32_t *dout = (uint32_t *)(base + OWN_ACC_BYTES + OWN_Q_BYTES);
    const uint32_t nq = (uint32_t)n;
    CCHK(hipMemcpyAsync(dacc, acc, sizeof(xkrt_hip_mcc_access_t) * n, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(c_owners_prep, dim3((nq + 255) / 256), dim3(256), 0, h->stream, dacc, nq, h->LDs,
                       (uint32_t)h->s, h->cfg.y_lo, dq);
    CCHK(hipGetLastError());
    CArgs a = make_cargs(h, 0);
    hipLaunchKernelGGL(c_owners, dim3(std::min<uint32_t>(nq, 4096)), dim3(256), 0, h->stream, a, dq, dq + n,
                       dq + 2 * n, dq + 3 * n, nq, dout);
    CCHK(hipGetLastError());
    CCHK(hipMemcpyAsync(owners, dout, 4ull * n, hipMemcpyDeviceToHost, h->stream));
    return cerr(hipStreamSynchronize(h->stream));
}

int xkrt_hip_mcc_who_owns(xkrt_hip_mcc_t h, const xkrt_hip_mcc_access_t *acc, uint64_t n, uint32_t *owners)
{
    if (!h || (!acc && n) || (!owners && n)) return -EINVAL;
    if (h->poisoned) return -ECANCELED;  // the reference has aborted
    if (n == 0) return 0;
    std::vector<uint32_t> q(4 * n);
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t xb = acc[i].addr % h->LDs, col = acc[i].addr / h->LDs;
        if (acc[i].m == 0 || acc[i].n == 0 || xb % h->s) return -EINVAL;
        if (xb + acc[i].m * h->s > h->LDs) return -ENOTSUP;
        if (col < h->cfg.y_lo || col + acc[i].n > h->cfg.y_hi) return -ERANGE;
        q[i] = (uint32_t)(xb / h->s);
        q[n + i] = (uint32_t)(xb / h->s + acc[i].m);
        q[2 * n + i] = (uint32_t)(col - h->cfg.y_lo);
        q[3 * n + i] = (uint32_t)(col - h->cfg.y_lo + acc[i].n);
    }
    uint32_t *dq = nullptr, *dout = nullptr;
    CCHK(hipMalloc(&dq, 4ull * 4 * n));
    int rc = cerr(hipMalloc(&dout, 4ull * n));
    if (!rc) rc = cerr(hipMemcpyAsync(dq, q.data(), 16ull * n, hipMemcpyHostToDevice, h->stream));
    if (!rc) {
        CArgs a = make_cargs(h, 0);
        hipLaunchKernelGGL(c_owners, dim3((uint32_t)std::min<uint64_t>(n, 4096)), dim3(256), 0, h->stream, a, dq,
                           dq + n, dq + 2 * n, dq + 3 * n, (uint32_t)n, dout);
        rc = cerr(hipGetLastError());
    }
    if (!rc) rc = cerr(hipMemcpyAsync(owners, dout, 4ull * n, hipMemcpyDeviceToHost, h->stream));
    if (!rc) rc = cerr(hipStreamSynchronize(h->stream));
    (void)hipFree(dq);
    (void)hipFree(dout);
    return rc;
}

int xkrt_hip_mcc_stats(xkrt_hip_mcc_t h, xkrt_hip_mcc_stats_t *out)
{
    if (!h || !out) return -EINVAL;
    memset(out, 0, sizeof(*out));
    out->accesses = h->naccess;
    out->gpu_accesses = h->nhist;
    CCHK(hipMemcpyAsync(h->ctr_host, h->ctr, sizeof(CCounters), hipMemcpyDeviceToHost, h->stream));
    CCHK(hipStreamSynchronize(h->stream));
    out->chunks = h->ctr_host->nchunks;
    out->last_parts = h->last_rows;
    out->last_fetches = h->last_fetch;
    out->last_ms = h->last_ms;
    out->device_bytes = h->device_bytes;
    out->poisoned = h->poisoned ? 1 : 0;
    out->last_fetch_base = h->last_fetch_base;
    out->in_flight = h->inflight;
    out->registrations = h->live.size();
    return 0;
}

} // extern "C"
