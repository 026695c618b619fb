/*
 * Deterministic task/access stream generators for the dependency resolver and
 * the coherency planner.
 *
 * Every stream is produced in spawn order as structure-of-arrays records of
 * column-major matrix tile accesses, exactly the arguments the reference's
 * BLAS access constructor takes (include/xkrt/memory/access/access.hpp:539-574:
 * addr, ld, m, n, sizeof_type, mode), plus the spawning task index, the slot
 * of the access inside its task and, for coherency streams, the device that
 * fetches it (0 = host, 1 = the GPU).
 *
 * Randomness comes from SplitMix64 only, with explicit integer transforms, so
 * the streams are identical whatever the C library's <random> does.
 *
 * Calling convention: each generator takes a capacity `cap` and output arrays.
 * With any output array NULL it only counts and returns the number of
 * accesses.  Otherwise it writes min(count, cap) records and returns count.
 */
#define _POSIX_C_SOURCE 199309L  /* clock_gettime */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <time.h>

#define MODE_R  1u
#define MODE_W  2u
#define MODE_RW 3u

typedef struct {
    uint32_t *task;
    uint16_t *slot;
    uint8_t  *mode;
    uint8_t  *dev;
    uint64_t *addr;
    uint64_t *ld;
    uint64_t *m;
    uint64_t *n;
    uint64_t *s;
    uint64_t  cap;
    uint64_t  count;
    int       fill;
} sink_t;

static void sink_init(sink_t *k, uint64_t cap, uint32_t *task, uint16_t *slot, uint8_t *mode,
                      uint8_t *dev, uint64_t *addr, uint64_t *ld, uint64_t *m, uint64_t *n, uint64_t *s)
{
    k->task = task; k->slot = slot; k->mode = mode; k->dev = dev;
    k->addr = addr; k->ld = ld; k->m = m; k->n = n; k->s = s;
    k->cap = cap; k->count = 0;
    k->fill = task && slot && mode && addr && ld && m && n && s;
}

static void emit(sink_t *k, uint32_t task, uint16_t slot, uint8_t mode, uint8_t dev,
                 uint64_t addr, uint64_t ld, uint64_t m, uint64_t n, uint64_t s)
{
    if (k->fill && k->count < k->cap) {
        uint64_t i = k->count;
        k->task[i] = task; k->slot[i] = slot; k->mode[i] = mode;
        if (k->dev) k->dev[i] = dev;
        k->addr[i] = addr; k->ld[i] = ld; k->m[i] = m; k->n[i] = n; k->s[i] = s;
    }
    k->count++;
}

/* SplitMix64 (Steele, Lea, Flood 2014). */
static inline uint64_t sm64(uint64_t *state)
{
    uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* uniform double in [0, 1) from the top 53 bits */
static inline double sm64_unit(uint64_t *state)
{
    return (double)(sm64(state) >> 11) * (1.0 / 9007199254740992.0);
}

/* log-uniform integer in [1, maxv]: floor(2^(u * log2(maxv + 1))) */
static inline uint64_t log_uniform(uint64_t *state, uint64_t maxv)
{
    double u = sm64_unit(state);
    uint64_t v = (uint64_t)floor(exp2(u * log2((double)maxv + 1.0)));
    if (v < 1) v = 1;
    if (v > maxv) v = maxv;
    return v;
}

uint64_t xks_splitmix64_next(uint64_t *state) { return sm64(state); }

/* address of tile (i, j) of a column-major matrix at `base`, tile size nb */
static inline uint64_t tile_addr(uint64_t base, uint64_t ld, uint64_t s, uint64_t i, uint64_t j, uint64_t nb)
{
    return base + (j * nb * ld + i * nb) * s;
}

/*
 * Tiled right-looking Cholesky on a T x T tile grid (T = N / nb), loop order
 *   for k: POTRF(RW Akk)
 *          for i > k: TRSM(R Akk, RW Aik)
 *          for i > k: SYRK(R Aik, RW Aii); for k < j < i: GEMM(R Aik, R Ajk, RW Aij)
 * Task indices start at task0.
 */
uint64_t xks_cholesky(uint64_t N, uint64_t nb, uint64_t s, uint64_t base, uint32_t task0,
                      uint64_t cap, uint32_t *task, uint16_t *slot, uint8_t *mode, uint8_t *dev,
                      uint64_t *addr, uint64_t *ld, uint64_t *m, uint64_t *n, uint64_t *sz)
{
    sink_t k; sink_init(&k, cap, task, slot, mode, dev, addr, ld, m, n, sz);
    const uint64_t T = N / nb, LD = N;
    uint32_t t = task0;
    for (uint64_t kk = 0; kk < T; ++kk) {
        emit(&k, t, 0, MODE_RW, 1, tile_addr(base, LD, s, kk, kk, nb), LD, nb, nb, s); t++;
        for (uint64_t i = kk + 1; i < T; ++i) {
            emit(&k, t, 0, MODE_R,  1, tile_addr(base, LD, s, kk, kk, nb), LD, nb, nb, s);
            emit(&k, t, 1, MODE_RW, 1, tile_addr(base, LD, s, i,  kk, nb), LD, nb, nb, s); t++;
        }
        for (uint64_t i = kk + 1; i < T; ++i) {
            emit(&k, t, 0, MODE_R,  1, tile_addr(base, LD, s, i, kk, nb), LD, nb, nb, s);
            emit(&k, t, 1, MODE_RW, 1, tile_addr(base, LD, s, i, i,  nb), LD, nb, nb, s); t++;
            for (uint64_t j = kk + 1; j < i; ++j) {
                emit(&k, t, 0, MODE_R,  1, tile_addr(base, LD, s, i, kk, nb), LD, nb, nb, s);
                emit(&k, t, 1, MODE_R,  1, tile_addr(base, LD, s, j, kk, nb), LD, nb, nb, s);
                emit(&k, t, 2, MODE_RW, 1, tile_addr(base, LD, s, i, j,  nb), LD, nb, nb, s); t++;
            }
        }
    }
    return k.count;
}

/*
 * Tiled right-looking LU without pivoting, loop order
 *   for k: GETRF(RW Akk)
 *          for j > k: TRSM-L(R Akk, RW Akj)
 *          for i > k: TRSM-U(R Akk, RW Aik)
 *          for i > k, j > k: GEMM(R Aik, R Akj, RW Aij)
 */
uint64_t xks_lu(uint64_t N, uint64_t nb, uint64_t s, uint64_t base, uint32_t task0,
                uint64_t cap, uint32_t *task, uint16_t *slot, uint8_t *mode, uint8_t *dev,
                uint64_t *addr, uint64_t *ld, uint64_t *m, uint64_t *n, uint64_t *sz)
{
    sink_t k; sink_init(&k, cap, task, slot, mode, dev, addr, ld, m, n, sz);
    const uint64_t T = N / nb, LD = N;
    uint32_t t = task0;
    for (uint64_t kk = 0; kk < T; ++kk) {
        emit(&k, t, 0, MODE_RW, 1, tile_addr(base, LD, s, kk, kk, nb), LD, nb, nb, s); t++;
        for (uint64_t j = kk + 1; j < T; ++j) {
            emit(&k, t, 0, MODE_R,  1, tile_addr(base, LD, s, kk, kk, nb), LD, nb, nb, s);
            emit(&k, t, 1, MODE_RW, 1, tile_addr(base, LD, s, kk, j,  nb), LD, nb, nb, s); t++;
        }
        for (uint64_t i = kk + 1; i < T; ++i) {
            emit(&k, t, 0, MODE_R,  1, tile_addr(base, LD, s, kk, kk, nb), LD, nb, nb, s);
            emit(&k, t, 1, MODE_RW, 1, tile_addr(base, LD, s, i,  kk, nb), LD, nb, nb, s); t++;
        }
        for (uint64_t i = kk + 1; i < T; ++i)
            for (uint64_t j = kk + 1; j < T; ++j) {
                emit(&k, t, 0, MODE_R,  1, tile_addr(base, LD, s, i,  kk, nb), LD, nb, nb, s);
                emit(&k, t, 1, MODE_R,  1, tile_addr(base, LD, s, kk, j,  nb), LD, nb, nb, s);
                emit(&k, t, 2, MODE_RW, 1, tile_addr(base, LD, s, i,  j,  nb), LD, nb, nb, s); t++;
            }
    }
    return k.count;
}

/*
 * Seeded random single-access tasks over a rows x cols element space of a
 * column-major buffer (leading dimension LD = rows, element size s).  Per
 * access, in this draw order: row origin (uniform), col origin (uniform),
 * height m and width n (log-uniform in [1, maxsz], clipped to the space),
 * mode (u < pr: R, u < pr + pw: W, else RW).
 */
uint64_t xks_random2d(uint64_t nacc, uint64_t rows, uint64_t cols, uint64_t s, uint64_t maxsz,
                      uint64_t seed, double pr, double pw, uint64_t base,
                      uint64_t cap, uint32_t *task, uint16_t *slot, uint8_t *mode, uint8_t *dev,
                      uint64_t *addr, uint64_t *ld, uint64_t *m, uint64_t *n, uint64_t *sz)
{
    sink_t k; sink_init(&k, cap, task, slot, mode, dev, addr, ld, m, n, sz);
    uint64_t st = seed;
    for (uint64_t a = 0; a < nacc; ++a) {
        uint64_t r0 = sm64(&st) % rows;
        uint64_t c0 = sm64(&st) % cols;
        uint64_t mm = log_uniform(&st, maxsz);
        uint64_t nn = log_uniform(&st, maxsz);
        double u = sm64_unit(&st);
        if (mm > rows - r0) mm = rows - r0;
        if (nn > cols - c0) nn = cols - c0;
        uint8_t md = u < pr ? MODE_R : (u < pr + pw ? MODE_W : MODE_RW);
        emit(&k, (uint32_t)a, 0, md, 1, base + (c0 * rows + r0) * s, rows, mm, nn, s);
    }
    return k.count;
}

/*
 * 2D heat-diffusion halo stream: an Ng x Ng grid of element size s (LD = Ng)
 * in tile x tile tiles.  Each step spawns, in row-major tile order, one task
 * per tile: RW own tile, then R of the 1-element-wide halo strips of the
 * existing neighbours in the order north (row r0-1), south (row r0+tile),
 * west (col c0-1), east (col c0+tile).
 */
uint64_t xks_heat(uint64_t Ng, uint64_t tile, uint64_t s, uint64_t steps, uint64_t base, uint32_t task0,
                  uint64_t cap, uint32_t *task, uint16_t *slot, uint8_t *mode, uint8_t *dev,
                  uint64_t *addr, uint64_t *ld, uint64_t *m, uint64_t *n, uint64_t *sz)
{
    sink_t k; sink_init(&k, cap, task, slot, mode, dev, addr, ld, m, n, sz);
    const uint64_t T = Ng / tile, LD = Ng;
    uint32_t t = task0;
    for (uint64_t st = 0; st < steps; ++st)
        for (uint64_t ti = 0; ti < T; ++ti)
            for (uint64_t tj = 0; tj < T; ++tj) {
                const uint64_t r0 = ti * tile, c0 = tj * tile;
                uint16_t sl = 0;
                emit(&k, t, sl++, MODE_RW, 1, base + (c0 * LD + r0) * s, LD, tile, tile, s);
                if (ti > 0)     emit(&k, t, sl++, MODE_R, 1, base + (c0 * LD + r0 - 1) * s,    LD, 1, tile, s);
                if (ti + 1 < T) emit(&k, t, sl++, MODE_R, 1, base + (c0 * LD + r0 + tile) * s, LD, 1, tile, s);
                if (tj > 0)     emit(&k, t, sl++, MODE_R, 1, base + ((c0 - 1) * LD + r0) * s,  LD, tile, 1, s);
                if (tj + 1 < T) emit(&k, t, sl++, MODE_R, 1, base + ((c0 + tile) * LD + r0) * s, LD, tile, 1, s);
                t++;
            }
    return k.count;
}

/*
 * Coherency stream over one byte buffer seen through two column-major views:
 * view 0 (LD = ld0, element s0) and view 1 (LD = ld1, element s1), both
 * spanning `cols` columns.  Per access, in this draw order: view (u < 0.5:
 * view 0), row origin, col origin (uniform in the view's element space),
 * height and width (log-uniform in [1, maxsz], clipped), device (u < pdev_host:
 * host 0, else GPU 1), mode (u < pr: R, u < pr + pw: W, else RW).  One task per
 * access.
 */
uint64_t xks_coherency(uint64_t nacc, uint64_t ld0, uint64_t s0, uint64_t ld1, uint64_t s1, uint64_t cols,
                       uint64_t maxsz, uint64_t seed, double pdev_host, double pr, double pw, uint64_t base,
                       uint64_t cap, uint32_t *task, uint16_t *slot, uint8_t *mode, uint8_t *dev,
                       uint64_t *addr, uint64_t *ld, uint64_t *m, uint64_t *n, uint64_t *sz)
{
    sink_t k; sink_init(&k, cap, task, slot, mode, dev, addr, ld, m, n, sz);
    uint64_t st = seed;
    for (uint64_t a = 0; a < nacc; ++a) {
        int v = sm64_unit(&st) < 0.5 ? 0 : 1;
        const uint64_t LD = v ? ld1 : ld0, s = v ? s1 : s0;
        uint64_t r0 = sm64(&st) % LD;
        uint64_t c0 = sm64(&st) % cols;
        uint64_t mm = log_uniform(&st, maxsz);
        uint64_t nn = log_uniform(&st, maxsz);
        uint8_t dv = sm64_unit(&st) < pdev_host ? 0 : 1;
        double u = sm64_unit(&st);
        if (mm > LD - r0) mm = LD - r0;
        if (nn > cols - c0) nn = cols - c0;
        uint8_t md = u < pr ? MODE_R : (u < pr + pw ? MODE_W : MODE_RW);
        emit(&k, (uint32_t)a, 0, md, dv, base + (c0 * LD + r0) * s, LD, mm, nn, s);
    }
    return k.count;
}

/*
 * Batch boundaries: sizes from a geometric distribution with the given mean
 * (support 1, 2, ...), capped at maxb, until `total` items are covered.  Writes
 * at most cap sizes; returns the number of batches.
 */
uint64_t xks_geometric_batches(uint64_t total, double mean, uint64_t maxb, uint64_t seed,
                               uint64_t cap, uint64_t *sizes)
{
    uint64_t st = seed, covered = 0, nb = 0;
    const double p = 1.0 / mean;
    const double lq = log1p(-p);
    while (covered < total) {
        double u = sm64_unit(&st);
        /* inverse CDF of the geometric law on {1, 2, ...}: 1 + floor(ln(1-u) / ln(1-p)) */
        double g = floor(log1p(-u) / lq);
        uint64_t b = (g >= (double)maxb) ? maxb : 1 + (uint64_t)g;
        if (b > maxb) b = maxb;
        if (b > total - covered) b = total - covered;
        if (sizes && nb < cap) sizes[nb] = b;
        covered += b;
        nb++;
    }
    return nb;
}

/*
 * Serving-loop clients: per client, the factorisation kind (0 = Cholesky,
 * 1 = LU, 50/50) and the tile count in [nmin, nmax], in this draw order.
 */
void xks_clients(uint64_t nclients, uint64_t nmin, uint64_t nmax, uint64_t seed, uint8_t *kind, uint32_t *ntiles)
{
    uint64_t st = seed;
    for (uint64_t c = 0; c < nclients; ++c) {
        kind[c] = sm64_unit(&st) < 0.5 ? 0 : 1;
        ntiles[c] = (uint32_t)(nmin + sm64(&st) % (nmax - nmin + 1));
    }
}

/*
 * Column-major tile -> rects in (byte-offset-in-column, column) space,
 * following include/xkrt/memory/access/access.hpp:205-287: with
 * A the begin address and LDs = ld * s, one rect [A % LDs, +m*s) x [A / LDs, +n)
 * when the tile does not wrap, otherwise [A % LDs, LDs) x [y0, y0+n) and
 * [0, m*s - (LDs - A % LDs)) x [y0+1, y0+n+1).  rects receives 8 values per
 * access (x0, x1, y0, y1 of rect 0 then rect 1; rect 1 is empty (0,0,0,0)
 * for single-rect tiles); nrects receives 1 or 2.
 */
void xks_matrix_to_rects(uint64_t count, const uint64_t *addr, const uint64_t *ld, const uint64_t *m,
                         const uint64_t *n, const uint64_t *s, uint64_t *rects, uint8_t *nrects)
{
    for (uint64_t i = 0; i < count; ++i) {
        const uint64_t LDs = ld[i] * s[i];
        const uint64_t A = addr[i];
        const uint64_t x0 = A % LDs, y0 = A / LDs;
        uint64_t *r = rects + 8 * i;
        if (x0 + m[i] * s[i] <= LDs) {
            r[0] = x0; r[1] = x0 + m[i] * s[i]; r[2] = y0; r[3] = y0 + n[i];
            r[4] = r[5] = r[6] = r[7] = 0;
            nrects[i] = 1;
        } else {
            r[0] = x0; r[1] = LDs; r[2] = y0; r[3] = y0 + n[i];
            r[4] = 0; r[5] = m[i] * s[i] - (LDs - x0); r[6] = y0 + 1; r[7] = y0 + n[i] + 1;
            nrects[i] = 2;
        }
    }
}

/*
 * Open-loop submission for the serving workload: batch i is handed to
 * submit(srv, client[i], acc[i], nacc[i], &req[i]) at t_s[i] seconds after the
 * call starts (a spin on CLOCK_MONOTONIC), and close_client(srv, client[i])
 * follows when last[i] is set.  lag_s[i] = how late the submit call started.
 * The server's entry points come in as function pointers, so this file links
 * against nothing.  Returns 0, or the first nonzero code of a call (the loop
 * stops there).
 */
typedef int (*xks_submit_fn)(void *srv, uint32_t client, const void *acc, uint64_t n, uint64_t *request);
typedef int (*xks_close_fn)(void *srv, uint32_t client);

static double xks_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int xks_open_loop(void *srv, xks_submit_fn submit, xks_close_fn close_client, uint64_t n, const double *t_s,
                  const uint32_t *client, const void *const *acc, const uint64_t *nacc, const uint8_t *last,
                  double *lag_s, uint64_t *req)
{
    const double t0 = xks_now();
    for (uint64_t i = 0; i < n; ++i) {
        double now = xks_now() - t0;
        while (now < t_s[i]) now = xks_now() - t0;
        lag_s[i] = now - t_s[i];
        int rc = submit(srv, client[i], acc[i], nacc[i], &req[i]);
        if (rc) return rc;
        if (last[i] && (rc = close_client(srv, client[i])) != 0) return rc;
    }
    return 0;
}
