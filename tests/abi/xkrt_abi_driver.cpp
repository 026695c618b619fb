// A compiled C++ caller of the C ABI (include/xkrt_hip.h): what a runtime
// linked against libxkrt_hip.so does, with no Python in between.  Built by
// xkaapi_amd/build.py; tests/test_gpu_abi_driver.py runs it on the GPU and
// checks what it prints against the oracle.
//
//   1. one dependency domain: a tiled Cholesky (4096^2 fp64, 256^2 tiles)
//      submitted in batches of whole tasks, sync, edges copied; an empty tile
//      is refused with -EINVAL;
//   2. a pool with two domains (Cholesky and LU, 1024^2 fp64, 128^2 tiles)
//      staged together and resolved by one launch, edges with their domain;
//   3. a coherency tree: who_owns and who_owns_static agree;
//   4. a task domain over a mixed program (program() below: tiles of two
//      (ld, sizeof_type) pairs and byte segments), submitted in batches.
//
// Output lines: "chol <pred> <succ>", "pool <domain> <pred> <succ>",
// "domain <name> <index>", "who_owns_equal <0|1>", "prog <pred> <succ>",
// "ok".  Exit status 0 on success.

#include "xkrt_hip.h"

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <utility>
#include <vector>

#define CHECK(x)                                                                        \
    do {                                                                                \
        const int rc_ = (x);                                                            \
        if (rc_) {                                                                      \
            std::fprintf(stderr, "%s:%d: %s returned %d\n", __FILE__, __LINE__, #x, rc_); \
            return 1;                                                                   \
        }                                                                               \
    } while (0)

namespace {

using Acc = xkrt_hip_access_t;

Acc tile(uint64_t base, uint64_t ld, uint64_t s, uint64_t i, uint64_t j, uint64_t nb, uint32_t task, uint8_t mode)
{
    Acc a{};
    a.addr = base + (j * nb * ld + i * nb) * s;  // column-major tile (i, j)
    a.m = nb;
    a.n = nb;
    a.task = task;
    a.mode = mode;
    a.kind = XKRT_HIP_KIND_TILE;
    return a;
}

// tiled right-looking Cholesky: POTRF(RW Akk), TRSM(R Akk, RW Aik), SYRK(R Aik, RW Aii), GEMM(R Aik, R Ajk, RW Aij)
std::vector<Acc> cholesky(uint64_t N, uint64_t nb, uint64_t s)
{
    std::vector<Acc> v;
    const uint64_t T = N / nb;
    uint32_t t = 0;
    for (uint64_t k = 0; k < T; ++k) {
        v.push_back(tile(0, N, s, k, k, nb, t++, XKRT_HIP_MODE_RW));
        for (uint64_t i = k + 1; i < T; ++i) {
            v.push_back(tile(0, N, s, k, k, nb, t, XKRT_HIP_MODE_R));
            v.push_back(tile(0, N, s, i, k, nb, t++, XKRT_HIP_MODE_RW));
        }
        for (uint64_t i = k + 1; i < T; ++i) {
            v.push_back(tile(0, N, s, i, k, nb, t, XKRT_HIP_MODE_R));
            v.push_back(tile(0, N, s, i, i, nb, t++, XKRT_HIP_MODE_RW));
            for (uint64_t j = k + 1; j < i; ++j) {
                v.push_back(tile(0, N, s, i, k, nb, t, XKRT_HIP_MODE_R));
                v.push_back(tile(0, N, s, j, k, nb, t, XKRT_HIP_MODE_R));
                v.push_back(tile(0, N, s, i, j, nb, t++, XKRT_HIP_MODE_RW));
            }
        }
    }
    return v;
}

// tiled right-looking LU without pivoting: GETRF(RW Akk), TRSM-L(R Akk, RW Akj), TRSM-U(R Akk, RW Aik),
// GEMM(R Aik, R Akj, RW Aij)
std::vector<Acc> lu(uint64_t N, uint64_t nb, uint64_t s)
{
    std::vector<Acc> v;
    const uint64_t T = N / nb;
    uint32_t t = 0;
    for (uint64_t k = 0; k < T; ++k) {
        v.push_back(tile(0, N, s, k, k, nb, t++, XKRT_HIP_MODE_RW));
        for (uint64_t j = k + 1; j < T; ++j) {
            v.push_back(tile(0, N, s, k, k, nb, t, XKRT_HIP_MODE_R));
            v.push_back(tile(0, N, s, k, j, nb, t++, XKRT_HIP_MODE_RW));
        }
        for (uint64_t i = k + 1; i < T; ++i) {
            v.push_back(tile(0, N, s, k, k, nb, t, XKRT_HIP_MODE_R));
            v.push_back(tile(0, N, s, i, k, nb, t++, XKRT_HIP_MODE_RW));
        }
        for (uint64_t i = k + 1; i < T; ++i)
            for (uint64_t j = k + 1; j < T; ++j) {
                v.push_back(tile(0, N, s, i, k, nb, t, XKRT_HIP_MODE_R));
                v.push_back(tile(0, N, s, k, j, nb, t, XKRT_HIP_MODE_R));
                v.push_back(tile(0, N, s, i, j, nb, t++, XKRT_HIP_MODE_RW));
            }
    }
    return v;
}

// [begin, end) access ranges of whole tasks, about `tasks` tasks each
std::vector<std::pair<size_t, size_t>> batches(const std::vector<Acc> &v, uint32_t tasks)
{
    std::vector<std::pair<size_t, size_t>> out;
    size_t a = 0;
    while (a < v.size()) {
        size_t b = a;
        const uint32_t t0 = v[a].task;
        while (b < v.size() && v[b].task < t0 + tasks) ++b;
        out.push_back({a, b});
        a = b;
    }
    return out;
}

int run_domain()
{
    const std::vector<Acc> acc = cholesky(4096, 256, 8);
    xkrt_hip_dep_config_t cfg{};
    cfg.bin_x_bytes = 2048;
    cfg.bin_y_cols = 256;
    cfg.kernel = 1;
    xkrt_hip_dep_domain_t d = nullptr;
    CHECK(xkrt_hip_dep_domain_create(4096, 8, &cfg, nullptr, &d));
    uint64_t expect_id = 0;
    for (auto [a, b] : batches(acc, 100)) {
        uint64_t first = 0;
        CHECK(xkrt_hip_dep_submit(d, acc.data() + a, b - a, &first));
        if (first != expect_id) {
            std::fprintf(stderr, "access ids: got %llu, want %llu\n", (unsigned long long)first,
                         (unsigned long long)expect_id);
            return 1;
        }
        expect_id += b - a;
    }
    uint64_t n = 0;
    CHECK(xkrt_hip_dep_sync(d, &n));
    std::vector<uint32_t> e(2 * n + 2);
    uint64_t got = 0;
    CHECK(xkrt_hip_dep_edges_copy(d, e.data(), n, &got));
    if (got != n) return 1;
    std::vector<std::pair<uint32_t, uint32_t>> pairs(n);
    for (uint64_t i = 0; i < n; ++i) pairs[i] = {e[2 * i], e[2 * i + 1]};
    std::sort(pairs.begin(), pairs.end());
    for (auto &p : pairs) std::printf("chol %u %u\n", p.first, p.second);

    // an empty tile is refused and nothing of its batch is enqueued
    Acc bad = acc[0];
    bad.m = 0;
    bad.task = acc.back().task + 1;
    if (xkrt_hip_dep_submit(d, &bad, 1, nullptr) != -EINVAL) {
        std::fprintf(stderr, "an empty tile must be refused with -EINVAL\n");
        return 1;
    }
    xkrt_hip_dep_stats_t st{};
    CHECK(xkrt_hip_dep_stats(d, &st));
    if (st.accesses != acc.size() || st.err != 0) {
        std::fprintf(stderr, "stats: %llu accesses, err %llu\n", (unsigned long long)st.accesses,
                     (unsigned long long)st.err);
        return 1;
    }
    CHECK(xkrt_hip_dep_domain_destroy(d));
    return 0;
}

int run_pool()
{
    xkrt_hip_pool_config_t pc{};
    pc.kernel = 1;
    pc.max_rects = 1 << 16;
    xkrt_hip_pool_t p = nullptr;
    CHECK(xkrt_hip_pool_create(&pc, nullptr, &p));
    const std::vector<Acc> ch = cholesky(1024, 128, 8), l = lu(1024, 128, 8);
    xkrt_hip_dep_domain_t dc = nullptr, dl = nullptr;
    CHECK(xkrt_hip_pool_domain_create(p, 1024, 8, 1024, 128, 0, 1024, &dc));
    CHECK(xkrt_hip_pool_domain_create(p, 1024, 8, 1024, 128, 0, 1024, &dl));
    uint32_t ic = 0, il = 0;
    CHECK(xkrt_hip_dep_index(dc, &ic));
    CHECK(xkrt_hip_dep_index(dl, &il));
    std::printf("domain cholesky %u\ndomain lu %u\n", ic, il);
    uint64_t f0 = 0, f1 = 0, ticket = 0;
    CHECK(xkrt_hip_pool_stage(p, dc, ch.data(), ch.size(), &f0));
    CHECK(xkrt_hip_pool_stage(p, dl, l.data(), l.size(), &f1));
    CHECK(xkrt_hip_pool_launch(p, &ticket));
    CHECK(xkrt_hip_pool_wait(p, ticket));
    uint64_t n = 0;
    CHECK(xkrt_hip_pool_sync(p, &n));
    std::vector<uint32_t> e(2 * n + 2), dom(n + 1);
    uint64_t got = 0;
    CHECK(xkrt_hip_pool_edges_copy(p, e.data(), dom.data(), n, &got));
    for (uint64_t i = 0; i < got; ++i) std::printf("pool %u %u %u\n", dom[i], e[2 * i], e[2 * i + 1]);
    CHECK(xkrt_hip_dep_domain_destroy(dc));
    CHECK(xkrt_hip_dep_domain_destroy(dl));
    CHECK(xkrt_hip_pool_destroy(p));
    return 0;
}

int run_coherency()
{
    xkrt_hip_mcc_config_t mc{};
    mc.y_lo = 0;
    mc.y_hi = 512;
    xkrt_hip_mcc_t h = nullptr;
    CHECK(xkrt_hip_mcc_create(512, 8, &mc, nullptr, &h));
    // GPU reads and writes of 64^2 tiles, then host reads: coherency spread over both devices
    std::vector<xkrt_hip_mcc_access_t> acc;
    for (uint64_t k = 0; k < 48; ++k) {
        xkrt_hip_mcc_access_t a{};
        const uint64_t i = (k * 5) % 8, j = (k * 3) % 8;
        a.addr = (j * 64 * 512 + i * 64) * 8;
        a.m = 64;
        a.n = 64;
        a.dev = k % 3 == 2 ? 0 : 1;
        a.mode = k % 4 == 0 ? XKRT_HIP_MODE_RW : XKRT_HIP_MODE_R;
        acc.push_back(a);
    }
    uint64_t np = 0, nf = 0;
    int64_t fail = 0;
    CHECK(xkrt_hip_mcc_plan(h, acc.data(), acc.size(), &np, &nf, &fail));
    if (fail != -1) return 1;
    std::vector<xkrt_hip_mcc_access_t> q;
    for (uint64_t k = 0; k < 64; ++k) {
        xkrt_hip_mcc_access_t a{};
        a.addr = ((k * 7 % 400) * 512 + (k * 13 % 400)) * 8;
        a.m = 1 + k % 100;
        a.n = 1 + k % 90;
        q.push_back(a);
    }
    std::vector<uint32_t> o1(q.size()), o2(q.size());
    CHECK(xkrt_hip_mcc_who_owns(h, q.data(), q.size(), o1.data()));
    CHECK(xkrt_hip_mcc_who_owns_static(h, q.data(), q.size(), o2.data()));
    std::printf("who_owns_equal %d\n", o1 == o2 ? 1 : 0);
    CHECK(xkrt_hip_mcc_destroy(h));
    return o1 == o2 ? 0 : 1;
}

// A mixed program of one task domain, generated the same way by the test:
// 160 tasks over a 256 x 256 fp64 matrix at 1 MiB (32 x 32 tiles) viewed
// with two leading dimensions (256 x 8 B and 512 x 4 B: two BLAS domains) and
// byte segments over the same bytes.
std::vector<xkrt_hip_task_access_t> program()
{
    std::vector<xkrt_hip_task_access_t> v;
    const uint64_t base = 1u << 20;
    for (uint32_t t = 0; t < 160; ++t) {
        xkrt_hip_task_access_t a{};
        a.task = t;
        const uint32_t k = t % 64, i = k % 8, j = (k * 3 + t / 64) % 8;
        if (t % 5 == 4) {  // a segment: [base + 4096 t mod 512 KiB, + 5000)
            a.kind = XKRT_HIP_KIND_SEGMENT;
            a.addr = base + (4096ull * t) % (512u << 10);
            a.m = 5000;
            a.mode = t % 2 ? XKRT_HIP_MODE_R : XKRT_HIP_MODE_W;
            v.push_back(a);
            continue;
        }
        a.kind = XKRT_HIP_KIND_TILE;
        if (t % 3 == 0) {  // tile (i, j) of the 512 x 4 B view: 64 rows of 4 B
            a.ld = 512;
            a.sizeof_type = 4;
            a.addr = base + (j * 32ull * 512 + i * 64) * 4;
            a.m = 64;
        } else {            // tile (i, j) of the 256 x 8 B view
            a.ld = 256;
            a.sizeof_type = 8;
            a.addr = base + (j * 32ull * 256 + i * 32) * 8;
            a.m = 32;
        }
        a.n = 32;
        a.mode = XKRT_HIP_MODE_RW;
        v.push_back(a);
        if (t % 7 == 0) {  // and a segment read of its first column
            xkrt_hip_task_access_t r = a;
            r.kind = XKRT_HIP_KIND_SEGMENT;
            r.m = 256;
            r.mode = XKRT_HIP_MODE_R;
            v.push_back(r);
        }
    }
    return v;
}

int run_taskdom()
{
    xkrt_hip_taskdom_config_t c{};
    c.kernel = 1;
    c.seg_col_bytes = 4096;
    c.seg_bin_x_bytes = 1024;
    c.seg_bin_y_cols = 1;
    c.blas_bin_x_bytes = 512;
    c.blas_bin_y_cols = 16;
    xkrt_hip_taskdom_t t = nullptr;
    CHECK(xkrt_hip_taskdom_create(&c, nullptr, &t));
    const std::vector<xkrt_hip_task_access_t> acc = program();
    size_t a = 0;
    while (a < acc.size()) {  // batches of 16 tasks
        size_t b = a;
        while (b < acc.size() && acc[b].task < acc[a].task + 16) ++b;
        uint64_t first = 0;
        CHECK(xkrt_hip_taskdom_submit(t, acc.data() + a, b - a, &first));
        if (first != a) return 1;
        a = b;
    }
    uint64_t n = 0;
    CHECK(xkrt_hip_taskdom_sync(t, &n));
    std::vector<uint64_t> e(2 * n + 2);
    uint64_t got = 0;
    CHECK(xkrt_hip_taskdom_edges_copy(t, e.data(), n, &got));
    for (uint64_t i = 0; i < got; ++i)
        std::printf("prog %llu %llu\n", (unsigned long long)e[2 * i], (unsigned long long)e[2 * i + 1]);
    xkrt_hip_taskdom_stats_t st{};
    CHECK(xkrt_hip_taskdom_stats(t, &st));
    if (st.blas_domains != 2 || st.accesses != acc.size()) return 1;
    CHECK(xkrt_hip_taskdom_destroy(t));
    return 0;
}

} // namespace

int main()
{
    CHECK(xkrt_hip_init(0));
    if (run_domain() || run_pool() || run_coherency() || run_taskdom()) return 1;
    std::printf("ok\n");
    return 0;
}
