// Test infrastructure only: a driver around the reference's own KHP-tree
// engine (include/xkrt/memory/access/common/khp-tree.hpp), compiled from the
// reference sources by oracle/Makefile into oracle/_ref/libkhpref.so.
//
// The engine (insert / intersect / red-black k-subtrees) is the reference's
// code, included unmodified.  The BLAS dependency-tree hooks and the edge
// materialisation cannot be included: dependency-tree.hpp pulls task.hpp ->
// logger.h -> the CMake-generated xkrt/support.h.  They are therefore restated
// here, line for line in behaviour, from
//   include/xkrt/memory/access/blas/dependency-tree.hpp:93-186  (node state)
//   include/xkrt/memory/access/blas/dependency-tree.hpp:229-256 (on_insert)
//   include/xkrt/memory/access/blas/dependency-tree.hpp:296-355 (stop test, on_intersect)
//   include/xkrt/memory/access/blas/dependency-tree.hpp:367-414 (link / put)
//   include/xkrt/task/task.hpp:494-512                         (__access_precedes dedup)
//   src/task/task.cc:285-299                                   (LINK pass then PUT pass)
// Tasks never complete in this harness, so every edge is materialised.
//
// The memory-tree partition helper inserts access rects and lists the nodes
// that intersect them (memory-tree.hpp:1738-1775 steps (1)-(2), with the
// host short path of 1183-1192: nothing is inserted while the tree is empty).

#include <xkrt/memory/access/common/khp-tree.hpp>

#include <algorithm>
#include <cstdint>
#include <ctime>
#include <memory>
#include <vector>

namespace {

enum : uint8_t { MODE_R = 1, MODE_W = 2 };

using Rect = KHyperrect<2>;

// an edge target: the successor access's task and index.  An access's list
// may be shared by its views in several trees (one program: the interval tree
// and the BLAS trees), so __access_precedes sees every edge of the access.
struct Succ {
    uint32_t task, idx;
};
using SuccList = std::vector<Succ>;

struct Acc {
    uint32_t idx;
    uint32_t task;
    uint8_t mode;
    int nrect;
    Rect rects[3];  // inserted and intersected (dependency-tree.hpp:366-426)
    int nmark;
    Rect marks[3];  // what on_insert tests a node against: access->rects() (dependency-tree.hpp:241-256)
    SuccList own;
    SuccList *succ = nullptr;  // shared list, else own
};

struct DepSearch {
    Acc *access;
};

using DepBase = KHPTree<2, DepSearch>;

class DepNode : public DepBase::Node {
    using Base = DepBase::Node;
  public:
    std::vector<Acc *> last_reads;
    Acc *last_write;
    int nwrites;

    DepNode(const Rect &h, int k, Color c) : Base(h, k, c), last_reads(), last_write(nullptr), nwrites(0) {}
    DepNode(const Rect &h, int k, Color c, const DepNode *inherit)
        : Base(h, k, c), last_reads(inherit->last_reads), last_write(inherit->last_write), nwrites(0) {}

    void update_includes(void) override
    {
        Base::update_includes();
        nwrites = last_write ? 1 : 0;
        for (int k = 0; k < 2; ++k)
            for (int d = 0; d < 2; ++d) {
                DepNode *c = reinterpret_cast<DepNode *>(this->st[k].children[d]);
                if (c) nwrites += c->nwrites;
            }
    }
};

static inline void precedes(SuccList &l, uint32_t task, uint32_t idx)
{
    if (!l.empty() && l.back().task == task) return;
    l.push_back({task, idx});
}

static inline void access_precedes(Acc *pred, Acc *succ)
{
    precedes(pred->succ ? *pred->succ : pred->own, succ->task, succ->idx);
}

class DepTree : public DepBase {
  public:
    Node *new_node(DepSearch &, const Rect &h, const int k, const Color c) const override
    {
        return new DepNode(h, k, c);
    }
    Node *new_node(DepSearch &, const Rect &h, const int k, const Color c, const Node *inherit) const override
    {
        return new DepNode(h, k, c, reinterpret_cast<const DepNode *>(inherit));
    }
    void on_insert(Node *nb, DepSearch &s) override
    {
        DepNode *node = reinterpret_cast<DepNode *>(nb);
        for (int r = 0; r < s.access->nmark; ++r) {
            if (s.access->marks[r].intersects(node->hyperrect)) {
                if (s.access->mode & MODE_W) {
                    node->last_reads.clear();
                    node->last_write = s.access;
                } else if (s.access->mode == MODE_R) {
                    node->last_reads.push_back(s.access);
                }
                break;
            }
        }
    }
    void on_shrink(Node *, const Interval &, int) override {}
    bool intersect_stop_test(Node *nb, DepSearch &s, const Rect &) const override
    {
        return s.access->mode == MODE_R && reinterpret_cast<DepNode *>(nb)->nwrites == 0;
    }
    void on_intersect(Node *nb, DepSearch &s, const Rect &) const override
    {
        DepNode *node = reinterpret_cast<DepNode *>(nb);
        if ((s.access->mode & MODE_W) && !node->last_reads.empty()) {
            for (Acc *p : node->last_reads) access_precedes(p, s.access);
        } else if (node->last_write) {
            access_precedes(node->last_write, s.access);
        }
    }
    void link(Acc *a)
    {
        DepSearch s{a};
        for (int r = 0; r < a->nrect; ++r) intersect(s, a->rects[r]);
    }
    void put(Acc *a)
    {
        DepSearch s{a};
        for (int r = 0; r < a->nrect; ++r) insert(s, a->rects[r]);
    }
};

// mode flag (above the access modes) of an access replayed without a LINK
constexpr uint8_t PUT_ONLY = 0x40;

// The segment (K = 1) tree: the hooks of IntervalDependencyTree
// (include/xkrt/memory/access/interval/dependency-tree.hpp:229-355) restated
// over the reference's KHPTree<1>: a node is marked when the segment
// intersects it, the LINK rule is the BLAS tree's.
using IRect = KHyperrect<1>;
struct IAcc {
    uint32_t idx;
    uint32_t task;
    uint8_t mode;
    IRect seg;
    SuccList own;
    SuccList *succ = nullptr;
};
struct ISearch {
    IAcc *access;
};
using IntBase = KHPTree<1, ISearch>;

class IntNode : public IntBase::Node {
    using Base = IntBase::Node;
  public:
    std::vector<IAcc *> last_reads;
    IAcc *last_write;
    int nwrites;
    IntNode(const IRect &h, int k, Color c) : Base(h, k, c), last_reads(), last_write(nullptr), nwrites(0) {}
    IntNode(const IRect &h, int k, Color c, const IntNode *inherit)
        : Base(h, k, c), last_reads(inherit->last_reads), last_write(inherit->last_write), nwrites(0) {}
    void update_includes(void) override
    {
        Base::update_includes();
        nwrites = last_write ? 1 : 0;
        for (int d = 0; d < 2; ++d) {
            IntNode *c = reinterpret_cast<IntNode *>(this->st[0].children[d]);
            if (c) nwrites += c->nwrites;
        }
    }
};

static inline void iaccess_precedes(IAcc *pred, IAcc *succ)
{
    precedes(pred->succ ? *pred->succ : pred->own, succ->task, succ->idx);
}

class IntTree : public IntBase {
  public:
    Node *new_node(ISearch &, const IRect &h, const int k, const Color c) const override { return new IntNode(h, k, c); }
    Node *new_node(ISearch &, const IRect &h, const int k, const Color c, const Node *inherit) const override
    {
        return new IntNode(h, k, c, reinterpret_cast<const IntNode *>(inherit));
    }
    void on_insert(Node *nb, ISearch &s) override
    {
        IntNode *node = reinterpret_cast<IntNode *>(nb);
        if (s.access->seg.intersects(node->hyperrect)) {
            if (s.access->mode & MODE_W) {
                node->last_reads.clear();
                node->last_write = s.access;
            } else if (s.access->mode == MODE_R) {
                node->last_reads.push_back(s.access);
            }
        }
    }
    void on_shrink(Node *, const Interval &, int) override {}
    bool intersect_stop_test(Node *nb, ISearch &s, const IRect &) const override
    {
        return s.access->mode == MODE_R && reinterpret_cast<IntNode *>(nb)->nwrites == 0;
    }
    void on_intersect(Node *nb, ISearch &s, const IRect &) const override
    {
        IntNode *node = reinterpret_cast<IntNode *>(nb);
        if ((s.access->mode & MODE_W) && !node->last_reads.empty()) {
            for (IAcc *p : node->last_reads) iaccess_precedes(p, s.access);
        } else if (node->last_write) {
            iaccess_precedes(node->last_write, s.access);
        }
    }
    void link(IAcc *a)
    {
        ISearch s{a};
        intersect(s, a->seg);
    }
    void put(IAcc *a)
    {
        ISearch s{a};
        insert(s, a->seg);
    }
};

// Plain fragment tree used for the memory-tree partition (hooks are no-ops,
// as memory-tree.hpp's INSERTING_BLOCKS on_insert is).
struct PartSearch { std::vector<Rect> *out; };
using PartBase = KHPTree<2, PartSearch>;
class PartTree : public PartBase {
  public:
    Node *new_node(PartSearch &, const Rect &h, const int k, const Color c) const override { return new Node(h, k, c); }
    Node *new_node(PartSearch &, const Rect &h, const int k, const Color c, const Node *) const override { return new Node(h, k, c); }
    void on_insert(Node *, PartSearch &) override {}
    void on_shrink(Node *, const Interval &, int) override {}
    bool intersect_stop_test(Node *, PartSearch &, const Rect &) const override { return false; }
    void on_intersect(Node *nb, PartSearch &s, const Rect &) const override { s.out->push_back(nb->hyperrect); }
};

static inline Rect make_rect(const uint64_t *r)
{
    Interval l[2] = {Interval(r[0], r[1]), Interval(r[2], r[3])};
    return Rect(l);
}

static inline uint64_t now_ns(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

static uint64_t count_frags(DepTree &t, const Acc &a)
{
    // Fragments are the tree's nodes; count those intersecting the access.
    uint64_t n = 0;
    std::vector<DepBase::Node *> stack;
    if (t.root) stack.push_back(t.root);
    while (!stack.empty()) {
        DepBase::Node *node = stack.back();
        stack.pop_back();
        bool hit = false, sub = false;
        for (int r = 0; r < a.nrect; ++r) {
            if (a.rects[r].intersects(node->hyperrect)) hit = true;
            if (a.rects[r].intersects(node->includes.hyperrect)) sub = true;
        }
        if (!sub) continue;
        if (hit) n++;
        for (int k = 0; k < 2; ++k)
            for (int d = 0; d < 2; ++d)
                if (node->st[k].children[d]) stack.push_back(node->st[k].children[d]);
    }
    return n;
}

} // namespace

extern "C" {

// Resolve one dependency domain.  Accesses are in spawn order, grouped by
// task (task[] non-decreasing); rects holds 8 u64 per access (rect 0 then
// rect 1), nrect 1 or 2.  Writes up to edge_cap (pred, succ) access-index
// pairs, grouped by predecessor in access order then successor-list order.
// Returns the number of edges (which may exceed edge_cap), or -1.
// frag (nullable) receives per access the number of tree nodes intersecting
// its rects after its task's PUT pass.  task_ns (nullable, one entry per
// task) receives the wall time of each task's LINK + PUT.
int64_t khpref_resolve(uint64_t nacc, const uint32_t *task, const uint8_t *mode, const uint8_t *nrect,
                       const uint64_t *rects, uint64_t edge_cap, uint32_t *edges, uint64_t *frag,
                       uint64_t *task_ns)
{
    std::vector<Acc> acc(nacc);
    for (uint64_t i = 0; i < nacc; ++i) {
        acc[i].idx = (uint32_t)i;
        acc[i].task = task[i];
        acc[i].mode = mode[i];
        acc[i].nrect = nrect[i];
        acc[i].rects[0] = make_rect(rects + 8 * i);
        acc[i].rects[1] = make_rect(rects + 8 * i + 4);
        acc[i].nmark = acc[i].nrect;
        acc[i].marks[0] = acc[i].rects[0];
        acc[i].marks[1] = acc[i].rects[1];
        acc[i].own.reserve(8);
    }
    DepTree tree;
    uint64_t i = 0, tix = 0;
    while (i < nacc) {
        uint64_t j = i;
        while (j < nacc && task[j] == task[i]) ++j;
        uint64_t t0 = task_ns ? now_ns() : 0;
        for (uint64_t a = i; a < j; ++a) tree.link(&acc[a]);
        for (uint64_t a = i; a < j; ++a) tree.put(&acc[a]);
        if (task_ns) task_ns[tix] = now_ns() - t0;
        if (frag)
            for (uint64_t a = i; a < j; ++a) frag[a] = count_frags(tree, acc[a]);
        i = j;
        tix++;
    }
    uint64_t ne = 0;
    for (uint64_t p = 0; p < nacc; ++p)
        for (const Succ &s : acc[p].own) {
            if (ne < edge_cap) {
                edges[2 * ne] = (uint32_t)p;
                edges[2 * ne + 1] = s.idx;
            }
            ne++;
        }
    return (int64_t)ne;
}

// Same, up to 3 rects per access (12 u64 each) inserted / intersected, and the
// rects on_insert tests against given separately (marks, 12 u64 each): for a
// segment seen by a BLAS domain, interval_to_rects in the domain frame and
// the segment's own rects (built with ld = SIZE_MAX).
int64_t khpref_resolve_marks(uint64_t nacc, const uint32_t *task, const uint8_t *mode, const uint8_t *nrect,
                             const uint64_t *rects, const uint8_t *nmark, const uint64_t *marks, uint64_t edge_cap,
                             uint32_t *edges)
{
    std::vector<Acc> acc(nacc);
    for (uint64_t i = 0; i < nacc; ++i) {
        acc[i].idx = (uint32_t)i;
        acc[i].task = task[i];
        acc[i].mode = mode[i] & (uint8_t)~PUT_ONLY;
        acc[i].nrect = nrect[i];
        acc[i].nmark = nmark[i];
        for (int r = 0; r < 3; ++r) {
            acc[i].rects[r] = make_rect(rects + 12 * i + 4 * r);
            acc[i].marks[r] = make_rect(marks + 12 * i + 4 * r);
        }
        acc[i].own.reserve(8);
    }
    DepTree tree;
    uint64_t i = 0;
    while (i < nacc) {
        uint64_t j = i;
        while (j < nacc && task[j] == task[i]) ++j;
        // an access flagged PUT_ONLY is replayed into the tree without a LINK: the segment
        // accesses put_interval'd into a BLAS tree when it is created (src/task/task.cc:177-198)
        for (uint64_t a = i; a < j; ++a)
            if (!(mode[a] & PUT_ONLY)) tree.link(&acc[a]);
        for (uint64_t a = i; a < j; ++a) tree.put(&acc[a]);
        i = j;
    }
    uint64_t ne = 0;
    for (uint64_t p = 0; p < nacc; ++p)
        for (const Succ &sc : acc[p].own) {
            if (ne < edge_cap) {
                edges[2 * ne] = (uint32_t)p;
                edges[2 * ne + 1] = sc.idx;
            }
            ne++;
        }
    return (int64_t)ne;
}

// Resolve one segment (K = 1) dependency domain: IntervalDependencyTree
// (include/xkrt/memory/access/interval/dependency-tree.hpp:187-363) over
// the reference's KHPTree<1>.  seg holds 2 u64 per access ([a, b) bytes).
// Edges as khpref_resolve.
int64_t khpref_resolve_intervals(uint64_t nacc, const uint32_t *task, const uint8_t *mode, const uint64_t *seg,
                                 uint64_t edge_cap, uint32_t *edges)
{
    std::vector<IAcc> acc(nacc);
    for (uint64_t i = 0; i < nacc; ++i) {
        acc[i].idx = (uint32_t)i;
        acc[i].task = task[i];
        acc[i].mode = mode[i];
        Interval l[1] = {Interval(seg[2 * i], seg[2 * i + 1])};
        acc[i].seg = IRect(l);
        acc[i].own.reserve(8);
    }
    IntTree tree;
    uint64_t i = 0;
    while (i < nacc) {
        uint64_t j = i;
        while (j < nacc && task[j] == task[i]) ++j;
        for (uint64_t a = i; a < j; ++a) tree.link(&acc[a]);
        for (uint64_t a = i; a < j; ++a) tree.put(&acc[a]);
        i = j;
    }
    uint64_t ne = 0;
    for (uint64_t p = 0; p < nacc; ++p)
        for (const Succ &sc : acc[p].own) {
            if (ne < edge_cap) {
                edges[2 * ne] = (uint32_t)p;
                edges[2 * ne + 1] = sc.idx;
            }
            ne++;
        }
    return (int64_t)ne;
}

// Total wall time (ns) of LINK + PUT over the stream, nothing else recorded.
uint64_t khpref_time(uint64_t nacc, const uint32_t *task, const uint8_t *mode, const uint8_t *nrect,
                     const uint64_t *rects, uint64_t *nedges)
{
    std::vector<Acc> acc(nacc);
    for (uint64_t i = 0; i < nacc; ++i) {
        acc[i].idx = (uint32_t)i;
        acc[i].task = task[i];
        acc[i].mode = mode[i];
        acc[i].nrect = nrect[i];
        acc[i].rects[0] = make_rect(rects + 8 * i);
        acc[i].rects[1] = make_rect(rects + 8 * i + 4);
        acc[i].nmark = acc[i].nrect;
        acc[i].marks[0] = acc[i].rects[0];
        acc[i].marks[1] = acc[i].rects[1];
        acc[i].own.reserve(8);
    }
    DepTree tree;
    uint64_t t0 = now_ns();
    uint64_t i = 0;
    while (i < nacc) {
        uint64_t j = i;
        while (j < nacc && task[j] == task[i]) ++j;
        for (uint64_t a = i; a < j; ++a) tree.link(&acc[a]);
        for (uint64_t a = i; a < j; ++a) tree.put(&acc[a]);
        i = j;
    }
    uint64_t dt = now_ns() - t0;
    uint64_t ne = 0;
    for (uint64_t p = 0; p < nacc; ++p) ne += acc[p].own.size();
    if (nedges) *nedges = ne;
    return dt;
}

// Memory-tree partition of each access of one coherency domain: host
// accesses (dev 0) are skipped while the tree is empty; otherwise every rect
// is inserted, then every node intersecting a rect is listed.  part_off gets
// nacc + 1 offsets; parts gets 4 u64 per partite (x0, x1, y0, y1).  Returns
// the total number of partites (may exceed part_cap).
int64_t khpref_partition(uint64_t nacc, const uint8_t *dev, const uint8_t *nrect, const uint64_t *rects,
                         uint64_t part_cap, uint64_t *part_off, uint64_t *parts)
{
    PartTree tree;
    std::vector<Rect> out;
    uint64_t np = 0;
    for (uint64_t i = 0; i < nacc; ++i) {
        part_off[i] = np;
        if (dev[i] == 0 && tree.root == nullptr) continue;
        PartSearch s{&out};
        Rect r[2] = {make_rect(rects + 8 * i), make_rect(rects + 8 * i + 4)};
        for (int k = 0; k < nrect[i]; ++k) tree.insert(s, r[k]);
        out.clear();
        for (int k = 0; k < nrect[i]; ++k) tree.intersect(s, r[k]);
        for (const Rect &p : out) {
            if (np < part_cap) {
                parts[4 * np + 0] = p[0].a; parts[4 * np + 1] = p[0].b;
                parts[4 * np + 2] = p[1].a; parts[4 * np + 3] = p[1].b;
            }
            np++;
        }
    }
    part_off[nacc] = np;
    return (int64_t)np;
}

// One task domain's dependency resolution over a mixed program: the dispatch
// of task_dependency_resolve_do (src/task/task.cc:204-277) over one interval
// tree and one BLAS tree per (ld, sizeof_type), created on first use
// (task_get_dependency_domain_blas_matrix, task.cc:153-201: a new tree gets
// the interval tree's accesses put, newest first, with no LINK).  Per task:
// every access LINKed in slot order, then every access PUT (task.cc:285-299).
// A segment links / puts into the interval tree and then into every BLAS tree
// existing at that moment (link_interval / put_interval, dependency-tree.hpp:
// 384-427: rects in the tree's frame, marks in the ld = SIZE_MAX frame).  Each
// access keeps one successor list across all its trees (__access_precedes).
// kind 0 = tile (addr, m rows, n columns, ld, s), 1 = segment [addr, addr + m).
// Edges as khpref_resolve, in program access indices.
int64_t khpref_resolve_program(uint64_t nacc, const uint32_t *task, const uint8_t *mode, const uint8_t *kind,
                               const uint64_t *addr, const uint64_t *m, const uint64_t *n, const uint64_t *ld,
                               const uint64_t *s, uint64_t edge_cap, uint32_t *edges)
{
    struct Blas {
        uint64_t ld, s;
        DepTree tree;
        std::vector<std::unique_ptr<Acc>> views;  // per program access, created on first use
    };
    std::vector<SuccList> succ(nacc);
    std::vector<IAcc> iacc(nacc);
    std::vector<std::unique_ptr<Blas>> blas;
    IntTree itree;
    std::vector<uint32_t> iput;  // accesses put into the interval tree, in put order
    auto interval_rects = [](uint64_t a, uint64_t size, uint64_t LD, uint64_t sz, Rect out[3]) {
        const uint64_t LDs = LD * sz, x0 = a % LDs, x1 = std::min(x0 + size, LDs), dx10 = x1 - x0;
        const uint64_t x3 = (size - dx10) % LDs, y0 = a / LDs, y1 = y0 + 1;
        const uint64_t y3 = y0 + (size - dx10 - x3) / LDs + 2, y2 = y3 - 1;
        const uint64_t r[3][4] = {{x0, x1, y0, y1}, {0, LDs, y1, y2}, {0, x3, y2, y3}};
        for (int k = 0; k < 3; ++k) out[k] = make_rect(r[k]);
    };
    auto view = [&](Blas &b, uint32_t i) -> Acc * {
        if (b.views.size() < nacc) b.views.resize(nacc);
        if (!b.views[i]) {
            auto a = std::make_unique<Acc>();
            a->idx = i;
            a->task = task[i];
            a->mode = mode[i];
            a->succ = &succ[i];
            if (kind[i] == 1) {
                interval_rects(addr[i], m[i], b.ld, b.s, a->rects);
                interval_rects(addr[i], m[i], UINT64_MAX, 1, a->marks);
                a->nrect = a->nmark = 3;
            } else {
                // matrix_to_rects (access.hpp:205-287): one rect, or two when the tile wraps
                const uint64_t LDs = b.ld * b.s, X0 = addr[i] % LDs, Y0 = addr[i] / LDs, ms = m[i] * b.s;
                if (X0 + ms <= LDs) {
                    const uint64_t r[4] = {X0, X0 + ms, Y0, Y0 + n[i]};
                    a->rects[0] = make_rect(r);
                    a->nrect = 1;
                } else {
                    const uint64_t r0[4] = {X0, LDs, Y0, Y0 + n[i]}, r1[4] = {0, ms - (LDs - X0), Y0 + 1, Y0 + n[i] + 1};
                    a->rects[0] = make_rect(r0);
                    a->rects[1] = make_rect(r1);
                    a->nrect = 2;
                }
                a->nmark = a->nrect;
                for (int k = 0; k < a->nrect; ++k) a->marks[k] = a->rects[k];
            }
            b.views[i] = std::move(a);
        }
        return b.views[i].get();
    };
    auto tree_of = [&](uint32_t i) -> Blas & {
        for (auto &b : blas)
            if (b->ld == ld[i] && b->s == s[i]) return *b;
        blas.push_back(std::make_unique<Blas>());
        Blas &b = *blas.back();
        b.ld = ld[i];
        b.s = s[i];
        for (auto it = iput.rbegin(); it != iput.rend(); ++it) b.tree.put(view(b, *it));
        return b;
    };
    for (uint64_t i = 0; i < nacc; ++i) {
        iacc[i].idx = (uint32_t)i;
        iacc[i].task = task[i];
        iacc[i].mode = mode[i];
        Interval l[1] = {Interval(addr[i], addr[i] + m[i])};
        iacc[i].seg = IRect(l);
        iacc[i].succ = &succ[i];
    }
    uint64_t i = 0;
    while (i < nacc) {
        uint64_t j = i;
        while (j < nacc && task[j] == task[i]) ++j;
        for (uint64_t a = i; a < j; ++a) {
            if (kind[a] == 1) {
                itree.link(&iacc[a]);
                for (auto &b : blas) b->tree.link(view(*b, (uint32_t)a));
            } else {
                tree_of((uint32_t)a).tree.link(view(tree_of((uint32_t)a), (uint32_t)a));
            }
        }
        for (uint64_t a = i; a < j; ++a) {
            if (kind[a] == 1) {
                itree.put(&iacc[a]);
                iput.push_back((uint32_t)a);
                for (auto &b : blas) b->tree.put(view(*b, (uint32_t)a));
            } else {
                tree_of((uint32_t)a).tree.put(view(tree_of((uint32_t)a), (uint32_t)a));
            }
        }
        i = j;
    }
    uint64_t ne = 0;
    for (uint64_t p = 0; p < nacc; ++p)
        for (const Succ &sc : succ[p]) {
            if (ne < edge_cap) {
                edges[2 * ne] = (uint32_t)p;
                edges[2 * ne + 1] = sc.idx;
            }
            ne++;
        }
    return (int64_t)ne;
}

} // extern "C"
