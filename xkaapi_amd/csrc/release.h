// Successor release on the device (private to the library): wait counters,
// successor lists and the ready queue of one task domain, fed by the edge
// buffer of its pool in stream order.  See release.hip.
#pragma once

#include <cstdint>

namespace xkrt_release {

struct Engine;

// capacities: tasks (numbered 0, 1, ... in submit order), accesses (program
// ids), successor-list nodes (one per edge resolved), empty writes that handle
// maps may insert (edge ends 0x80000000 | k: empty tasks, completed as soon as
// they are released)
int create(void *stream, uint64_t max_tasks, uint64_t max_acc, uint64_t max_nodes, uint64_t max_virtual,
           Engine **out);
void destroy(Engine *e);

// a batch of whole tasks: accesses [acc0, acc0 + n) of tasks task[0 .. n)
// (consecutive task numbers, continuing the previous batch) with their modes
int add(Engine *e, uint64_t acc0, const uint32_t *task, const uint8_t *mode, uint64_t n);

// a handle map's edges on the device: `stripes` stripes of stripe_cap pairs,
// the count of each, and the device word counting its inserted empty writes
struct HandleEdges {
    const uint32_t *edges;
    const uint64_t *count;
    uint32_t stripes;
    uint64_t stripe_cap;
    const uint32_t *nvirtual;
};

// after a pool launch, on its stream: applies the unconsumed edges of the
// pool's buffer (count in device memory) and of the handle map (nullable),
// drops the guards of the empty writes inserted since the last call, then
// the commit guard of the launch's tasks [t0, t1)
int apply(Engine *e, const uint32_t *edges, const uint64_t *count, const HandleEdges *h, uint32_t t0, uint32_t t1);

// completion of ready tasks (stream order)
int complete(Engine *e, const uint32_t *tasks, uint64_t n);

// the same, then waits until the device has released their successors (and
// those of every completion enqueued before): they are in the queue on return
int complete_wait(Engine *e, const uint32_t *tasks, uint64_t n);

// released tasks not read yet, in release order (non-blocking)
int ready(Engine *e, uint32_t *out, uint64_t cap, uint64_t *n);

// synchronises the stream; 0 or -errno of the device's error bits
int check(Engine *e);

struct Counts {
    uint64_t tasks, accesses, nodes, released, read, completed;
};
int counts(Engine *e, Counts *out);

} // namespace xkrt_release
