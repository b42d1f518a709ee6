// kp_layout.h -- records shared by the kernels and the host side of the
// device pool (plain C++ layouts, identical on both sides).
#pragma once
#include <cstdint>

namespace kpd {

struct RootInfo {          // decomposition result per (job, root position)
    int32_t status;        // 1 valid, 0 not, <0 error
    uint32_t n, h, l;
    uint64_t wr;           // CSR entries read: d(r) + sum over later neighbours, then + sum over B and L
};

// One job's graph in HBM (local vertex ids 0..pn-1) and its parameters.
struct JobDev {
    const uint64_t *off;
    const uint32_t *nbr, *deg, *orig;
    const int32_t *dpos, *dseq;
    uint32_t pn;
    int32_t k, q;
    uint32_t pad;
};

// Per-job result accumulators (sums mod 2^64), one 64-byte line each.
struct JobAcc {
    unsigned long long count, ha, hb, tasks, alg, pad[3];
};
static_assert(sizeof(JobAcc) == 64, "JobAcc is one line");

struct RootDesc {          // one valid root; frames name roots by their index in the class table
    uint64_t addr;         // device address of the root's subgraph
    uint32_t n, h, l, slot;
    uint32_t job;
    uint16_t k, q;
};
static_assert(sizeof(RootDesc) == 32, "RootDesc layout");

struct Sched {             // persistent-kernel scheduler state (one per launch)
    uint32_t next_root;    // cursor over the start list
    uint32_t next_seed;    // cursor over the resumed frames
    uint32_t active;
    uint32_t q_count;      // published frames over all sub-queues
    uint32_t hungry;       // waves currently looking for work
    uint32_t pause;        // set once a pause is requested; waves dump and exit
    uint32_t exited;
    uint32_t dump_used;    // frames written to the dump array
    int32_t error;
    uint32_t pad0;
    uint64_t epoch_frames; // frames processed, in steps of the check interval
    uint64_t list_unused;
    uint64_t donated;
    uint64_t frames, spills, pops;
    uint64_t busy_ticks, idle_ticks;   // wall-clock ticks summed over waves
};

} // namespace kpd
