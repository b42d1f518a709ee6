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

// Per-job result accumulators (sums mod 2^64).  The second line holds what
// the monolithic per-root kernel adds for the roots it completes (the pooled
// path counts its roots on the host while staging).
struct JobAcc {
    unsigned long long count, ha, hb, tasks, alg, pad0[3];
    unsigned long long roots, adj, dalg;   // valid roots, 4 W_r, 4 W_r + ceil(n^2/8) + n^2
    unsigned int max_n, max_l;
    unsigned long long pad1[4];
};
static_assert(sizeof(JobAcc) == 128, "JobAcc is two lines");

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
    uint32_t retired;      // waves that exited early as surplus
    uint32_t started;      // waves that have become resident
    uint64_t epoch_frames; // frames processed, in steps of the check interval
    uint64_t list_unused;
    uint64_t donated;
    uint64_t frames, spills, pops;
    uint64_t busy_ticks, idle_ticks;   // wall-clock ticks summed over waves
};


// Words of a task frame of a W-word size class: meta, P, C, X, C2, rounded
// to an even count so that frames stay 16-byte aligned.
constexpr int FrameWords(int W) { return (4 * W + 2) & ~1; }

// ---- kernel arguments (built on the host, passed by value) ----

// One shard of the frame queue: a bounded MPMC ring with sequence-numbered
// slots.  Shards sit on separate 128-byte lines so that waves polling and
// claiming different shards do not contend on one address.
struct SubQ {
    uint32_t count;        // published frames
    uint32_t size;         // reserved slots
    uint64_t head, tail;   // tickets
    uint32_t pad[26];
};
static_assert(sizeof(SubQ) == 128, "SubQ is one 128-byte line");
constexpr uint32_t NSUBQ = 32;

struct DecompArgs {
    const JobDev *jobs;
    const uint32_t *slot_job, *slot_pos;   // per slot: job index and root position
    uint32_t nslots;
    uint32_t *blk, *lft;
    uint32_t lcap;
    RootInfo *info;
    uint32_t *ws_cnt, *ws_b1c;
    uint64_t *ws_a1;
    uint32_t cnt_stride;                   // counter words per wave (largest pn)
    uint32_t d1cap, w1cap, nwaves;
};

struct BuildArgs {
    const JobDev *jobs;
    const uint32_t *blk, *lft;
    uint32_t lcap;
    const RootDesc *desc;
    uint32_t ndesc;
    uint32_t *ws_cnt;
    uint32_t cnt_stride, nwaves;
    uint64_t *wr_extra;   // per desc: sum of degrees over B and L
};

// Narrowing: a frame of a wide size class whose live vertices (P, C, the
// k-search candidates and the excluded vertices that can still extend a plex
// of its subtree) fit one word is re-rooted as a derived root of the one-word
// class.  The derived subgraph keeps those vertices in block order, so the
// one-word search walks the same tree.  An epoch fills one derived-root buffer
// (subgraph arena, root entries, one frame per root); the one-word class
// resumes those frames in the next epoch.  The root entries are written into a
// region of the one-word class's own root table reserved past its regular
// roots (buffer b from base + b * cap), so its frames name derived roots like
// any other.
constexpr uint32_t NDERIVED_BUF = 4;
struct NarrowArgs {
    uint64_t *arena;              // derived subgraphs, bump-allocated
    uint64_t arena_cap;           // words
    unsigned long long *ctr;      // [0] arena words reserved, [1] derived roots reserved
    RootDesc *desc;               // derived root i: one-word class table entry tag + i
    uint64_t *frames;             // its frame (one-word layout, FrameWords(1) words)
    uint32_t cap;                 // derived roots (0 = narrowing off)
    uint32_t tag;                 // the buffer's first index in the one-word class table
    uint32_t min_slack;           // only frames with |P| + |C ∪ D| >= q + min_slack are narrowed
    uint32_t pause_when_full;     // a wave that finds the buffer full pauses its class
    uint64_t after_ticks;         // narrowing starts this long after the epoch's first wave
};

struct SearchArgs {
    const RootDesc *desc;      // class root table
    const uint32_t *todo;      // roots to start, in order (indices into desc)
    uint32_t ntodo;
    const uint64_t *seed;      // frames to resume: seed[0, nseed_dev) on the device,
    const uint64_t *hseed;     // then hseed[0, nseed - nseed_dev) in mapped pinned host memory
    uint32_t nseed, nseed_dev;
    uint32_t seed_chunk;       // resumed frames a wave claims at once (<= the stack room for them)
    uint64_t *dump;            // frames written on pause
    uint32_t dump_cap;
    JobAcc *acc;               // per-job accumulators
    uint64_t *qframes, *qseq;  // NSUBQ * qcap slots
    uint32_t qcap;             // slots per shard
    SubQ *subq;
    Sched *sched;
    uint32_t *list;
    uint64_t list_cap;
    uint64_t *list_used;       // shared by all size-class launches
    uint32_t nwaves;
    uint32_t ring_cap;         // frames per wave in the LDS ring
    uint32_t back_cap;         // frames per wave in the HBM backing stack
    uint64_t *gstack;          // nwaves * back_cap frames
    uint32_t donate_period, donate_target;
    uint64_t limit_ticks;      // 0 = no limit; else wall-clock ticks per launch
    const uint32_t *host_pause;   // mapped host word (nonzero = pause) or null
    const uint32_t *class_pause;  // mapped host word of this class alone, or null
    uint64_t pause_frames;     // 0 = none; else pause once this many frames were processed
    unsigned long long *t0;    // shared start tick of the epoch (0 until the first wave of any class starts)
    uint32_t retire_margin;    // hungry waves kept beyond the busy ones once nothing is left to start
    uint64_t retire_ticks;     // wall-clock ticks a wave stays hungry before it may retire
    NarrowArgs nar;            // wide classes only
};

struct RootKernArgs {
    DecompArgs d;              // jobs, slot lists, workspaces; blk/lft/info hold one entry per wave
    BuildArgs b;               // jobs, blk/lft, lcap; wr_extra one entry per wave
    uint64_t *arena;           // arena_words per wave
    uint64_t arena_words;
    RootDesc *desc;            // one entry per wave (the search reads it as its root table)
    uint32_t *deferred, *ndeferred;
    uint64_t budget;           // frames per root before it is deferred (0 = none)
    SearchArgs s;
};

} // namespace kpd
