// kp_engine.h -- host side of the device pool shared by the plain, batched
// and served entry points (compiled with hipcc only).
//
// A job is (preprocessed graph resident in HBM, k, q, window of degeneracy
// positions).  The engine decomposes the roots of any set of jobs, builds
// their subgraphs, files the valid roots into per-size-class root tables and
// runs the persistent search kernels over all of them in epochs.  An epoch
// ends when every class has drained, or when a pause was requested (a frame
// budget, or the host pause word): then the pending frames of each class are
// carried into the next epoch, at most `seed_budget` of them on the device and
// the rest in pinned host memory.
#pragma once
#include <hip/hip_runtime.h>

#include "kp_internal.h"
#include "kp_layout.h"

#include <functional>
#include <vector>

namespace kp {

constexpr int kNumClass = 7;
// narrowing defaults: classes of >= 2 words, frames with |P| + |C ∪ D| >= q + 10;
// buffers of a 4 GiB arena and 4M derived roots
constexpr int kNarrowFrom = 2;
constexpr uint32_t kNarrowMinSlack = 10;
constexpr size_t kNarrowArenaBytes = size_t(4) << 30;
constexpr uint32_t kNarrowRoots = 1u << 22;
// narrowing and regrid pay only in long epochs (measured: soc-epinions 184 ->
// 310 ms and email-euall 184 -> 348 ms with narrowing from the first frame)
constexpr double kNarrowAfterMs = 250.0;
constexpr double kRegridAfterMs = 250.0;
extern const int kClassW[kNumClass];   // bitset words per size class
int class_of(uint32_t n);              // -1 when n > 64 * 16

// Launch settings of one size class (0 = the built-in default); set by
// kp_set_tune / kp_autotune and read by every engine at launch.
struct ClassTune {
    int blocks_per_cu = 0;       // <= the occupancy limit
    uint32_t donate_period = 0;  // frames between donation checks (default 256)
    uint32_t ring_frames = 0;    // LDS ring frames per wave (default: 8 KiB)
    uint32_t queue_slots = 0;    // frame-queue slots per shard
};
extern ClassTune g_tune[kNumClass];

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);            // grows (contents are not kept)
    void release();
    template <class T> T *as() const { return static_cast<T *>(p); }
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
};

// A preprocessed graph in HBM (local ids 0..pn-1) plus host-side sizes.
struct GraphView {
    const uint64_t *off = nullptr;
    const uint32_t *nbr = nullptr, *deg = nullptr, *orig = nullptr;
    const int32_t *dpos = nullptr, *dseq = nullptr;
    uint32_t pn = 0;
    uint64_t pm = 0;
    uint32_t maxdeg = 0, d1max = 0;   // largest degree / largest later-neighbour count
};

// Host or GPU preprocessing (KP_FLAG_GPU_PREP in p.flags) of g for p's k, q, truss.
int prep_graph(const kp_graph *g, const kp_params &p, Prep *out);

// Packs preprocessed graphs into one device allocation (one host->HBM copy).
size_t packed_graph_bytes(const Prep &p);
int upload_graphs(const std::vector<const Prep *> &preps, DevBuf &buf, std::vector<GraphView> &views,
                  hipStream_t st);

struct JobIn {
    uint32_t id;               // index in the engine's job table
    GraphView g;
    int k, q;
    int64_t pos_begin, pos_end;   // root positions, already clipped to [0, pn - q + 2)
    const std::vector<uint32_t> *pos_list = nullptr;   // if set: exactly these positions instead
};

struct JobStats {
    uint64_t valid_roots = 0, adj_bytes = 0, alg_bytes = 0;
    uint32_t max_n = 0, max_l = 0;
    uint64_t arena_bytes = 0;
};

struct SearchOpts {
    uint64_t mono_budget = 0;     // > 0: a per-root pass first, roots over this many frames go to the pool
    bool want_list = false;
    uint64_t list_cap = 0;        // words
    uint64_t limit_ticks = 0;
    uint64_t pause_frames = 0;    // frame budget per launch (0 = none)
    bool host_pause = false;      // waves poll the engine's pinned pause word
    bool pausable = false;        // allocate dump arrays (needed by either pause source)
    uint64_t seed_budget = 1;     // resumed frames kept on the device per class
    // narrowing: frames of classes of >= narrow_from words whose live set fits
    // one word become derived roots of the one-word class (0 = off)
    int narrow_from = 0;
    uint32_t narrow_min_slack = 0;   // ... when |P| + |C ∪ D| - q is at least this
    bool narrow_pause = false;       // a wide class whose buffer is full pauses (needs pausable)
    double narrow_after_ms = 0;      // narrowing starts this long into an epoch
    double regrid_after_ms = 0;      // regrid pauses only in epochs older than this
    // once only classes launched on a reduced grid are still running, pause
    // them so that the next epoch gives them the whole chip
    bool regrid = false;
};

struct ClassState {
    DevBuf q, qseq, sched, stack, subq, desc, todo, seed[2], dump;
    hipStream_t stream = nullptr;
    int blocks_per_cu = 0, nb_max = 0;  // blocks per CU of the last launch / the occupancy limit
    bool regrid_sent = false;           // the host asked the running launch to pause for a larger grid
    bool demoted = false;               // the last launch ran at one block per CU beside a primary class
    std::vector<kpd::RootDesc> hdesc;   // root table (frames name roots by index)
    std::vector<uint32_t> htodo;        // start list
    size_t todo_done = 0;               // entries of htodo already started
    size_t desc_uploaded = 0;           // entries of hdesc copied to the device
    uint32_t maxn = 0;
    int cur_seed = 0;
    uint32_t nseed = 0;                 // resumed frames in seed[cur_seed]
    uint64_t *spill = nullptr;          // pinned host frames, mapped and coherent
    const uint64_t *spill_dev = nullptr;  // the same frames' device address
    size_t spill_cap = 0, spill_n = 0;  // in frames
    // last launch
    bool launched = false;
    uint32_t nwaves = 0, ring = 0, back = 0, dump_cap = 0, launch_todo = 0;
    uint32_t launch_seeds = 0;          // device seeds, then host frames [0, launch_seeds - nseed)
    hipEvent_t done_ev = nullptr;
    bool has_work() const { return todo_done < htodo.size() || nseed > 0 || spill_n > 0; }
    ~ClassState();
};

struct EpochStats {
    uint64_t frames = 0, donated = 0, spilled = 0, refilled = 0, narrowed = 0;
    bool paused = false;
};

struct Engine {
    int device = 0, cus = 0;
    uint64_t ticks_per_ms = 100000;    // device wall clock (s_memrealtime) rate
    hipStream_t stream = nullptr, side = nullptr;
    DevBuf jobs, acc;                  // JobDev table, JobAcc per job
    std::vector<kpd::JobDev> hjobs;
    size_t jobs_uploaded = 0;
    DevBuf blk, lft, info, slot_job, slot_pos, ws_cnt, ws_b1c, ws_a1, bdesc, wr_extra;
    size_t ws_cnt_zeroed = 0;          // bytes of ws_cnt known to be zero
    std::vector<DevBuf *> segs;        // arena segments owned by the engine (plain / batch)
    size_t segs_used = 0;              // segs[0, segs_used) hold the current tables' subgraphs
    DevBuf list, list_used;
    DevBuf t0;                         // shared start tick of the running epoch (time limit)
    ClassState cls[kNumClass];
    uint32_t *hpause = nullptr, *dpause = nullptr;   // pinned host pause word
    uint64_t total_spilled = 0, total_pauses = 0;
    // Derived-root buffers of narrowed frames.  An epoch fills one; its roots
    // stay in use until the one-word class has run an epoch through without a
    // pause and holds no pending frame (then every older buffer is free).
    struct NarBuf {
        DevBuf arena, frames;
        bool in_use = false;
    };
    NarBuf nar[kpd::NDERIVED_BUF];
    size_t nar_base = 0;               // one-word class table index of buffer 0's first derived root
    bool nar_base_set = false;         // (set once per run, after staging)
    DevBuf nar_ctr;
    int nar_cur = -1;                  // buffer filled by the running epoch (-1 = narrowing off)
    uint32_t nar_cap = 0;              // derived roots per buffer
    uint64_t total_narrowed = 0;
    int take_derived(EpochStats &es, bool c0_clean);

    int init();
    ~Engine();
    uint64_t device_bytes() const;
    // job table: make room for `n` jobs, set entry i (uploaded by sync_jobs)
    int reserve_jobs(size_t n);
    void set_job(uint32_t id, const GraphView &g, int k, int q);
    int sync_jobs();
    int zero_acc(uint32_t first, uint32_t count);
    // Decomposes and builds every root of `jobs`; appends the valid roots of
    // each class (largest first) to its root table and start list.  `alloc`
    // returns zeroed device memory for `words` 64-bit words of job index i.
    int stage(const std::vector<JobIn> &jobs, std::vector<JobStats> &st,
              const std::function<uint64_t *(size_t i, uint64_t words)> &alloc);
    // Monolithic per-root pass over the roots of `jobs` (root_kernel, one wave
    // per root, widths 1-4 words in turn).  Roots it completes are added to
    // the jobs' accumulators; the positions it leaves (blocks wider than four
    // words, subtrees over `budget` frames) are returned per job in `left`.
    int run_roots(const std::vector<JobIn> &jobs, uint64_t budget, uint64_t limit_ticks,
                  std::vector<std::vector<uint32_t>> &left);
    DevBuf rp_blk, rp_lft, rp_info, rp_wr, rp_arena, rp_desc, rp_stack, rp_def, rp_sched;
    // One epoch: launch every class with work; finish waits and carries pauses.
    int launch_epoch(const SearchOpts &o);
    int finish_epoch(const SearchOpts &o, EpochStats &es);
    bool has_work() const;
    void reset_tables();               // forget every root table, start list and pending frame
    // zeroed device memory for `words` 64-bit words of subgraph arena, from a
    // segment kept across runs (reset_tables frees them all for reuse)
    uint64_t *arena_alloc(uint64_t words);
    void release_arena();              // gives the segments back to the device
    int read_acc(uint32_t first, uint32_t count, std::vector<kpd::JobAcc> &out);
    // pending frames (device seeds and host spill) per job index; per_job must have room
    int count_pending(std::vector<uint64_t> &per_job);
    // allocates the pinned, device-mapped pause word that SearchOpts::host_pause polls
    int enable_host_pause();
    bool epoch_done();                 // every launched class has finished (does not block)
    // waits for the running epoch; once every class still running was launched
    // on a reduced grid (and another has finished), pauses them (regrid)
    int watch_epoch(const SearchOpts &o);
    // launches class c alone over its whole start list and resumed frames and
    // returns the kernel time; the class's tables are left as they were
    int time_class(int c, const SearchOpts &o, float *ms);
};

int dev_batch(const kp_graph *const *g, const kp_params *p, size_t n, kp_result *r);

// root positions [0, pn - q + 2) of a run, clipped to the window of *p
void root_window(uint32_t pn, const kp_params &p, int64_t &b, int64_t &e);
// LDS ring and HBM backing frames per wave of the per-root kernel at width W
void root_stack(int W, uint32_t &ring, uint32_t &back);
// launches the per-root kernel of width W (1, 2, 3 or 4 words) over ra's slots
int launch_roots_w(int W, hipStream_t st, const kpd::RootKernArgs &ra, uint32_t nwaves);
// fills `words` 64-bit words at p with v (a kernel launch, no runtime allocation)
int fill_words(hipStream_t st, void *p, uint64_t words, uint64_t v);

} // namespace kp
