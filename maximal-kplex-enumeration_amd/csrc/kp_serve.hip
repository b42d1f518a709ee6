// kp_serve.hip -- served mode: continuous batching of (graph, k, q) queries
// over the shared device pool (no reference counterpart: the reference runs
// one graph per process).
//
// Queries wait in a FIFO.  kp_server_poll is one scheduling step:
//   - an epoch (one launch per size class over every admitted query's roots
//     and resumed frames) that has ended is finished: frames of a paused epoch
//     are carried, queries with no unstarted root and no pending frame are
//     complete and leave with their result;
//   - waiting queries are admitted while a job slot is free and their
//     subgraph arena fits the memory budget: the (graph, q-k, truss)
//     preprocessing is done once and kept resident in HBM, then the roots are
//     decomposed and built into the arena;
//   - the next epoch starts with the carried frames and the new roots;
//   - while an epoch runs and queries wait, the epoch is paused once it is
//     slice_ms old, so the waiting queries join at the next step.
// Each query has its own job slot and accumulators, so its result does not
// depend on what else runs.
#include "kp_device.h"
#include "kp_engine.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <tuple>
#include <vector>

using namespace kp;

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) {                                                        \
            kp::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
            return KP_EHIP;                                                            \
        }                                                                              \
    } while (0)

namespace {

// First-fit allocator over one device buffer; freed ranges coalesce.
struct ArenaPool {
    DevBuf buf;
    uint64_t cap = 0, used = 0;
    std::map<uint64_t, uint64_t> free_;   // offset -> bytes
    int init(uint64_t bytes)
    {
        int rc = buf.ensure(bytes);
        if (rc) return rc;
        cap = bytes;
        free_.clear();
        free_[0] = bytes;
        return KP_OK;
    }
    uint64_t *alloc(uint64_t bytes)
    {
        bytes = (bytes + 255) & ~uint64_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const uint64_t off = it->first, sz = it->second;
            free_.erase(it);
            if (sz > bytes) free_[off + bytes] = sz - bytes;
            used += bytes;
            return reinterpret_cast<uint64_t *>(static_cast<char *>(buf.p) + off);
        }
        return nullptr;
    }
    void release(uint64_t *p, uint64_t bytes)
    {
        bytes = (bytes + 255) & ~uint64_t(255);
        uint64_t off = uint64_t(reinterpret_cast<char *>(p) - static_cast<char *>(buf.p));
        used -= bytes;
        auto nx = free_.lower_bound(off);
        if (nx != free_.end() && off + bytes == nx->first) {
            bytes += nx->second;
            free_.erase(nx);
        }
        auto pv = free_.lower_bound(off);
        if (pv != free_.begin()) {
            --pv;
            if (pv->first + pv->second == off) {
                pv->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }
};

struct Resident {          // one preprocessed graph kept in HBM
    Prep prep;
    DevBuf buf;
    GraphView gv;
    double prep_ms = 0, upload_ms = 0;
};

struct Query {
    uint64_t ticket;
    uint32_t graph;
    kp_params p;
    double t_submit;
};

struct Live {              // an admitted query in its job slot
    bool used = false;
    Query q{};
    Resident *res = nullptr;
    JobStats st;
    std::vector<std::pair<uint64_t *, uint64_t>> regions;   // arena ranges
    uint64_t roots_left = 0;
    uint64_t pauses0 = 0, spilled0 = 0;   // engine totals at admission
    double t_admit = 0, stage_ms = 0, prep_ms = 0;
};

} // namespace

struct kp_server {
    kp_server_config cfg{};
    Engine E;
    ArenaPool pool;
    SearchOpts opts;
    std::vector<kp_graph> graphs;
    std::map<std::tuple<uint32_t, int, int>, std::unique_ptr<Resident>> resident;
    std::deque<Query> pending;
    std::vector<Live> live;
    std::vector<uint32_t> free_slots;
    std::deque<std::pair<uint64_t, kp_result>> done;
    bool epoch_running = false, pause_sent = false;
    double epoch_t0 = 0;
    uint64_t next_ticket = 1;
    ~kp_server()
    {
        for (auto &g : graphs) kp_graph_free(&g);
    }
};

namespace {

int get_resident(kp_server *s, const Query &q, Resident **out, double *prep_ms)
{
    // the (q-k)-core and the (q-2k)-truss decide the preprocessed graph
    const auto key = std::make_tuple(q.graph, q.p.q - q.p.k, q.p.truss ? q.p.q - 2 * q.p.k : -1);
    auto it = s->resident.find(key);
    *prep_ms = 0;
    if (it != s->resident.end()) {
        *out = it->second.get();
        return KP_OK;
    }
    auto r = std::make_unique<Resident>();
    const double t0 = now_ms();
    int rc = prep_graph(&s->graphs[q.graph], q.p, &r->prep);
    if (rc) return rc;
    r->prep_ms = now_ms() - t0;
    const double t1 = now_ms();
    std::vector<GraphView> views;
    if ((rc = upload_graphs({&r->prep}, r->buf, views, s->E.stream))) return rc;
    r->gv = views[0];
    r->upload_ms = now_ms() - t1;
    *prep_ms = r->prep_ms + r->upload_ms;
    *out = r.get();
    s->resident[key] = std::move(r);
    return KP_OK;
}

void free_regions(kp_server *s, Live &L)
{
    for (auto &rg : L.regions) s->pool.release(rg.first, rg.second);
    L.regions.clear();
}

// A completed query leaves its slot with its result.
int complete(kp_server *s, uint32_t slot, double now)
{
    Live &L = s->live[slot];
    std::vector<kpd::JobAcc> acc;
    int rc = s->E.read_acc(slot, 1, acc);
    if (rc) return rc;
    kp_result r;
    std::memset(&r, 0, sizeof r);
    r.count = acc[0].count;
    r.hash_a = acc[0].ha;
    r.hash_b = acc[0].hb;
    r.tasks = acc[0].tasks;
    r.search_alg_bytes = acc[0].alg;
    r.alg_bytes = L.st.alg_bytes + acc[0].alg;
    r.adj_bytes = L.st.adj_bytes;
    r.valid_roots = L.st.valid_roots;
    r.max_n = L.st.max_n;
    r.max_l = L.st.max_l;
    r.pn = L.res->gv.pn;
    r.pm = L.res->gv.pm;
    r.pn_first = L.res->prep.pn_first;
    r.preprocess_ms = L.prep_ms;
    r.decompose_ms = L.stage_ms;
    r.search_ms = now - L.t_admit;
    r.wall_ms = now - L.q.t_submit;
    r.pauses = s->E.total_pauses - L.pauses0;
    r.spilled = s->E.total_spilled - L.spilled0;
    s->done.emplace_back(L.q.ticket, r);
    free_regions(s, L);
    L = Live{};
    s->free_slots.push_back(slot);
    return KP_OK;
}

// A query that cannot run leaves its slot with an error result (status = code,
// counts zero).
void fail_query(kp_server *s, uint32_t slot, int code, double now)
{
    Live &L = s->live[slot];
    kp_result r;
    std::memset(&r, 0, sizeof r);
    r.status = code;
    r.wall_ms = now - L.q.t_submit;
    s->done.emplace_back(L.q.ticket, r);
    free_regions(s, L);
    L = Live{};
    s->free_slots.push_back(slot);
}

// A query refused before admission (no slot) leaves with an error result.
void fail_waiting(kp_server *s, const Query &q, int code)
{
    kp_result r;
    std::memset(&r, 0, sizeof r);
    r.status = code;
    r.wall_ms = now_ms() - q.t_submit;
    s->done.emplace_back(q.ticket, r);
}

// Finishes the ended epoch: carries frames, retires complete queries.  If a
// search kernel reported an error, every admitted query fails: their waves
// may already have added partial counts to the accumulators, and resuming
// the epoch's roots would count them again.
int finish(kp_server *s)
{
    Engine &E = s->E;
    size_t before[kNumClass];
    for (int c = 0; c < kNumClass; c++) before[c] = E.cls[c].todo_done;
    EpochStats es;
    int rc = E.finish_epoch(s->opts, es);
    s->epoch_running = false;
    if (rc) {
        const double now = now_ms();
        for (uint32_t j = 0; j < s->live.size(); j++)
            if (s->live[j].used) fail_query(s, j, rc, now);
        E.reset_tables();
        return rc == KP_EHIP ? rc : KP_OK;
    }
    for (int c = 0; c < kNumClass; c++) {
        const ClassState &C = E.cls[c];
        for (size_t t = before[c]; t < C.todo_done; t++) {
            const uint32_t job = C.hdesc[C.htodo[t]].job;
            if (job < s->live.size() && s->live[job].roots_left) s->live[job].roots_left--;
        }
    }
    std::vector<uint64_t> pend(s->live.size(), 0);
    if ((rc = E.count_pending(pend))) return rc;
    const double now = now_ms();
    for (uint32_t j = 0; j < s->live.size(); j++)
        if (s->live[j].used && s->live[j].roots_left == 0 && pend[j] == 0)
            if ((rc = complete(s, j, now))) return rc;
    bool any = false;
    for (const Live &L : s->live) any |= L.used;
    if (!any && !E.has_work()) E.reset_tables();
    return KP_OK;
}

// Records a staged query's statistics at admission.
void admitted(kp_server *s, uint32_t slot, const JobStats &st, double stage_ms, double now)
{
    Live &L = s->live[slot];
    L.st = st;
    L.roots_left = st.valid_roots;
    L.stage_ms = stage_ms;
    L.t_admit = now;
    L.pauses0 = s->E.total_pauses;
    L.spilled0 = s->E.total_spilled;
}

// Stages the queries of `part` one at a time after their group failed for a
// reason other than arena memory.  A query that fails alone leaves with an
// error result; one that does not fit the arena goes back to the queue with
// every query after it.
int admit_one_by_one(kp_server *s, const std::vector<JobIn> &part, double t0)
{
    Engine &E = s->E;
    for (size_t i = 0; i < part.size(); i++) {
        const uint32_t slot = part[i].id;
        bool oom = false;
        auto alloc = [&](size_t, uint64_t words) -> uint64_t * {
            uint64_t *p = s->pool.alloc(8 * words);
            if (!p) {
                oom = true;
                return nullptr;
            }
            if (hipMemsetAsync(p, 0, 8 * words, E.stream) != hipSuccess) return nullptr;
            s->live[slot].regions.emplace_back(p, 8 * words);
            return p;
        };
        std::vector<JobStats> st;
        const int rc = E.stage({part[i]}, st, alloc);
        const double now = now_ms();
        if (rc == KP_EHIP) return rc;
        if (rc && oom) {
            // this query and the ones after it wait for arena memory
            for (size_t j = part.size(); j-- > i;) {
                Live &L = s->live[part[j].id];
                free_regions(s, L);
                s->pending.push_front(L.q);
                s->free_slots.push_back(part[j].id);
                L = Live{};
            }
            return KP_OK;
        }
        if (rc) {
            fail_query(s, slot, rc, now);
            continue;
        }
        admitted(s, slot, st[0], now - t0, now);
        if (st[0].valid_roots == 0) {
            int rc2 = complete(s, slot, now);
            if (rc2) return rc2;
        }
    }
    return KP_OK;
}

// Admits waiting queries while slots and arena memory allow.
int admit(kp_server *s)
{
    Engine &E = s->E;
    std::vector<uint32_t> slots;
    std::vector<JobIn> jobs;
    while (!s->pending.empty() && !s->free_slots.empty()) {
        const Query q = s->pending.front();
        Resident *res = nullptr;
        double pms = 0;
        int rc = get_resident(s, q, &res, &pms);
        if (rc) {
            // the query cannot run (its preprocessing failed): it leaves the
            // queue with an error result
            fail_waiting(s, q, rc);
            s->pending.pop_front();
            continue;
        }
        const uint32_t slot = s->free_slots.back();
        s->free_slots.pop_back();
        s->pending.pop_front();
        Live &L = s->live[slot];
        L = Live{};
        L.used = true;
        L.q = q;
        L.res = res;
        L.prep_ms = pms;
        E.set_job(slot, res->gv, q.p.k, q.p.q);
        JobIn j{slot, res->gv, q.p.k, q.p.q, 0, std::max<int64_t>(0, int64_t(res->gv.pn) - q.p.q + 2)};
        jobs.push_back(j);
        slots.push_back(slot);
    }
    if (jobs.empty()) return KP_OK;
    int rc;
    if ((rc = E.sync_jobs())) return rc;
    for (uint32_t slot : slots)
        if ((rc = E.zero_acc(slot, 1))) return rc;
    const double t0 = now_ms();
    // Stage the group; if the arena cannot hold it, the queries that do not
    // fit go back to the front of the queue, one by one from the end.  Any
    // other failure is isolated by staging the queries one at a time: a query
    // that fails alone leaves with an error result, the others run.
    size_t take = jobs.size();
    std::vector<JobStats> st;
    for (;;) {
        std::vector<JobIn> part(jobs.begin(), jobs.begin() + take);
        bool oom = false;
        auto alloc = [&](size_t i, uint64_t words) -> uint64_t * {
            uint64_t *p = s->pool.alloc(8 * words);
            if (!p) {
                oom = true;
                return nullptr;
            }
            if (hipMemsetAsync(p, 0, 8 * words, E.stream) != hipSuccess) return nullptr;
            s->live[part[i].id].regions.emplace_back(p, 8 * words);
            return p;
        };
        rc = E.stage(part, st, alloc);
        if (!rc) break;
        for (size_t i = 0; i < take; i++) free_regions(s, s->live[part[i].id]);
        if (rc == KP_EHIP) return rc;
        if (!oom) return admit_one_by_one(s, part, t0);
        if (take == 0) return rc;
        // put the last query of the group back
        Live &L = s->live[jobs[take - 1].id];
        s->pending.push_front(L.q);
        s->free_slots.push_back(jobs[take - 1].id);
        L = Live{};
        take--;
        if (take == 0) {
            bool any = false;
            for (const Live &x : s->live) any |= x.used;
            if (!any) {
                set_error("query %llu needs more subgraph arena than the server budget of %llu bytes",
                          (unsigned long long)s->pending.front().ticket, (unsigned long long)s->pool.cap);
                fail_waiting(s, s->pending.front(), KP_ENOMEM);
                s->pending.pop_front();
                return KP_OK;
            }
            return KP_OK;   // waits until running queries free the arena
        }
    }
    const double ms = now_ms() - t0;
    const double now = now_ms();
    for (size_t i = 0; i < take; i++) admitted(s, jobs[i].id, st[i], ms, now);
    for (size_t i = 0; i < take; i++)
        if (s->live[jobs[i].id].roots_left == 0)
            if ((rc = complete(s, jobs[i].id, now))) return rc;
    return KP_OK;
}

} // namespace

extern "C" int kp_server_create(const kp_server_config *cfg, kp_server **out)
{
    if (!out) return KP_EINVAL;
    *out = nullptr;
    auto *s = new (std::nothrow) kp_server();
    if (!s) return KP_ENOMEM;
    if (cfg) s->cfg = *cfg;
    if (!s->cfg.max_inflight) s->cfg.max_inflight = 64;
    if (!s->cfg.mem_budget_bytes) s->cfg.mem_budget_bytes = 4ull << 30;
    // Each admission pause dumps and relaunches the running classes; on the
    // served workload (281 queries at 200/s) the makespan was 4.60 s at 2 ms,
    // 4.33 s at 10 ms and 4.30 s at 30 ms (p50 3.05 / 2.75 / 2.50 s).
    if (s->cfg.slice_ms <= 0) s->cfg.slice_ms = 30.0;
    int rc;
    if ((rc = s->E.init()) || (rc = s->E.enable_host_pause()) || (rc = s->E.reserve_jobs(s->cfg.max_inflight)) ||
        (rc = s->pool.init(s->cfg.mem_budget_bytes))) {
        delete s;
        return rc;
    }
    s->live.resize(s->cfg.max_inflight);
    for (size_t i = s->cfg.max_inflight; i-- > 0;) s->free_slots.push_back(uint32_t(i));
    s->opts.pausable = true;
    s->opts.host_pause = true;
    s->opts.seed_budget = 1ull << 20;
    *out = s;
    return KP_OK;
}

extern "C" int kp_server_add_graph(kp_server *s, const kp_graph *g, uint32_t *graph_id)
{
    if (!s || !g || !graph_id) return KP_EINVAL;
    kp_graph c{};
    int rc = kp_graph_from_csr(g->n, g->deg, g->nbr, &c);
    if (rc) return rc;
    *graph_id = uint32_t(s->graphs.size());
    s->graphs.push_back(c);
    return KP_OK;
}

extern "C" int kp_server_preload(kp_server *s, const uint32_t *graph_ids, const kp_params *p, size_t n)
{
    if (!s || (n && (!graph_ids || !p))) return KP_EINVAL;
    using Key = std::tuple<uint32_t, int, int>;
    std::vector<Key> keys;
    std::vector<size_t> which;
    for (size_t i = 0; i < n; i++) {
        if (graph_ids[i] >= s->graphs.size() || p[i].k < 1 || p[i].q < 2 * p[i].k) return KP_EINVAL;
        const Key key = std::make_tuple(graph_ids[i], p[i].q - p[i].k, p[i].truss ? p[i].q - 2 * p[i].k : -1);
        if (s->resident.count(key) || std::find(keys.begin(), keys.end(), key) != keys.end()) continue;
        keys.push_back(key);
        which.push_back(i);
    }
    std::vector<std::unique_ptr<Resident>> made(keys.size());
    std::vector<int> rcs(keys.size(), KP_OK);
    // host threads take graphs from a shared counter (each preprocessing run
    // uses OpenMP inside as well)
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t t; (t = next.fetch_add(1)) < keys.size();) {
            const size_t i = which[t];
            made[t] = std::make_unique<Resident>();
            const double t0 = now_ms();
            rcs[t] = prep_graph(&s->graphs[graph_ids[i]], p[i], &made[t]->prep);
            made[t]->prep_ms = now_ms() - t0;
        }
    };
    const size_t nt = std::min<size_t>(keys.size(), std::max(1u, std::thread::hardware_concurrency() / 4));
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    for (size_t t = 0; t < keys.size(); t++) {
        if (rcs[t]) return rcs[t];
        std::vector<GraphView> views;
        const double t1 = now_ms();
        int rc = upload_graphs({&made[t]->prep}, made[t]->buf, views, s->E.stream);
        if (rc) return rc;
        made[t]->gv = views[0];
        made[t]->upload_ms = now_ms() - t1;
        s->resident[keys[t]] = std::move(made[t]);
    }
    return KP_OK;
}

extern "C" int kp_server_submit(kp_server *s, uint32_t graph_id, const kp_params *p, uint64_t *ticket)
{
    if (!s || !p || !ticket || graph_id >= s->graphs.size()) return KP_EINVAL;
    if (p->k < 1 || p->q < 2 * p->k || p->q > int(KP_MAX_BLOCK)) {
        set_error("invalid parameters k=%d q=%d (need k >= 1 and 2k <= q <= %d)", p->k, p->q, KP_MAX_BLOCK);
        return KP_EINVAL;
    }
    if ((p->flags & KP_FLAG_LIST) || p->root_begin || p->root_end) {
        set_error("the server returns counts and hashes of whole graphs (no lists, no root windows)");
        return KP_EINVAL;
    }
    *ticket = s->next_ticket++;
    s->pending.push_back({*ticket, graph_id, *p, now_ms()});
    return KP_OK;
}

extern "C" int kp_server_poll(kp_server *s, uint64_t *tickets, kp_result *results, size_t cap, size_t *ndone)
{
    if (!s || !ndone || (cap && (!tickets || !results))) return KP_EINVAL;
    *ndone = 0;
    int rc = KP_OK;
    if (hipSetDevice(s->E.device) != hipSuccess) return KP_EHIP;
    if (s->epoch_running) {
        if (!s->E.epoch_done()) {
            if (!s->pending.empty() && !s->free_slots.empty() && !s->pause_sent &&
                now_ms() - s->epoch_t0 >= s->cfg.slice_ms) {
                __atomic_store_n(s->E.hpause, 1u, __ATOMIC_SEQ_CST);
                s->pause_sent = true;
            }
        } else {
            rc = finish(s);
        }
    }
    if (!rc && !s->epoch_running) {
        rc = admit(s);
        if (!rc && s->E.has_work()) {
            __atomic_store_n(s->E.hpause, 0u, __ATOMIC_SEQ_CST);
            s->pause_sent = false;
            rc = s->E.launch_epoch(s->opts);
            if (!rc) {
                s->epoch_running = true;
                s->epoch_t0 = now_ms();
            }
        }
    }
    while (*ndone < cap && !s->done.empty()) {
        tickets[*ndone] = s->done.front().first;
        results[*ndone] = s->done.front().second;
        s->done.pop_front();
        (*ndone)++;
    }
    return rc;
}

extern "C" size_t kp_server_outstanding(const kp_server *s)
{
    if (!s) return 0;
    size_t n = s->pending.size() + s->done.size();
    for (const Live &L : s->live) n += L.used;
    return n;
}

extern "C" int kp_server_destroy(kp_server *s)
{
    if (!s) return KP_OK;
    if (s->epoch_running) {
        __atomic_store_n(s->E.hpause, 1u, __ATOMIC_SEQ_CST);
        for (int c = 0; c < kNumClass; c++)
            if (s->E.cls[c].launched) (void)hipStreamSynchronize(s->E.cls[c].stream);
    }
    delete s;
    return KP_OK;
}
