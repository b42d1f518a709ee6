// kp_device.hip -- device driver of libkplex: engine, launches, prepared
// graphs, plain and batched runs (the kernels are in kp_kernels.h).
#include <hip/hip_runtime.h>

#include "kp_device.h"
#include "kp_engine.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "kp_kernels.h"

namespace kpd {
// The one-word list-free search kernel for epochs whose primary class is the
// one-word class is compiled in kp_search1.hip.
extern template __global__ void search_kernel<1, 2, false, 1>(SearchArgs);
} // namespace kpd


// ======================================================================
// host side: the engine (kp_engine.h), prepared graphs, plain and batched runs
// ======================================================================
using namespace kpd;

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) {                                                        \
            kp::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
            return KP_EHIP;                                                            \
        }                                                                              \
    } while (0)

namespace kp {

const int kClassW[kNumClass] = {1, 2, 3, 4, 6, 8, 16};
ClassTune g_tune[kNumClass];

int class_of(uint32_t n)
{
    const uint32_t w = (n + 63) / 64;
    for (int c = 0; c < kNumClass; c++)
        if (uint32_t(kClassW[c]) >= w) return c;
    return -1;
}

int DevBuf::ensure(size_t need)
{
    if (need <= bytes && p) return KP_OK;
    release();
    if (need == 0) need = 8;
    if (hipMalloc(&p, need) != hipSuccess) {
        set_error("hipMalloc(%zu) failed", need);
        p = nullptr;
        return KP_ENOMEM;
    }
    bytes = need;
    return KP_OK;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

namespace {

size_t al256(size_t x) { return (x + 255) & ~size_t(255); }

// per-root pass of plain runs: graphs of at most kMonoMaxPn preprocessed
// vertices; a root whose subtree passes kMonoBudget frames goes to the pool
constexpr uint32_t kMonoMaxPn = 4096;
constexpr uint64_t kMonoBudget = 4096;

// the reference's task-queue capacity MAX_CAP = 2048 * 2048 (inc/common.h:11,
// the overflow queue of inc/host_funcs.h:982), the default for
// kp_params.queue_capacity
constexpr uint64_t kDefaultQueueCap = 4194304;

} // namespace

int prep_graph(const kp_graph *g, const kp_params &p, Prep *out)
{
    if (p.flags & KP_FLAG_GPU_PREP) return gpu_preprocess(g, p.k, p.q, p.truss, out);
    return preprocess(g, p.k, p.q, p.truss, out);
}

size_t packed_graph_bytes(const Prep &p)
{
    const size_t n = p.g.n, m = p.g.m;
    return al256(8 * (n + 1)) + al256(4 * m) + 4 * al256(4 * n);
}

int upload_graphs(const std::vector<const Prep *> &preps, DevBuf &buf, std::vector<GraphView> &views, hipStream_t st)
{
    std::vector<size_t> base(preps.size());
    size_t total = 0;
    for (size_t i = 0; i < preps.size(); i++) {
        base[i] = total;
        total += packed_graph_bytes(*preps[i]);
    }
    int rc = buf.ensure(std::max<size_t>(total, 256));
    if (rc) return rc;
    std::vector<uint8_t> host(total);
    views.assign(preps.size(), GraphView{});
#pragma omp parallel for schedule(dynamic, 1) if (preps.size() > 1)
    for (size_t i = 0; i < preps.size(); i++) {
        const Prep &P = *preps[i];
        const kp_graph &G = P.g;
        const size_t n = G.n, m = G.m;
        uint8_t *h = host.data() + base[i];
        const char *d = static_cast<const char *>(buf.p) + base[i];
        size_t o = 0;
        auto put = [&](const void *src, size_t bytes) -> const char * {
            const size_t at = o;
            if (bytes && src) std::memcpy(h + o, src, bytes);
            else if (bytes) std::memset(h + o, 0, bytes);
            o += al256(bytes);
            return d + at;
        };
        GraphView &v = views[i];
        v.off = reinterpret_cast<const uint64_t *>(put(G.off, 8 * (n + 1)));
        v.nbr = reinterpret_cast<const uint32_t *>(put(G.nbr, 4 * m));
        v.deg = reinterpret_cast<const uint32_t *>(put(G.deg, 4 * n));
        v.dpos = reinterpret_cast<const int32_t *>(put(P.dpos.data(), 4 * n));
        v.dseq = reinterpret_cast<const int32_t *>(put(P.dseq.data(), 4 * n));
        v.orig = reinterpret_cast<const uint32_t *>(put(P.orig.data(), 4 * n));
        v.pn = uint32_t(n);
        v.pm = m;
        uint32_t maxdeg = P.maxdeg, d1 = P.d1max;
        if (!P.sizes_known) {
            maxdeg = d1 = 0;
            for (uint32_t u = 0; u < G.n; u++) {
                maxdeg = std::max(maxdeg, G.deg[u]);
                uint32_t later = 0;
                for (uint64_t e = G.off[u]; e < G.off[u + 1]; e++) later += P.dpos[G.nbr[e]] > P.dpos[u];
                d1 = std::max(d1, later);
            }
        }
        v.maxdeg = maxdeg;
        v.d1max = d1;
    }
    if (total) HIPCHK(hipMemcpyAsync(buf.p, host.data(), total, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return KP_OK;
}

ClassState::~ClassState()
{
    if (spill) (void)hipHostFree(spill);
    if (done_ev) (void)hipEventDestroy(done_ev);
}

namespace {

// The streams of one device, created once per process and shared by every
// engine on it.  HIP maps streams onto its hardware queues (GPU_MAX_HW_QUEUES,
// 4 by default) round-robin in creation order, so the streams of the four
// narrowest size classes, which carry nearly all the work, are created first
// and get a queue each; classes on one queue would run one after the other.
// Engines on one device share these streams: their launches are ordered on
// them, never lost.
struct DeviceStreams {
    bool ready = false;
    int cus = 0;
    uint64_t ticks_per_ms = 100000;
    hipStream_t cls[kNumClass] = {};
    hipStream_t main = nullptr, side = nullptr;
};
std::mutex g_streams_mu;
DeviceStreams g_streams[64];

int device_streams(int device, DeviceStreams **out)
{
    if (device < 0 || device >= 64) return KP_ENODEV;
    std::lock_guard<std::mutex> lk(g_streams_mu);
    DeviceStreams &d = g_streams[device];
    if (!d.ready) {
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        d.cus = prop.multiProcessorCount;
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
            d.ticks_per_ms = uint64_t(khz);
        for (int c = 0; c < kNumClass; c++) HIPCHK(hipStreamCreateWithFlags(&d.cls[c], hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&d.main, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&d.side, hipStreamNonBlocking));
        d.ready = true;
    }
    *out = &d;
    return KP_OK;
}

} // namespace

int Engine::init()
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        set_error("no HIP device visible: the HIP backend is required");
        return KP_ENODEV;
    }
    HIPCHK(hipGetDevice(&device));
    DeviceStreams *ds = nullptr;
    int rc = device_streams(device, &ds);
    if (rc) return rc;
    cus = ds->cus;
    ticks_per_ms = ds->ticks_per_ms;
    stream = ds->main;
    side = ds->side;
    for (int c = 0; c < kNumClass; c++) {
        cls[c].stream = ds->cls[c];
        if (!cls[c].done_ev) HIPCHK(hipEventCreateWithFlags(&cls[c].done_ev, hipEventDisableTiming));
    }
    return KP_OK;
}

Engine::~Engine()
{
    for (DevBuf *b : segs) delete b;
    if (hpause) (void)hipHostFree(hpause);
}

uint64_t Engine::device_bytes() const
{
    const DevBuf *all[] = {&jobs, &acc, &blk, &lft, &info, &slot_job, &slot_pos, &ws_cnt, &ws_b1c, &ws_a1,
                           &bdesc, &wr_extra, &list, &list_used, &nar_ctr};
    uint64_t s = 0;
    for (const DevBuf *b : all) s += b->bytes;
    for (const NarBuf &b : nar) s += b.arena.bytes + b.frames.bytes;
    for (const DevBuf *b : segs) s += b->bytes;
    for (const auto &c : cls) {
        const DevBuf *cb[] = {&c.q, &c.qseq, &c.sched, &c.stack, &c.subq, &c.desc, &c.todo, &c.seed[0], &c.seed[1], &c.dump};
        for (const DevBuf *b : cb) s += b->bytes;
    }
    return s;
}

int Engine::reserve_jobs(size_t n)
{
    n = std::max<size_t>(n, 1);
    if (hjobs.size() < n) hjobs.resize(n, JobDev{});
    int rc;
    if ((rc = jobs.ensure(sizeof(JobDev) * hjobs.size())) || (rc = acc.ensure(sizeof(JobAcc) * hjobs.size()))) return rc;
    return KP_OK;
}

void Engine::set_job(uint32_t id, const GraphView &g, int k, int q)
{
    hjobs[id] = JobDev{g.off, g.nbr, g.deg, g.orig, g.dpos, g.dseq, g.pn, k, q, 0};
}

int Engine::sync_jobs()
{
    HIPCHK(hipMemcpyAsync(jobs.p, hjobs.data(), sizeof(JobDev) * hjobs.size(), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return KP_OK;
}

int Engine::zero_acc(uint32_t first, uint32_t count)
{
    if (count) HIPCHK(hipMemsetAsync(acc.as<JobAcc>() + first, 0, sizeof(JobAcc) * count, stream));
    return KP_OK;
}

int Engine::read_acc(uint32_t first, uint32_t count, std::vector<JobAcc> &out)
{
    out.assign(count, JobAcc{});
    if (!count) return KP_OK;
    HIPCHK(hipMemcpyAsync(out.data(), acc.as<JobAcc>() + first, sizeof(JobAcc) * count, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return KP_OK;
}

bool Engine::has_work() const
{
    for (const auto &c : cls)
        if (c.has_work()) return true;
    return false;
}

void Engine::reset_tables()
{
    for (auto &c : cls) {
        c.hdesc.clear();
        c.htodo.clear();
        c.todo_done = 0;
        c.desc_uploaded = 0;
        c.maxn = 0;
        c.nseed = 0;
        c.spill_n = 0;
        c.launched = false;
    }
    segs_used = 0;
    for (auto &b : nar) b.in_use = false;
    nar_base_set = false;
}

uint64_t *Engine::arena_alloc(uint64_t words)
{
    const size_t need = 8 * std::max<uint64_t>(words, 1);
    size_t pick = segs.size();
    for (size_t i = segs_used; i < segs.size(); i++)
        if (segs[i]->bytes >= need && (pick == segs.size() || segs[i]->bytes < segs[pick]->bytes)) pick = i;
    if (pick == segs.size()) {
        auto *b = new (std::nothrow) DevBuf();
        if (!b) return nullptr;
        if (b->ensure(need)) {
            delete b;
            return nullptr;
        }
        segs.push_back(b);
    }
    std::swap(segs[pick], segs[segs_used]);
    DevBuf *b = segs[segs_used++];
    if (hipMemsetAsync(b->p, 0, need, stream) != hipSuccess) return nullptr;
    return b->as<uint64_t>();
}

void Engine::release_arena()
{
    for (DevBuf *b : segs) delete b;
    segs.clear();
    segs_used = 0;
}

namespace {

template <int WC>
int launch_build(hipStream_t st, const BuildArgs &ba, uint32_t nwg)
{
    hipLaunchKernelGGL((build_kernel<WC>), dim3(nwg), dim3(256), 0, st, ba);
    HIPCHK(hipGetLastError());
    return KP_OK;
}

int run_build_class(hipStream_t st, int c, const BuildArgs &ba, uint32_t nwg)
{
    switch (kClassW[c]) {
    case 1: return launch_build<1>(st, ba, nwg);
    case 2: return launch_build<2>(st, ba, nwg);
    case 3: return launch_build<3>(st, ba, nwg);
    case 4: return launch_build<4>(st, ba, nwg);
    case 6: return launch_build<6>(st, ba, nwg);
    case 8: return launch_build<8>(st, ba, nwg);
    case 16: return launch_build<16>(st, ba, nwg);
    }
    return KP_EINVAL;
}

} // namespace

int Engine::stage(const std::vector<JobIn> &jv, std::vector<JobStats> &st,
                  const std::function<uint64_t *(size_t i, uint64_t words)> &alloc)
{
    st.assign(jv.size(), JobStats{});
    size_t S = 0;
    uint32_t lcap = 1, d1cap = 1, stride = 1;
    for (const JobIn &j : jv) {
        S += j.pos_list ? j.pos_list->size() : size_t(std::max<int64_t>(0, j.pos_end - j.pos_begin));
        const uint32_t pn1 = std::max<uint32_t>(j.g.pn, 1);
        lcap = std::max(lcap, std::min<uint32_t>(pn1, std::max<uint32_t>(2048, 2 * j.g.maxdeg + 1024)));
        d1cap = std::max(d1cap, j.g.d1max);
        stride = std::max(stride, pn1);
    }
    if (S == 0) return KP_OK;
    const uint32_t w1cap = (d1cap + 63) / 64;
    std::vector<uint32_t> slot_i, slot_id, slot_p;
    slot_i.reserve(S);
    slot_id.reserve(S);
    slot_p.reserve(S);
    for (size_t i = 0; i < jv.size(); i++) {
        auto add = [&](uint32_t pos) {
            slot_i.push_back(uint32_t(i));
            slot_id.push_back(jv[i].id);
            slot_p.push_back(pos);
        };
        if (jv[i].pos_list)
            for (uint32_t pos : *jv[i].pos_list) add(pos);
        else
            for (int64_t pos = jv[i].pos_begin; pos < jv[i].pos_end; pos++) add(uint32_t(pos));
    }
    const size_t per_slot = size_t(NCAP + lcap) * 4;
    const size_t chunk = std::max<size_t>(1, std::min<size_t>({S, 65536, (size_t(2) << 30) / per_slot}));
    uint32_t nwaves = uint32_t(std::min<size_t>(chunk, size_t(cus) * 16));
    nwaves = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(nwaves, (8ull << 30) / (uint64_t(stride) * 4))));
    int rc;
    const size_t cnt_bytes = size_t(nwaves) * stride * 4;
    if ((rc = blk.ensure(chunk * NCAP * 4)) || (rc = lft.ensure(chunk * lcap * 4)) ||
        (rc = info.ensure(chunk * sizeof(RootInfo))) || (rc = slot_job.ensure(chunk * 4)) ||
        (rc = slot_pos.ensure(chunk * 4)) || (rc = ws_b1c.ensure(size_t(nwaves) * d1cap * 4)) ||
        (rc = ws_a1.ensure(size_t(nwaves) * d1cap * w1cap * 8)))
        return rc;
    {
        void *before = ws_cnt.p;
        if ((rc = ws_cnt.ensure(cnt_bytes))) return rc;
        if (ws_cnt.p != before || cnt_bytes > ws_cnt_zeroed) {
            HIPCHK(hipMemsetAsync(ws_cnt.p, 0, ws_cnt.bytes, stream));
            ws_cnt_zeroed = ws_cnt.bytes;
        }
    }
    std::vector<RootInfo> hinfo(chunk);
    std::vector<RootDesc> newd[kNumClass];
    struct Pend {
        int c;
        size_t i;
        uint64_t words;
    };
    for (size_t c0 = 0; c0 < S; c0 += chunk) {
        const size_t cn = std::min(chunk, S - c0);
        HIPCHK(hipMemcpyAsync(slot_job.p, slot_id.data() + c0, 4 * cn, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(slot_pos.p, slot_p.data() + c0, 4 * cn, hipMemcpyHostToDevice, stream));
        const uint32_t nw = uint32_t(std::min<size_t>(nwaves, cn));
        DecompArgs da{};
        da.jobs = jobs.as<JobDev>();
        da.slot_job = slot_job.as<uint32_t>();
        da.slot_pos = slot_pos.as<uint32_t>();
        da.nslots = uint32_t(cn);
        da.blk = blk.as<uint32_t>();
        da.lft = lft.as<uint32_t>();
        da.lcap = lcap;
        da.info = info.as<RootInfo>();
        da.ws_cnt = ws_cnt.as<uint32_t>();
        da.ws_b1c = ws_b1c.as<uint32_t>();
        da.ws_a1 = ws_a1.as<uint64_t>();
        da.cnt_stride = stride;
        da.d1cap = d1cap;
        da.w1cap = w1cap;
        da.nwaves = nw;
        hipLaunchKernelGGL(decompose_kernel, dim3((nw + WPB - 1) / WPB), dim3(256), 0, stream, da);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(hinfo.data(), info.p, sizeof(RootInfo) * cn, hipMemcpyDeviceToHost, stream));
        if (hipStreamSynchronize(stream) != hipSuccess) {
            ws_cnt_zeroed = 0;
            set_error("decompose kernel failed");
            return KP_EHIP;
        }
        // classes, per-job arena sizes
        std::vector<RootDesc> cd[kNumClass];
        std::vector<Pend> pend[kNumClass];
        std::vector<uint64_t> jwords(jv.size(), 0);
        for (size_t s = 0; s < cn; s++) {
            const RootInfo &ri = hinfo[s];
            const size_t i = slot_i[c0 + s];
            if (ri.status < 0) {
                if (ri.status == -1) {
                    set_error("job %zu, root at degeneracy position %u: block exceeds %u vertices", i, slot_p[c0 + s], NCAP);
                    return KP_ERANGE;
                }
                set_error("job %zu, root at degeneracy position %u: left set exceeds %u vertices", i, slot_p[c0 + s], lcap);
                return KP_EOVERFLOW;
            }
            if (ri.status != 1) continue;
            JobStats &js = st[i];
            js.valid_roots++;
            js.max_n = std::max(js.max_n, ri.n);
            js.max_l = std::max(js.max_l, ri.l);
            const int c = class_of(ri.n);
            const uint64_t W = uint64_t(kClassW[c]);
            const uint64_t words = 3ull * ri.n * W + uint64_t(ri.l) * W + (ri.n + 1) / 2;
            RootDesc d{};
            d.n = ri.n;
            d.h = ri.h;
            d.l = ri.l;
            d.slot = uint32_t(s);
            d.job = jv[i].id;
            d.k = uint16_t(jv[i].k);
            d.q = uint16_t(jv[i].q);
            cd[c].push_back(d);
            pend[c].push_back({c, i, words});
            jwords[i] += words;
        }
        std::vector<uint64_t *> jbase(jv.size(), nullptr), jcur(jv.size(), nullptr);
        for (size_t i = 0; i < jv.size(); i++)
            if (jwords[i]) {
                jbase[i] = alloc(i, jwords[i]);
                if (!jbase[i]) {
                    set_error("arena allocation of %llu bytes failed", (unsigned long long)(8 * jwords[i]));
                    return KP_ENOMEM;
                }
                jcur[i] = jbase[i];
                st[i].arena_bytes += 8 * jwords[i];
            }
        std::vector<RootDesc> all;
        std::vector<uint32_t> cbeg(kNumClass + 1, 0);
        for (int c = 0; c < kNumClass; c++) {
            cbeg[c] = uint32_t(all.size());
            for (size_t t = 0; t < cd[c].size(); t++) {
                RootDesc d = cd[c][t];
                d.addr = reinterpret_cast<uint64_t>(jcur[pend[c][t].i]);
                jcur[pend[c][t].i] += pend[c][t].words;
                all.push_back(d);
            }
        }
        cbeg[kNumClass] = uint32_t(all.size());
        const size_t nd = all.size();
        std::vector<uint64_t> wr(nd);
        if (nd) {
            if ((rc = bdesc.ensure(sizeof(RootDesc) * nd)) || (rc = wr_extra.ensure(8 * nd))) return rc;
            HIPCHK(hipMemcpyAsync(bdesc.p, all.data(), sizeof(RootDesc) * nd, hipMemcpyHostToDevice, stream));
            for (int c = 0; c < kNumClass; c++) {
                const uint32_t b = cbeg[c], e = cbeg[c + 1];
                if (b == e) continue;
                BuildArgs ba{};
                ba.jobs = jobs.as<JobDev>();
                ba.blk = blk.as<uint32_t>();
                ba.lft = lft.as<uint32_t>();
                ba.lcap = lcap;
                ba.desc = bdesc.as<RootDesc>() + b;
                ba.ndesc = e - b;
                ba.ws_cnt = ws_cnt.as<uint32_t>();
                ba.cnt_stride = stride;
                ba.nwaves = std::min<uint32_t>(nw, ((e - b) + WPB - 1) / WPB * WPB);
                ba.wr_extra = wr_extra.as<uint64_t>() + b;
                if ((rc = run_build_class(stream, c, ba, (ba.nwaves + WPB - 1) / WPB))) return rc;
            }
            HIPCHK(hipMemcpyAsync(wr.data(), wr_extra.p, 8 * nd, hipMemcpyDeviceToHost, stream));
            if (hipStreamSynchronize(stream) != hipSuccess) {
                ws_cnt_zeroed = 0;
                set_error("build kernel failed");
                return KP_EHIP;
            }
        }
        for (int c = 0; c < kNumClass; c++)
            for (uint32_t t = cbeg[c]; t < cbeg[c + 1]; t++) {
                const RootDesc &d = all[t];
                const RootInfo &ri = hinfo[d.slot];
                JobStats &js = st[pend[c][t - cbeg[c]].i];
                const uint64_t w = ri.wr + wr[t];
                js.adj_bytes += 4 * w;
                js.alg_bytes += 4 * w + (uint64_t(ri.n) * ri.n + 7) / 8 + uint64_t(ri.n) * ri.n;
                newd[c].push_back(d);
            }
    }
    // largest blocks first within what this call adds (longest-processing-time order)
    for (int c = 0; c < kNumClass; c++) {
        auto &v = newd[c];
        if (v.empty()) continue;
        std::stable_sort(v.begin(), v.end(), [](const RootDesc &x, const RootDesc &y) { return x.n > y.n; });
        ClassState &C = cls[c];
        for (const RootDesc &d : v) {
            C.htodo.push_back(uint32_t(C.hdesc.size()));
            C.hdesc.push_back(d);
            C.maxn = std::max(C.maxn, d.n);
        }
    }
    return KP_OK;
}

namespace {

// Untuned, a one-word class with at least this many pending roots (and
// frames) but fewer than a full grid has waves runs six blocks per CU rather
// than four: enough roots that the extra waves find work (75k window, 2,094
// roots: 179 -> 163-167 ms with the longer donation period below).
constexpr uint64_t kManyRoots = 1024;

template <int W, int LW>
int launch_class(Engine &E, int c, const SearchOpts &o, bool secondary, bool alone)
{
    ClassState &C = E.cls[c];
    constexpr int FW = FrameLayout<W>::FW;
    const ClassTune &T = g_tune[c];
    // DFS depth bound (every child shrinks C2 or C), plus the resumed frames a
    // wave claims at once below its first
    const uint32_t need = C.maxn + 3 + kSeedChunk;
    // LDS ring: 8 KiB of frames per wave unless tuned; at most 64 KiB per block
    const uint32_t ring_max = std::max<uint32_t>(4, (64 * 1024) / (FW * 8 * 4));
    uint32_t ring = T.ring_frames ? std::max<uint32_t>(4, T.ring_frames) : std::max<uint32_t>(16, 8192 / (FW * 8));
    if (W == 1) ring = need;   // the one-word kernel keeps its whole stack in the ring
    ring = std::min(std::min(ring, ring_max), need);
    const uint32_t back = need > ring ? need - ring + 1 : 0;
    const size_t dyn = size_t(ring) * FW * 8 * 4;
    auto kern = o.want_list ? search_kernel<W, LW, true> : search_kernel<W, LW, false>;
    // The one-word class as the primary class takes the build of
    // kp_search1.hip (75k window 223 -> 190 ms); beside wider primary classes
    // it keeps this one (email-euall: 173 ms, against 187 ms with the other).
    if constexpr (W == 1)
        if (!secondary && !o.want_list) kern = search_kernel<1, LW, false, 1>;
    int nb = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 256, dyn));
    if (nb < 1) nb = 1;
    C.nb_max = nb;
    // Classes launch together on their own streams, and idle waves of a class
    // that has nothing left to start retire, so that blocks of the other
    // classes take their slots.  Untuned, when the epoch's primary class (the
    // one with the most pending roots) is a wide one (two or more words, so at
    // most two blocks per CU and VGPRs to spare), each other class gets one
    // block per CU and runs beside it.  Measured: email-euall with the
    // two-word class primary 150 ms this way against 250 ms with every class
    // on its full grid; a window of the Chung-Lu 75k graph whose one-word
    // class holds most frames, 12.6 s with the one-word class primary against
    // 29 s with a wider class primary (chosen then by the sum of n^2).
    // Untuned, a class with fewer pending roots and frames than the waves of
    // a full grid runs at most four blocks per CU (one-word class) or two
    // (wide classes): its few heavy-tailed roots leave most waves polling and
    // pulling donations (measured: soc-epinions with the two-word class
    // primary, 198 ms at four blocks and 166 ms at two; a 4,357-vertex graph
    // of the many-graph workload, 36 ms at four one-word blocks and 195 ms at
    // seven).  Batches and served epochs with many roots keep the full grid.
    // A class with nothing but paused frames left resumes on the grid it
    // paused on: a pause changes nothing about the work that remains
    // (measured on the Chung-Lu 320k window paused every 200M frames: with
    // the resumed classes re-ranked, the three-word class ran at one block
    // per CU and set the pace of every epoch).
    const uint64_t pending = uint64_t(C.htodo.size() - C.todo_done) + C.nseed + C.spill_n;
    const bool resume_only = C.nseed + C.spill_n > 0 && C.todo_done == C.htodo.size() && C.blocks_per_cu > 0;
    // (a resumed class keeps its demotion: it runs on the grid it paused on)
    if (T.blocks_per_cu > 0) {
        nb = std::min(nb, T.blocks_per_cu);
        C.demoted = false;
    } else if (resume_only) {
        nb = std::min(nb, C.blocks_per_cu);
    } else if (secondary) {
        nb = 1;
        C.demoted = nb < C.nb_max;
    } else {
        C.demoted = false;
        if (pending < uint64_t(nb) * uint64_t(E.cus) * 4u)
            nb = std::min(nb, W >= 2 ? 2 : (pending >= kManyRoots ? 6 : 4));
    }
    const uint32_t nwg = uint32_t(nb) * uint32_t(E.cus);
    SearchArgs sa{};
    sa.nwaves = nwg * 4;
    sa.seed_chunk = kSeedChunk;
    if (const char *e = std::getenv("KP_SEED_CHUNK")) sa.seed_chunk = std::min<uint32_t>(kSeedChunk, std::max(1, std::atoi(e)));
    sa.qcap = T.queue_slots ? std::max<uint32_t>(16, T.queue_slots) : std::max<uint32_t>(256, sa.nwaves * 4 / NSUBQ);
    sa.ring_cap = ring;
    sa.back_cap = back;
    // donation check period in frames: the one-word frame loop runs a frame
    // in a fraction of the wide classes' time (75k window: 179 ms at 512
    // against 188 ms at 256), and a one-word grid of six or more blocks per
    // CU, with more waves to feed, checks less often still (75k window at
    // six blocks: 164-167 ms at 2,048 frames)
    sa.donate_period = T.donate_period ? T.donate_period : (W == 1 ? (nb >= 6 ? 2048u : 512u) : 256u);
    sa.donate_target = std::max<uint32_t>(16, sa.nwaves / 4);
    sa.retire_margin = std::max<uint32_t>(32, sa.nwaves / 16);
    // a class alone on the chip keeps its waves: retiring them frees slots
    // for no one, and waves that stay are the ones donations can reach
    sa.retire_ticks = alone ? 0 : E.ticks_per_ms * 10;
    const size_t nslots = size_t(sa.qcap) * NSUBQ;
    int rc;
    if ((rc = C.q.ensure(nslots * FW * 8)) || (rc = C.qseq.ensure(nslots * 8)) || (rc = C.sched.ensure(sizeof(Sched))) ||
        (rc = C.subq.ensure(sizeof(SubQ) * NSUBQ)))
        return rc;
    if (back && (rc = C.stack.ensure(size_t(sa.nwaves) * back * FW * 8))) return rc;
    uint32_t dump_cap = 0;
    if (o.pausable) {
        dump_cap = uint32_t(std::min<uint64_t>(0xffffffffu, uint64_t(sa.nwaves) * (need + 2) + nslots));
        if ((rc = C.dump.ensure(size_t(dump_cap) * FW * 8))) return rc;
    }
    // root table (grown by whole re-upload) and the start list
    if (C.hdesc.size() * sizeof(RootDesc) > C.desc.bytes) {
        if ((rc = C.desc.ensure(2 * C.hdesc.size() * sizeof(RootDesc)))) return rc;
        C.desc_uploaded = 0;
    }
    if (C.desc_uploaded < C.hdesc.size()) {
        HIPCHK(hipMemcpyAsync(C.desc.as<RootDesc>() + C.desc_uploaded, C.hdesc.data() + C.desc_uploaded,
                              sizeof(RootDesc) * (C.hdesc.size() - C.desc_uploaded), hipMemcpyHostToDevice, C.stream));
        C.desc_uploaded = C.hdesc.size();
    }
    const size_t ntodo = C.htodo.size() - C.todo_done;
    if ((rc = C.todo.ensure(4 * std::max<size_t>(ntodo, 1)))) return rc;
    if (ntodo)
        HIPCHK(hipMemcpyAsync(C.todo.p, C.htodo.data() + C.todo_done, 4 * ntodo, hipMemcpyHostToDevice, C.stream));
    sa.desc = C.desc.as<RootDesc>();
    sa.todo = C.todo.as<uint32_t>();
    sa.ntodo = uint32_t(ntodo);
    // resumed frames: the device seeds, then every spilled frame read in place
    // from pinned host memory
    const size_t nhost = std::min<size_t>(C.spill_n, 0x7fffffffu - C.nseed);
    sa.seed = C.nseed ? C.seed[C.cur_seed].as<uint64_t>() : nullptr;
    sa.hseed = nhost ? C.spill_dev : nullptr;
    sa.nseed_dev = C.nseed;
    sa.nseed = C.nseed + uint32_t(nhost);
    sa.dump = o.pausable ? C.dump.as<uint64_t>() : nullptr;
    sa.dump_cap = dump_cap;
    sa.acc = E.acc.as<JobAcc>();
    sa.qframes = C.q.as<uint64_t>();
    sa.qseq = C.qseq.as<uint64_t>();
    sa.sched = C.sched.as<Sched>();
    sa.subq = C.subq.as<SubQ>();
    sa.gstack = back ? C.stack.as<uint64_t>() : nullptr;
    sa.list = o.want_list ? E.list.as<uint32_t>() : nullptr;
    sa.list_cap = o.list_cap;
    sa.list_used = E.list_used.as<uint64_t>();
    sa.limit_ticks = o.limit_ticks;
    sa.host_pause = o.host_pause ? E.dpause : nullptr;
    sa.class_pause = (o.regrid && E.dpause) ? E.dpause + 1 + c : nullptr;
    if (sa.class_pause) __atomic_store_n(&E.hpause[1 + c], 0u, __ATOMIC_RELEASE);
    sa.pause_frames = o.pause_frames;
    sa.t0 = E.t0.as<unsigned long long>();
    if (W >= 2 && E.nar_cur >= 0 && W >= o.narrow_from) {
        Engine::NarBuf &B = E.nar[E.nar_cur];
        sa.nar.arena = B.arena.as<uint64_t>();
        sa.nar.arena_cap = B.arena.bytes / 8;
        sa.nar.ctr = E.nar_ctr.as<unsigned long long>();
        const size_t first = E.nar_base + size_t(E.nar_cur) * E.nar_cap;
        sa.nar.desc = E.cls[0].desc.as<RootDesc>() + first;
        sa.nar.frames = B.frames.as<uint64_t>();
        sa.nar.cap = E.nar_cap;
        sa.nar.tag = uint32_t(first);
        sa.nar.min_slack = o.narrow_min_slack;
        sa.nar.pause_when_full = o.narrow_pause && o.pausable;
        sa.nar.after_ticks = uint64_t(o.narrow_after_ms * double(E.ticks_per_ms));
    }
    HIPCHK(hipMemsetAsync(sa.sched, 0, sizeof(Sched), C.stream));
    HIPCHK(hipMemsetAsync(sa.subq, 0, sizeof(SubQ) * NSUBQ, C.stream));
    hipLaunchKernelGGL(iota_kernel, dim3(uint32_t((nslots + 255) / 256)), dim3(256), 0, C.stream, sa.qseq,
                       uint32_t(nslots), sa.qcap);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(kern, dim3(nwg), dim3(256), dyn, C.stream, sa);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(C.done_ev, C.stream));
    C.launched = true;
    C.nwaves = sa.nwaves;
    C.ring = ring;
    C.back = back;
    C.dump_cap = dump_cap;
    C.launch_todo = uint32_t(ntodo);
    C.launch_seeds = sa.nseed;
    C.blocks_per_cu = nb;
    return KP_OK;
}

int launch_class_w(Engine &E, int c, const SearchOpts &o, bool secondary = false, bool alone = false)
{
    switch (kClassW[c]) {
    case 1: return launch_class<1, 2>(E, c, o, secondary, alone);
    case 2: return launch_class<2, 2>(E, c, o, secondary, alone);
    case 3: return launch_class<3, 2>(E, c, o, secondary, alone);
    case 4: return launch_class<4, 1>(E, c, o, secondary, alone);
    case 6: return launch_class<6, 0>(E, c, o, secondary, alone);
    case 8: return launch_class<8, 0>(E, c, o, secondary, alone);
    case 16: return launch_class<16, 0>(E, c, o, secondary, alone);
    }
    return KP_EINVAL;
}

// the epoch's shared start tick back to 0 (the first wave to start sets it)
hipError_t zero_t0(Engine &E, hipStream_t st)
{
    if (E.t0.ensure(256) != KP_OK) return hipErrorOutOfMemory;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, st, E.t0.as<uint64_t>(), 1u, 0ull);
    return hipGetLastError();
}

int grow_spill(ClassState &C, size_t need, size_t fbytes)
{
    if (need <= C.spill_cap) return KP_OK;
    const size_t cap = std::max<size_t>(need, 2 * C.spill_cap);
    void *p = nullptr, *d = nullptr;
    // mapped and coherent: search kernels read spilled frames in place
    if (hipHostMalloc(&p, cap * fbytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
        set_error("pinned host allocation of %zu bytes failed", cap * fbytes);
        return KP_ENOMEM;
    }
    if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess) {
        (void)hipHostFree(p);
        set_error("pinned host frames have no device address");
        return KP_EHIP;
    }
    if (C.spill_n) std::memcpy(p, C.spill, C.spill_n * fbytes);
    if (C.spill) (void)hipHostFree(C.spill);
    C.spill = static_cast<uint64_t *>(p);
    C.spill_dev = static_cast<const uint64_t *>(d);
    C.spill_cap = cap;
    return KP_OK;
}

// Pending frames after a launch: `D` dumped frames plus the resumed frames it
// did not take.  The launch offered the device seeds [0, nseed) and then the
// host frames [0, launch_seeds - nseed); `ns` of them, in that order, were
// taken.  Up to `budget` pending frames become the next launch's device seeds;
// the rest stay in (or go to) pinned host memory, which the next launch reads
// in place.  All copies run on the side stream.
int carry_frames(Engine &E, ClassState &C, size_t fbytes, uint32_t D, uint32_t ns, uint64_t budget, EpochStats &es)
{
    const size_t R = size_t(std::max<uint64_t>(1, std::min<uint64_t>(budget, 0x7fffffffu)));
    const uint32_t cd = std::min(ns, C.nseed);                       // device seeds taken
    const size_t ch = std::min<size_t>(ns - cd, C.spill_n);          // host frames taken
    if (ch) {
        std::memmove(C.spill, reinterpret_cast<const char *>(C.spill) + ch * fbytes, (C.spill_n - ch) * fbytes);
        C.spill_n -= ch;
    }
    const uint32_t U = C.nseed - cd;
    const size_t want = std::min<size_t>(R, size_t(D) + U + C.spill_n);
    DevBuf &nb = C.seed[1 - C.cur_seed];
    int rc;
    if (want && (rc = nb.ensure(std::max<size_t>(want, 1) * fbytes))) return rc;
    char *dst = static_cast<char *>(nb.p);
    const char *dump = static_cast<const char *>(C.dump.p);
    const char *old = static_cast<const char *>(C.seed[C.cur_seed].p);
    size_t fill = 0;
    const size_t kd = std::min<size_t>(D, R);
    if (kd) HIPCHK(hipMemcpyAsync(dst, dump, kd * fbytes, hipMemcpyDeviceToDevice, E.side));
    fill = kd;
    const size_t ku = std::min<size_t>(U, R - fill);
    if (ku) HIPCHK(hipMemcpyAsync(dst + fill * fbytes, old + size_t(cd) * fbytes, ku * fbytes, hipMemcpyDeviceToDevice, E.side));
    fill += ku;
    const size_t sd = D - kd, su = U - ku;
    if (sd + su) {
        if ((rc = grow_spill(C, C.spill_n + sd + su, fbytes))) return rc;
        char *hs = reinterpret_cast<char *>(C.spill) + C.spill_n * fbytes;
        if (sd) HIPCHK(hipMemcpyAsync(hs, dump + kd * fbytes, sd * fbytes, hipMemcpyDeviceToHost, E.side));
        if (su)
            HIPCHK(hipMemcpyAsync(hs + sd * fbytes, old + (size_t(cd) + ku) * fbytes, su * fbytes, hipMemcpyDeviceToHost, E.side));
        C.spill_n += sd + su;
        es.spilled += sd + su;
        E.total_spilled += sd + su;
    } else if (fill < R && C.spill_n) {
        const size_t kr = std::min(R - fill, C.spill_n);
        const char *hs = reinterpret_cast<const char *>(C.spill) + (C.spill_n - kr) * fbytes;
        HIPCHK(hipMemcpyAsync(dst + fill * fbytes, hs, kr * fbytes, hipMemcpyHostToDevice, E.side));
        C.spill_n -= kr;
        fill += kr;
        es.refilled += kr;
    }
    HIPCHK(hipStreamSynchronize(E.side));
    C.cur_seed = 1 - C.cur_seed;
    C.nseed = uint32_t(fill);
    return KP_OK;
}

} // namespace

int fill_words(hipStream_t st, void *p, uint64_t words, uint64_t v)
{
    if (!words) return KP_OK;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(uint32_t((words + 255) / 256)), dim3(256), 0, st, static_cast<uint64_t *>(p),
                       uint32_t(words), v);
    HIPCHK(hipGetLastError());
    return KP_OK;
}

// Stack of a per-root wave at width W: the DFS depth is at most n + 2 frames
// (every child shrinks C2 or C), the newest in the LDS ring.
void root_stack(int W, uint32_t &ring, uint32_t &back)
{
    const uint32_t FW = uint32_t(FrameWords(W));
    const uint32_t need = 64u * uint32_t(W) + 3;
    ring = std::min(need, std::max<uint32_t>(16, 6144 / (FW * 8)));
    back = need - ring + 1;
}

template <int W, int LW>
int launch_roots(hipStream_t st, const RootKernArgs &ra, uint32_t nwaves)
{
    constexpr int FW = FrameLayout<W>::FW;
    const size_t dyn = size_t(ra.s.ring_cap) * FW * 8 * WPB;
    hipLaunchKernelGGL((root_kernel<W, LW>), dim3((nwaves + WPB - 1) / WPB), dim3(256), dyn, st, ra);
    HIPCHK(hipGetLastError());
    return KP_OK;
}

// Per-root pass at width W (1, 2, 3, 4 words)
int launch_roots_w(int W, hipStream_t st, const RootKernArgs &ra, uint32_t nwaves)
{
    switch (W) {
    case 1: return launch_roots<1, 2>(st, ra, nwaves);
    case 2: return launch_roots<2, 2>(st, ra, nwaves);
    case 3: return launch_roots<3, 2>(st, ra, nwaves);
    case 4: return launch_roots<4, 1>(st, ra, nwaves);
    }
    return KP_EINVAL;
}

// root positions [0, pn - q + 2) (src/device_funcs.cu:44-45), optionally windowed
void root_window(uint32_t pn, const kp_params &P, int64_t &b, int64_t &e)
{
    e = std::max<int64_t>(0, int64_t(pn) - P.q + 2);
    b = 0;
    if (P.root_end > 0 || P.root_begin > 0) {
        b = std::min<int64_t>(P.root_begin, e);
        if (P.root_end > 0) e = std::min<int64_t>(P.root_end, e);
        e = std::max(e, b);
    }
}

int Engine::run_roots(const std::vector<JobIn> &jv, uint64_t budget, uint64_t limit_ticks,
                      std::vector<std::vector<uint32_t>> &left)
{
    left.assign(jv.size(), {});
    std::vector<uint32_t> sj, sp, si;   // slot: job id, position, index in jv
    uint32_t lcap = 1, d1cap = 1, stride = 1;
    for (size_t i = 0; i < jv.size(); i++) {
        const JobIn &j = jv[i];
        for (int64_t pos = j.pos_begin; pos < j.pos_end; pos++) {
            sj.push_back(j.id);
            sp.push_back(uint32_t(pos));
            si.push_back(uint32_t(i));
        }
        const uint32_t pn1 = std::max<uint32_t>(j.g.pn, 1);
        lcap = std::max(lcap, std::min<uint32_t>(pn1, std::max<uint32_t>(2048, 2 * j.g.maxdeg + 1024)));
        d1cap = std::max(d1cap, j.g.d1max);
        stride = std::max(stride, pn1);
    }
    if (sj.empty()) return KP_OK;
    const uint32_t w1cap = (d1cap + 63) / 64;
    const uint32_t nwaves = uint32_t(std::min<size_t>(sj.size(), size_t(cus) * 16));
    // sized for the widest pass (four words)
    uint32_t ring4 = 0, back4 = 0;
    root_stack(4, ring4, back4);
    const uint64_t arena_words = 3ull * 256 * 4 + uint64_t(lcap) * 4 + 128;
    int rc;
    const size_t cnt_bytes = size_t(nwaves) * stride * 4;
    if ((rc = slot_job.ensure(4 * sj.size())) || (rc = slot_pos.ensure(4 * sj.size())) ||
        (rc = rp_blk.ensure(size_t(nwaves) * NCAP * 4)) || (rc = rp_lft.ensure(size_t(nwaves) * lcap * 4)) ||
        (rc = rp_info.ensure(size_t(nwaves) * sizeof(RootInfo))) || (rc = rp_wr.ensure(size_t(nwaves) * 8)) ||
        (rc = ws_b1c.ensure(size_t(nwaves) * d1cap * 4)) || (rc = ws_a1.ensure(size_t(nwaves) * d1cap * w1cap * 8)) ||
        (rc = rp_arena.ensure(size_t(nwaves) * arena_words * 8)) || (rc = rp_desc.ensure(size_t(nwaves) * sizeof(RootDesc))) ||
        (rc = rp_stack.ensure(size_t(nwaves) * back4 * FrameWords(4) * 8)) || (rc = rp_def.ensure(4 * (sj.size() + 64))) ||
        (rc = rp_sched.ensure(sizeof(Sched))))
        return rc;
    {
        void *before = ws_cnt.p;
        if ((rc = ws_cnt.ensure(cnt_bytes))) return rc;
        if (ws_cnt.p != before || cnt_bytes > ws_cnt_zeroed) {
            HIPCHK(hipMemsetAsync(ws_cnt.p, 0, ws_cnt.bytes, stream));
            ws_cnt_zeroed = ws_cnt.bytes;
        }
    }
    HIPCHK(hipMemsetAsync(rp_sched.p, 0, sizeof(Sched), stream));
    RootKernArgs ra{};
    ra.d.jobs = jobs.as<JobDev>();
    ra.d.slot_job = slot_job.as<uint32_t>();
    ra.d.slot_pos = slot_pos.as<uint32_t>();
    ra.d.blk = rp_blk.as<uint32_t>();
    ra.d.lft = rp_lft.as<uint32_t>();
    ra.d.lcap = lcap;
    ra.d.info = rp_info.as<RootInfo>();
    ra.d.ws_cnt = ws_cnt.as<uint32_t>();
    ra.d.ws_b1c = ws_b1c.as<uint32_t>();
    ra.d.ws_a1 = ws_a1.as<uint64_t>();
    ra.d.cnt_stride = stride;
    ra.d.d1cap = d1cap;
    ra.d.w1cap = w1cap;
    ra.b.jobs = jobs.as<JobDev>();
    ra.b.blk = rp_blk.as<uint32_t>();
    ra.b.lft = rp_lft.as<uint32_t>();
    ra.b.lcap = lcap;
    ra.b.cnt_stride = stride;
    ra.b.wr_extra = rp_wr.as<uint64_t>();
    ra.arena = rp_arena.as<uint64_t>();
    ra.arena_words = arena_words;
    ra.desc = rp_desc.as<RootDesc>();
    ra.deferred = rp_def.as<uint32_t>() + 1;
    ra.ndeferred = rp_def.as<uint32_t>();
    ra.budget = budget;
    ra.s.desc = ra.desc;
    ra.s.acc = acc.as<JobAcc>();
    ra.s.sched = rp_sched.as<Sched>();
    ra.s.gstack = rp_stack.as<uint64_t>();
    ra.s.limit_ticks = limit_ticks;
    // With a budget only the one-word pass runs and every root it leaves goes
    // to the pool; without one, the roots wider than W words go on to the
    // next width, and those wider than four words to the pool.
    for (int W : {1, 2, 3, 4}) {
        if (sj.empty()) break;
        HIPCHK(hipMemcpyAsync(slot_job.p, sj.data(), 4 * sj.size(), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(slot_pos.p, sp.data(), 4 * sj.size(), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemsetAsync(rp_def.p, 0, 4, stream));
        const uint32_t nw = uint32_t(std::min<size_t>(nwaves, sj.size()));
        ra.d.nslots = uint32_t(sj.size());
        ra.d.nwaves = nw;
        ra.s.nwaves = nw;
        root_stack(W, ra.s.ring_cap, ra.s.back_cap);
        if ((rc = launch_roots_w(W, stream, ra, nw))) return rc;
        uint32_t nd = 0;
        Sched hs{};
        HIPCHK(hipMemcpyAsync(&nd, rp_def.p, 4, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(&hs, rp_sched.p, sizeof(Sched), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (hs.error) {
            set_error("per-root kernel (W=%d) reported error %d", W, hs.error);
            return hs.error;
        }
        std::vector<uint32_t> def(nd);
        if (nd) HIPCHK(hipMemcpy(def.data(), rp_def.as<uint32_t>() + 1, 4 * nd, hipMemcpyDeviceToHost));
        std::vector<uint32_t> nj, np, ni;
        for (uint32_t s : def) {
            if (s >= sj.size()) continue;
            nj.push_back(sj[s]);
            np.push_back(sp[s]);
            ni.push_back(si[s]);
        }
        sj.swap(nj);
        sp.swap(np);
        si.swap(ni);
        if (budget) break;
    }
    for (size_t t = 0; t < sj.size(); t++) left[si[t]].push_back(sp[t]);
    for (auto &v : left) std::sort(v.begin(), v.end());
    return KP_OK;
}

int Engine::launch_epoch(const SearchOpts &o)
{
    if (o.regrid) {
        int rc = enable_host_pause();
        if (rc) return rc;
    }
    // Narrowing: the epoch fills a free derived-root buffer (none free: no
    // narrowing in this epoch); its counters restart here.
    nar_cur = -1;
    if (o.narrow_from > 0) {
        bool wide = false;
        for (int c = 1; c < kNumClass; c++) wide |= kClassW[c] >= o.narrow_from && cls[c].has_work();
        int pick = -1;
        for (int b = 0; b < int(NDERIVED_BUF) && pick < 0 && wide; b++)
            if (!nar[b].in_use) pick = b;
        if (pick >= 0) {
            NarBuf &B = nar[pick];
            if (!nar_cap) {
                nar_cap = kNarrowRoots;
                if (const char *e = std::getenv("KP_NARROW_ROOTS"))
                    nar_cap = std::min<uint32_t>(1u << 28, uint32_t(std::max(1, std::atoi(e))));
            }
            size_t arena = kNarrowArenaBytes;
            if (const char *e = std::getenv("KP_NARROW_ARENA_MB")) arena = size_t(std::max(1, std::atoi(e))) << 20;
            int rc;
            if ((rc = nar_ctr.ensure(256)) || (rc = B.frames.ensure(size_t(nar_cap) * FrameWords(1) * 8)) ||
                (rc = B.arena.ensure(arena)))
                return rc;
            // The derived roots' entries go past the one-word class's regular
            // roots, in its own table; the table is reserved once per run for
            // every buffer (no regular root is added after staging).
            ClassState &C0 = cls[0];
            if (!nar_base_set) {
                nar_base = C0.hdesc.size();
                const size_t need = (nar_base + size_t(NDERIVED_BUF) * nar_cap) * sizeof(RootDesc);
                if (C0.desc.bytes < need) {
                    if ((rc = C0.desc.ensure(need))) return rc;
                    C0.desc_uploaded = 0;
                }
                nar_base_set = true;
            } else if (C0.hdesc.size() != nar_base) {
                set_error("one-word root table grew after derived roots were placed");
                return KP_EINVAL;
            }
            if (C0.desc_uploaded < C0.hdesc.size()) {
                HIPCHK(hipMemcpyAsync(C0.desc.as<RootDesc>() + C0.desc_uploaded, C0.hdesc.data() + C0.desc_uploaded,
                                      sizeof(RootDesc) * (C0.hdesc.size() - C0.desc_uploaded), hipMemcpyHostToDevice, stream));
                C0.desc_uploaded = C0.hdesc.size();
            }
            if ((rc = fill_words(stream, nar_ctr.p, 2, 0))) return rc;
            nar_cur = pick;
        }
    }
    // the class streams start after everything queued on the main stream
    HIPCHK(zero_t0(*this, stream));
    hipEvent_t ev;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(ev, stream));
    int rc = KP_OK;
    // Classes with the most pending roots and resumed frames launch first:
    // each launch fills the device, and later ones start on the CUs that the
    // earlier ones free.  (The size of a block says little about its work: on
    // the Chung-Lu 75k graph, roots of 35-60 dense community vertices hold far
    // more frames than hub roots of 200-380 sparse ones.)
    std::vector<std::pair<double, int>> order;
    for (int c = 0; c < kNumClass; c++) {
        cls[c].launched = false;
        if (!cls[c].has_work()) continue;
        const ClassState &C = cls[c];
        order.emplace_back(-double(C.htodo.size() - C.todo_done + C.nseed + C.spill_n), c);
    }
    std::stable_sort(order.begin(), order.end());
    // A resumed epoch offers every pending frame (device seeds and spilled
    // frames read in place), so it is launched like a fresh one.
    const bool demote = !order.empty() && order[0].second > 0;
    for (size_t oi = 0; oi < order.size() && !rc; oi++) {
        const int c = order[oi].second;
        if (hipStreamWaitEvent(cls[c].stream, ev, 0) != hipSuccess) rc = KP_EHIP;
        else rc = launch_class_w(*this, c, o, oi > 0 && demote, order.size() == 1);
        if (!rc && std::getenv("KP_SERIAL_CLASSES") && hipStreamSynchronize(cls[c].stream) != hipSuccess) rc = KP_EHIP;
    }
    (void)hipEventDestroy(ev);
    return rc;
}

int Engine::enable_host_pause()
{
    if (hpause) return KP_OK;
    void *p = nullptr;
    if (hipHostMalloc(&p, 256, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
        set_error("pinned pause word allocation failed");
        return KP_ENOMEM;
    }
    hpause = static_cast<uint32_t *>(p);
    *hpause = 0;
    void *d = nullptr;
    HIPCHK(hipHostGetDevicePointer(&d, p, 0));
    dpause = static_cast<uint32_t *>(d);
    return KP_OK;
}

int Engine::watch_epoch(const SearchOpts &o)
{
    if (!o.regrid || !hpause) return KP_OK;
    const double t_launch = now_ms();
    for (;;) {
        bool running = false, finished = false, blocker = false;
        int cand[kNumClass], nc = 0;
        for (int c = 0; c < kNumClass; c++) {
            ClassState &C = cls[c];
            if (!C.launched) continue;
            const hipError_t q = hipEventQuery(C.done_ev);
            if (q == hipSuccess) {
                finished = true;
                continue;
            }
            if (q != hipErrorNotReady) return KP_OK;   // finish_epoch reports the failure
            running = true;
            if (C.regrid_sent) continue;
            if (C.demoted) cand[nc++] = c;
            else blocker = true;
        }
        if (!running) return KP_OK;
        // Every class still running was demoted to one block per CU beside a
        // primary class that has drained: pause them, and the next epoch
        // launches the one with the most pending work on the whole chip.  Not
        // in a young epoch: a pause and relaunch cost more than what a short
        // remainder gains (email-euall: 184 -> 207 ms when always on).
        if (finished && !blocker && nc && now_ms() - t_launch >= o.regrid_after_ms) {
            for (int i = 0; i < nc; i++) {
                __atomic_store_n(&hpause[1 + cand[i]], 1u, __ATOMIC_RELEASE);
                cls[cand[i]].regrid_sent = true;
            }
        }
        std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
}

bool Engine::epoch_done()
{
    for (const auto &c : cls)
        if (c.launched && hipEventQuery(c.done_ev) == hipErrorNotReady) return false;
    return true;
}

int Engine::count_pending(std::vector<uint64_t> &per_job)
{
    std::vector<uint64_t> meta;
    for (int c = 0; c < kNumClass; c++) {
        ClassState &C = cls[c];
        const size_t fb = size_t(FrameWords(kClassW[c])) * 8;
        meta.assign(C.nseed, 0);
        if (C.nseed)
            HIPCHK(hipMemcpy2DAsync(meta.data(), 8, C.seed[C.cur_seed].p, fb, 8, C.nseed, hipMemcpyDeviceToHost, side));
        HIPCHK(hipStreamSynchronize(side));
        for (size_t i = 0; i < C.spill_n; i++) meta.push_back(C.spill[i * (fb / 8)]);
        for (uint64_t m : meta) {
            const uint32_t root = uint32_t(m);
            if (root < C.hdesc.size() && C.hdesc[root].job < per_job.size()) per_job[C.hdesc[root].job]++;
        }
    }
    return KP_OK;
}

int Engine::time_class(int c, const SearchOpts &o, float *ms)
{
    ClassState &C = cls[c];
    nar_cur = -1;   // a timed class runs alone, without narrowing
    const size_t saved_done = C.todo_done;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } eg{a, b};
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(zero_t0(*this, C.stream));
    HIPCHK(hipEventRecord(a, C.stream));
    int rc = launch_class_w(*this, c, o);
    if (rc) return rc;
    HIPCHK(hipEventRecord(b, C.stream));
    HIPCHK(hipEventSynchronize(b));
    C.launched = false;
    C.todo_done = saved_done;
    Sched hs{};
    HIPCHK(hipMemcpy(&hs, C.sched.p, sizeof(Sched), hipMemcpyDeviceToHost));
    if (hs.error) {
        set_error("search kernel (class W=%d) reported error %d", kClassW[c], hs.error);
        return hs.error;
    }
    HIPCHK(hipEventElapsedTime(ms, a, b));
    return KP_OK;
}

int Engine::finish_epoch(const SearchOpts &o, EpochStats &es)
{
    int first_err = KP_OK;
    std::vector<std::pair<int, Sched>> got;
    for (int c = 0; c < kNumClass; c++) {
        ClassState &C = cls[c];
        if (!C.launched) continue;
        Sched hs{};
        if (hipMemcpyAsync(&hs, C.sched.p, sizeof(Sched), hipMemcpyDeviceToHost, C.stream) != hipSuccess ||
            hipStreamSynchronize(C.stream) != hipSuccess) {
            set_error("search kernel (class W=%d) failed: %s", kClassW[c], hipGetErrorString(hipGetLastError()));
            if (!first_err) first_err = KP_EHIP;
            continue;
        }
        got.emplace_back(c, hs);
    }
    for (auto &cg : got) {
        const int c = cg.first;
        const Sched &hs = cg.second;
        ClassState &C = cls[c];
        C.launched = false;
        if (C.regrid_sent) {
            C.regrid_sent = false;
            C.blocks_per_cu = 0;   // relaunched by the ordinary rules, not on the grid it paused on
        }
        if (std::getenv("KP_STATS"))
            std::fprintf(stderr,
                         "[kp] search W=%d todo=%u seeds=%u waves=%u (%d blocks/CU) ring=%u back=%u frames=%llu "
                         "donated=%llu pops=%llu spills=%llu busy=%.1f%% wave-s=%.1f pause=%u dumped=%u retired=%u\n",
                         kClassW[c], C.launch_todo, C.nseed, C.nwaves, C.blocks_per_cu, C.ring, C.back,
                         (unsigned long long)hs.frames, (unsigned long long)hs.donated, (unsigned long long)hs.pops,
                         (unsigned long long)hs.spills,
                         100.0 * double(hs.busy_ticks) / double(std::max<uint64_t>(1, hs.busy_ticks + hs.idle_ticks)),
                         double(hs.busy_ticks + hs.idle_ticks) / 1e8, hs.pause, hs.dump_used, hs.retired);
        if (hs.error) {
            if (!first_err) {
                set_error("search kernel (class W=%d) reported error %d", kClassW[c], hs.error);
                first_err = hs.error;
            }
            continue;
        }
        es.frames += hs.frames;
        es.donated += hs.donated;
        const uint32_t started = std::min(hs.next_root, C.launch_todo);
        const uint32_t ns = std::min(hs.next_seed, C.launch_seeds);
        C.todo_done += started;
        if (!hs.pause && (started < C.launch_todo || ns < C.launch_seeds || hs.dump_used)) {
            if (!first_err) {
                set_error("search kernel (class W=%d) ended with work left", kClassW[c]);
                first_err = KP_EINVAL;
            }
            continue;
        }
        if (hs.pause) {
            es.paused = true;
            total_pauses++;
        }
        const uint32_t D = hs.pause ? hs.dump_used : 0;
        int rc = carry_frames(*this, C, size_t(FrameWords(kClassW[c])) * 8, D, ns, o.seed_budget, es);
        if (rc && !first_err) first_err = rc;
    }
    if (!first_err) {
        // the one-word class ran this epoch through (or had nothing to run) and
        // holds no pending frame: no frame names an older derived root
        bool c0_clean = cls[0].nseed == 0 && cls[0].spill_n == 0;
        for (auto &cg : got)
            if (cg.first == 0 && cg.second.pause) c0_clean = false;
        int rc = take_derived(es, c0_clean);
        if (rc) first_err = rc;
    }
    nar_cur = -1;
    return first_err;
}

// Hands the derived roots of the epoch that ended (narrowed frames of wide
// classes) to the one-word class: their frames join its resumed frames, on the
// device; their roots stay in the buffer, which stays in use until no frame
// can name them.
int Engine::take_derived(EpochStats &es, bool c0_clean)
{
    if (c0_clean)
        for (int b = 0; b < int(NDERIVED_BUF); b++)
            if (b != nar_cur) nar[b].in_use = false;
    if (nar_cur < 0) return KP_OK;
    unsigned long long ctr[2] = {0, 0};
    HIPCHK(hipMemcpy(ctr, nar_ctr.p, sizeof(ctr), hipMemcpyDeviceToHost));
    const uint32_t nd = uint32_t(std::min<unsigned long long>(ctr[1], nar_cap));
    NarBuf &B = nar[nar_cur];
    if (std::getenv("KP_STATS"))
        std::fprintf(stderr, "[kp] narrowed %u frames into the one-word class (buffer %d, arena %.1f of %.1f MiB)\n", nd,
                     nar_cur, double(std::min<unsigned long long>(ctr[0], B.arena.bytes / 8)) * 8 / 1048576.0,
                     double(B.arena.bytes) / 1048576.0);
    if (!nd) return KP_OK;
    B.in_use = true;
    ClassState &C0 = cls[0];
    const size_t fb = size_t(FrameWords(1)) * 8;
    DevBuf &nb = C0.seed[1 - C0.cur_seed];
    int rc;
    if ((rc = nb.ensure(size_t(C0.nseed + nd) * fb))) return rc;
    if (C0.nseed)
        HIPCHK(hipMemcpyAsync(nb.p, C0.seed[C0.cur_seed].p, size_t(C0.nseed) * fb, hipMemcpyDeviceToDevice, side));
    HIPCHK(hipMemcpyAsync(static_cast<char *>(nb.p) + size_t(C0.nseed) * fb, B.frames.p, size_t(nd) * fb,
                          hipMemcpyDeviceToDevice, side));
    HIPCHK(hipStreamSynchronize(side));
    C0.cur_seed = 1 - C0.cur_seed;
    C0.nseed += nd;
    C0.maxn = std::max<uint32_t>(C0.maxn, 64);
    C0.blocks_per_cu = 0;   // a fresh grid, not the one it last ran on
    total_narrowed += nd;
    es.narrowed += nd;
    return KP_OK;
}

} // namespace kp

// ----------------------------------------------------------------------
// prepared graphs (kp_prepare / kp_run) and batches (kp_enumerate_batch)
// ----------------------------------------------------------------------
struct kp_prepared {
    kp_params p{};
    std::unique_ptr<kp::Prep> prep{new kp::Prep()};
    kp::DevBuf gbuf;
    kp::GraphView gv;
    kp::Engine E;
    double preprocess_ms = 0, upload_ms = 0;
};

namespace kp {

int dev_info(char *buf, size_t len)
{
    int dev = 0, count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        set_error("no HIP device visible");
        return KP_ENODEV;
    }
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    std::snprintf(buf, len, "%s (%s), %d CUs, %.1f GiB", prop.name, prop.gcnArchName, prop.multiProcessorCount,
                  double(prop.totalGlobalMem) / (1024.0 * 1024.0 * 1024.0));
    return KP_OK;
}

uint64_t dev_bytes(const kp_prepared *h) { return h ? h->gbuf.bytes + h->E.device_bytes() : 0; }

int dev_set_window(kp_prepared *h, int64_t root_begin, int64_t root_end)
{
    h->p.root_begin = root_begin;
    h->p.root_end = root_end;
    return KP_OK;
}

void dev_free(kp_prepared *h) { delete h; }

int dev_prepare(const kp_graph *g, const kp_params *p, kp_prepared **out)
{
    *out = nullptr;
    auto *h = new (std::nothrow) kp_prepared();
    if (!h) return KP_ENOMEM;
    h->p = *p;
    int rc = h->E.init();
    if (rc) { delete h; return rc; }
    const double t0 = now_ms();
    if ((rc = prep_graph(g, *p, h->prep.get()))) { delete h; return rc; }
    h->preprocess_ms = now_ms() - t0;
    const double t1 = now_ms();
    std::vector<GraphView> views;
    if ((rc = upload_graphs({h->prep.get()}, h->gbuf, views, h->E.stream))) { delete h; return rc; }
    h->gv = views[0];
    h->upload_ms = now_ms() - t1;
    *out = h;
    return KP_OK;
}

namespace {

SearchOpts search_opts(const Engine &E, const kp_params &P, int device, int &rc)
{
    SearchOpts o;
    rc = KP_OK;
    o.want_list = (P.flags & KP_FLAG_LIST) != 0;
    o.list_cap = o.want_list ? (64ull << 20) : 0;
    if (P.time_limit_s > 0) {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) != hipSuccess) rc = KP_EHIP;
        o.limit_ticks = uint64_t(P.time_limit_s * double(khz) * 1e3);
    }
    o.pause_frames = P.pause_every;
    o.pausable = P.pause_every > 0;
    const double t = (P.thres > 0 && P.thres <= 1) ? double(P.thres) : 1.0;
    const uint64_t cap = P.queue_capacity ? P.queue_capacity : kDefaultQueueCap;
    o.seed_budget = std::max<uint64_t>(1, uint64_t(t * double(cap)));
    o.narrow_from = kNarrowFrom;
    o.narrow_min_slack = kNarrowMinSlack;
    if (const char *e = std::getenv("KP_NARROW_FROM")) o.narrow_from = std::atoi(e);
    if (const char *e = std::getenv("KP_NARROW_SLACK")) o.narrow_min_slack = uint32_t(std::atoi(e));
    o.regrid = !(std::getenv("KP_REGRID") && std::atoi(std::getenv("KP_REGRID")) == 0);
    o.narrow_pause = true;
    o.narrow_after_ms = kNarrowAfterMs;
    o.regrid_after_ms = kRegridAfterMs;
    if (const char *e = std::getenv("KP_NARROW_AFTER_MS")) o.narrow_after_ms = std::atof(e);
    if (const char *e = std::getenv("KP_REGRID_AFTER_MS")) o.regrid_after_ms = std::atof(e);
    o.pausable = o.pausable || o.regrid;
    if (P.flags & KP_FLAG_NO_NARROW) o.narrow_from = 0;
    (void)E;
    return o;
}

// Plex records are (len, job, ids...); each job's plexes become sorted id
// lists in lexicographic order.
int decode_lists(const std::vector<uint32_t> &listed, const std::vector<uint32_t> &job_slot, std::vector<kp_result *> &res)
{
    std::vector<std::vector<std::pair<uint64_t, uint32_t>>> recs(res.size());
    std::vector<uint32_t> buf(listed);
    for (uint64_t i = 0; i + 1 < buf.size();) {
        const uint32_t len = buf[i], job = buf[i + 1];
        if (i + 2 + len > buf.size()) break;
        std::sort(buf.begin() + i + 2, buf.begin() + i + 2 + len);
        if (job < job_slot.size() && job_slot[job] != ~0u) recs[job_slot[job]].emplace_back(i + 2, len);
        i += 2 + len;
    }
    for (size_t j = 0; j < res.size(); j++) {
        auto &rv = recs[j];
        std::sort(rv.begin(), rv.end(), [&](const auto &x, const auto &y) {
            return std::lexicographical_compare(buf.begin() + x.first, buf.begin() + x.first + x.second,
                                                buf.begin() + y.first, buf.begin() + y.first + y.second);
        });
        kp_result *r = res[j];
        r->nlisted = rv.size();
        r->list_off = static_cast<uint64_t *>(std::malloc(8 * (rv.size() + 1)));
        uint64_t tot = 0;
        for (auto &x : rv) tot += x.second;
        r->list_ids = static_cast<uint32_t *>(std::malloc(4 * std::max<uint64_t>(tot, 1)));
        if (!r->list_off || !r->list_ids) return KP_ENOMEM;
        uint64_t o = 0;
        for (size_t i = 0; i < rv.size(); i++) {
            r->list_off[i] = o;
            std::memcpy(r->list_ids + o, buf.data() + rv[i].first, 4 * rv[i].second);
            o += rv[i].second;
        }
        r->list_off[rv.size()] = o;
    }
    return KP_OK;
}

// Runs the jobs already in E's job table (ids 0..n-1) to completion: the
// per-root pass when mono_budget > 0 (small graphs), then staging of the roots
// left and search epochs until no class holds work.  search_ms covers the
// per-root pass and the epochs, decompose_ms the staging.
int run_pool(Engine &E, const std::vector<JobIn> &jobs, const kp_params &P0, std::vector<kp_result *> &res,
             float &decompose_ms, float &search_ms, uint64_t mono_budget = 0)
{
    int rc;
    E.reset_tables();
    if ((rc = E.zero_acc(0, uint32_t(jobs.size())))) return rc;
    hipEvent_t e0, em, e1, e2;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&em));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventCreate(&e2));
    struct EvGuard {
        hipEvent_t a, b, c, d;
        ~EvGuard()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            (void)hipEventDestroy(c);
            (void)hipEventDestroy(d);
        }
    } eg{e0, em, e1, e2};
    int orc = KP_OK;
    SearchOpts o = search_opts(E, P0, E.device, orc);
    if (orc) return orc;
    HIPCHK(hipEventRecord(e0, E.stream));
    // the per-root pass takes the jobs on small graphs; the roots it leaves,
    // and every root of the other jobs, go to the pool
    std::vector<std::vector<uint32_t>> left;
    std::vector<JobIn> jobs2 = jobs;
    if (mono_budget && !o.want_list && P0.pause_every == 0) {
        std::vector<JobIn> small;
        std::vector<size_t> which;
        for (size_t i = 0; i < jobs.size(); i++)
            if (jobs[i].g.pn <= kMonoMaxPn) {
                small.push_back(jobs[i]);
                which.push_back(i);
            }
        std::vector<std::vector<uint32_t>> lsmall;
        if (!small.empty() && (rc = E.run_roots(small, mono_budget, o.limit_ticks, lsmall))) return rc;
        left.assign(jobs.size(), {});
        for (size_t t = 0; t < which.size(); t++) {
            left[which[t]].swap(lsmall[t]);
            jobs2[which[t]].pos_list = &left[which[t]];
        }
    }
    HIPCHK(hipEventRecord(em, E.stream));
    std::vector<JobStats> st;
    // plain and batched runs: one arena segment per staging chunk, reused across runs
    auto alloc = [&E](size_t, uint64_t words) -> uint64_t * { return E.arena_alloc(words); };
    if ((rc = E.stage(jobs2, st, alloc))) return rc;
    HIPCHK(hipEventRecord(e1, E.stream));
    if (o.want_list) {
        if ((rc = E.list.ensure(o.list_cap * 4))) return rc;
    }
    if ((rc = E.list_used.ensure(8))) return rc;
    HIPCHK(hipMemsetAsync(E.list_used.p, 0, 8, E.stream));
    EpochStats es;
    const uint64_t spilled0 = E.total_spilled, pauses0 = E.total_pauses, narrowed0 = E.total_narrowed;
    while (E.has_work()) {
        if ((rc = E.launch_epoch(o))) return rc;
        if ((rc = E.watch_epoch(o))) return rc;
        if ((rc = E.finish_epoch(o, es))) return rc;
    }
    HIPCHK(hipEventRecord(e2, E.stream));
    HIPCHK(hipEventSynchronize(e2));
    float mono_ms = 0;
    HIPCHK(hipEventElapsedTime(&mono_ms, e0, em));
    HIPCHK(hipEventElapsedTime(&decompose_ms, em, e1));
    HIPCHK(hipEventElapsedTime(&search_ms, e1, e2));
    search_ms += mono_ms;
    std::vector<JobAcc> acc;
    if ((rc = E.read_acc(0, uint32_t(jobs.size()), acc))) return rc;
    for (size_t i = 0; i < jobs.size(); i++) {
        kp_result *r = res[i];
        const JobAcc &A = acc[i];
        r->count = A.count;
        r->hash_a = A.ha;
        r->hash_b = A.hb;
        r->tasks = A.tasks;
        r->search_alg_bytes = A.alg;
        r->alg_bytes = st[i].alg_bytes + A.dalg + A.alg;
        r->adj_bytes = st[i].adj_bytes + A.adj;
        r->valid_roots = st[i].valid_roots + A.roots;
        r->max_n = std::max(st[i].max_n, A.max_n);
        r->max_l = std::max(st[i].max_l, A.max_l);
        r->pn = jobs[i].g.pn;
        r->pm = jobs[i].g.pm;
        r->spilled = E.total_spilled - spilled0;
        r->pauses = E.total_pauses - pauses0;
        r->narrowed = E.total_narrowed - narrowed0;
    }
    if (o.want_list) {
        uint64_t used = 0;
        HIPCHK(hipMemcpy(&used, E.list_used.p, 8, hipMemcpyDeviceToHost));
        if (used > o.list_cap) {
            set_error("plex list exceeds %llu words", (unsigned long long)o.list_cap);
            return KP_EOVERFLOW;
        }
        std::vector<uint32_t> listed(used);
        if (used) HIPCHK(hipMemcpy(listed.data(), E.list.p, 4 * used, hipMemcpyDeviceToHost));
        std::vector<uint32_t> slot(jobs.size());
        for (size_t i = 0; i < jobs.size(); i++) slot[i] = uint32_t(i);
        if ((rc = decode_lists(listed, slot, res))) return rc;
    }
    return KP_OK;
}

} // namespace

int dev_run(kp_prepared *h, kp_result *res)
{
    std::memset(res, 0, sizeof(*res));
    HIPCHK(hipSetDevice(h->E.device));
    int rc;
    if ((rc = h->E.reserve_jobs(1))) return rc;
    h->E.set_job(0, h->gv, h->p.k, h->p.q);
    if ((rc = h->E.sync_jobs())) return rc;
    JobIn j{0, h->gv, h->p.k, h->p.q, 0, 0};
    root_window(h->gv.pn, h->p, j.pos_begin, j.pos_end);
    std::vector<kp_result *> rv{res};
    float dms = 0, sms = 0;
    // Small graphs take the per-root pass first: one launch walks the roots
    // whose subtrees stay under the frame budget; the rest go to the pool.
    if ((rc = run_pool(h->E, {j}, h->p, rv, dms, sms, kMonoBudget))) return rc;
    res->preprocess_ms = h->preprocess_ms;
    res->upload_ms = h->upload_ms;
    res->pn_first = h->prep->pn_first;
    res->decompose_ms = dms;
    res->search_ms = sms;
    res->wall_ms = res->preprocess_ms + dms + sms;
    return KP_OK;
}

namespace {

// The engine behind kp_enumerate, kept per device between calls: its
// streams, workspaces and arena segments are reused, so a call on a small
// graph costs preprocessing, one upload and the kernels.  A call that finds
// it busy (another host thread) runs on an engine of its own.
struct EnumCache {
    std::mutex mu;
    kp_prepared *h = nullptr;
};
EnumCache g_enum_cache[64];
constexpr uint64_t kEnumCacheKeepBytes = 1ull << 30;   // larger engines are freed after the call

} // namespace

int dev_enumerate(const kp_graph *g, const kp_params *p, kp_result *r)
{
    std::memset(r, 0, sizeof(*r));
    int count = 0, dev = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        set_error("no HIP device visible: the HIP backend is required");
        return KP_ENODEV;
    }
    HIPCHK(hipGetDevice(&dev));
    std::unique_lock<std::mutex> lk;
    if (dev < 64) lk = std::unique_lock<std::mutex>(g_enum_cache[dev].mu, std::try_to_lock);
    if (!lk.owns_lock()) {
        kp_prepared *h = nullptr;
        int rc = dev_prepare(g, p, &h);
        if (rc) return rc;
        rc = dev_run(h, r);
        dev_free(h);
        return rc;
    }
    EnumCache &C = g_enum_cache[dev];
    int rc;
    if (!C.h) {
        C.h = new (std::nothrow) kp_prepared();
        if (!C.h) return KP_ENOMEM;
        if ((rc = C.h->E.init())) {
            delete C.h;
            C.h = nullptr;
            return rc;
        }
    }
    kp_prepared *h = C.h;
    h->p = *p;
    h->prep.reset(new Prep());
    const double t0 = now_ms();
    rc = prep_graph(g, *p, h->prep.get());
    h->preprocess_ms = now_ms() - t0;
    if (!rc) {
        const double t1 = now_ms();
        std::vector<GraphView> views;
        rc = upload_graphs({h->prep.get()}, h->gbuf, views, h->E.stream);
        if (!rc) {
            h->gv = views[0];
            h->upload_ms = now_ms() - t1;
            rc = dev_run(h, r);
        }
    }
    h->prep.reset(new Prep());
    if (rc || dev_bytes(h) > kEnumCacheKeepBytes) {   // after an error the next call starts afresh
        delete h;
        C.h = nullptr;
    }
    return rc;
}

// One root's decomposition and subgraph as the GPU builds them (the staging
// path: decompose_kernel, then build_kernel of the root's size class), copied
// back: peeled ids of the block and the left set in canonical order and, per
// pair of block vertices, adjacency and the two class thresholds the search
// tests.  Returns 1 for a valid root, 0 for an invalid one.
int dev_root_subgraph(kp_prepared *h, int64_t pos, uint32_t *block, uint32_t *n, uint32_t *hop, uint32_t *left,
                      uint32_t *l, uint32_t left_cap, uint8_t *pairs, size_t pairs_cap)
{
    HIPCHK(hipSetDevice(h->E.device));
    Engine &E = h->E;
    *n = *hop = *l = 0;
    if (pos < 0 || pos >= int64_t(h->gv.pn)) return KP_EINVAL;
    int rc;
    if ((rc = E.reserve_jobs(1))) return rc;
    E.set_job(0, h->gv, h->p.k, h->p.q);
    if ((rc = E.sync_jobs())) return rc;
    E.reset_tables();
    JobIn j{0, h->gv, h->p.k, h->p.q, pos, pos + 1};
    std::vector<JobStats> st;
    auto alloc = [&E](size_t, uint64_t words) -> uint64_t * { return E.arena_alloc(words); };
    if ((rc = E.stage({j}, st, alloc))) return rc;
    RootInfo ri{};
    HIPCHK(hipMemcpy(&ri, E.info.p, sizeof ri, hipMemcpyDeviceToHost));
    int c = -1;
    for (int x = 0; x < kNumClass; x++)
        if (!E.cls[x].hdesc.empty()) c = x;
    if (ri.status != 1 || c < 0) {
        E.reset_tables();
        return 0;
    }
    const RootDesc d = E.cls[c].hdesc[0];
    const uint32_t W = uint32_t(kClassW[c]);
    if (d.l > left_cap || (pairs && size_t(d.n) * d.n > pairs_cap)) {
        E.reset_tables();
        set_error("root at position %lld: block %u / left set %u exceed the caller's buffers", (long long)pos, d.n, d.l);
        return KP_ERANGE;
    }
    std::vector<uint64_t> rows(size_t(3) * d.n * W);
    HIPCHK(hipMemcpy(block, E.blk.p, 4 * size_t(d.n), hipMemcpyDeviceToHost));
    if (d.l) HIPCHK(hipMemcpy(left, E.lft.p, 4 * size_t(d.l), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(rows.data(), reinterpret_cast<const void *>(d.addr), 8 * rows.size(), hipMemcpyDeviceToHost));
    if (pairs) {
        const uint64_t *A = rows.data(), *C2 = A + size_t(d.n) * W, *C4 = C2 + size_t(d.n) * W;
        for (uint32_t i = 0; i < d.n; i++)
            for (uint32_t t = 0; t < d.n; t++) {
                const size_t wd = size_t(i) * W + (t >> 6);
                const uint64_t b = 1ull << (t & 63);
                pairs[size_t(i) * d.n + t] = uint8_t(((A[wd] & b) ? 1 : 0) | ((C2[wd] & b) ? 2 : 0) | ((C4[wd] & b) ? 4 : 0));
            }
    }
    *n = d.n;
    *hop = d.h;
    *l = d.l;
    E.reset_tables();
    return 1;
}

// Coordinate descent over the launch settings of each size class, timing the
// class alone on this prepared workload; the best settings become the
// process-wide ones.
int dev_autotune(kp_prepared *h, kp_tune *out)
{
    std::memset(out, 0, sizeof(*out));
    HIPCHK(hipSetDevice(h->E.device));
    Engine &E = h->E;
    int rc;
    if ((rc = E.reserve_jobs(1))) return rc;
    E.set_job(0, h->gv, h->p.k, h->p.q);
    if ((rc = E.sync_jobs())) return rc;
    JobIn j{0, h->gv, h->p.k, h->p.q, 0, 0};
    root_window(h->gv.pn, h->p, j.pos_begin, j.pos_end);
    E.reset_tables();
    if ((rc = E.zero_acc(0, 1))) return rc;
    std::vector<JobStats> st;
    auto alloc = [&E](size_t, uint64_t words) -> uint64_t * { return E.arena_alloc(words); };
    if ((rc = E.stage({j}, st, alloc))) return rc;
    kp_params P = h->p;
    P.pause_every = 0;
    P.flags &= ~KP_FLAG_LIST;
    int orc = KP_OK;
    SearchOpts o = search_opts(E, P, E.device, orc);
    if (orc) return orc;
    if ((rc = E.list_used.ensure(8))) return rc;
    for (int c = 0; c < kNumClass; c++) {
        if (E.cls[c].htodo.empty()) continue;
        ClassTune best = g_tune[c];
        float t_best = 0, t = 0, t2 = 0;
        if ((rc = E.time_class(c, o, &t_best))) return rc;   // warm-up
        if ((rc = E.time_class(c, o, &t_best))) return rc;
        if ((rc = E.time_class(c, o, &t2))) return rc;
        t_best = std::min(t_best, t2);
        // a candidate's time is the better of two runs (single runs vary by
        // a few percent from one to the next)
        auto trial = [&](ClassTune cand) -> int {
            g_tune[c] = cand;
            int r = E.time_class(c, o, &t);
            if (!r) r = E.time_class(c, o, &t2);
            g_tune[c] = best;
            if (r) return r;
            t = std::min(t, t2);
            if (t < 0.97f * t_best) {
                best = cand;
                t_best = t;
                g_tune[c] = best;
            }
            return KP_OK;
        };
        // The grid (every size from one below the untuned grid up to the
        // occupancy limit) jointly with the donation period: fewer donation
        // checks and more waves go together, so one coordinate at a time
        // misses the pair (75k window: 4 blocks per CU at 512 frames 182 ms,
        // 6 blocks at 2,048-4,096 frames 164-167 ms, either change alone
        // 176 ms).  Then short periods on the grid found.
        const uint32_t dp_default = kClassW[c] == 1 ? 512u : 256u;
        const int nb_now = std::max(1, E.cls[c].blocks_per_cu);
        const int nb_lim = std::max(nb_now, E.cls[c].nb_max);
        for (int nb = nb_lim; nb >= std::max(1, nb_now - 1); nb--) {
            for (uint32_t dp : {256u, 512u, 1024u, 2048u, 4096u}) {
                if (nb == nb_now && dp == dp_default) continue;
                ClassTune x = best;
                x.blocks_per_cu = nb;
                x.donate_period = dp;
                if ((rc = trial(x))) return rc;
            }
        }
        for (uint32_t dp : {64u, 128u}) {
            ClassTune x = best;
            x.donate_period = dp;
            if ((rc = trial(x))) return rc;
        }
        for (int nb = nb_now - 2; nb >= 1; nb--) {
            ClassTune x = best;
            x.blocks_per_cu = nb;
            if ((rc = trial(x))) return rc;
        }
        for (uint32_t rf : {8u, 32u, 64u}) {
            if (kClassW[c] == 1) break;   // the one-word ring always holds the whole stack
            ClassTune x = best;
            x.ring_frames = rf;
            if ((rc = trial(x))) return rc;
        }
        for (uint32_t qs : {64u, 2048u}) {
            ClassTune x = best;
            x.queue_slots = qs;
            if ((rc = trial(x))) return rc;
        }
        g_tune[c] = best;
        out->blocks_per_cu[c] = best.blocks_per_cu;
        out->donate_period[c] = best.donate_period;
        out->ring_frames[c] = best.ring_frames;
        out->queue_slots[c] = best.queue_slots;
        out->class_ms[c] = t_best;
    }
    E.reset_tables();
    return KP_OK;
}

void dev_set_tune(const kp_tune *t)
{
    for (int c = 0; c < kNumClass; c++) {
        g_tune[c] = ClassTune{};
        if (!t) continue;
        g_tune[c].blocks_per_cu = t->blocks_per_cu[c];
        g_tune[c].donate_period = t->donate_period[c];
        g_tune[c].ring_frames = t->ring_frames[c];
        g_tune[c].queue_slots = t->queue_slots[c];
    }
}

// Many graphs in one pool: host preprocessing in parallel over the graphs,
// one packed upload, one staging pass over every root, one set of search
// launches.  Each result's decompose/search times are the batch's.
int dev_batch(const kp_graph *const *g, const kp_params *p, size_t n, kp_result *r)
{
    for (size_t i = 0; i < n; i++) std::memset(&r[i], 0, sizeof(kp_result));
    if (!n) return KP_OK;
    Engine E;
    int rc = E.init();
    if (rc) return rc;
    const double t0 = now_ms();
    std::vector<Prep> preps(n);
    std::vector<int> prc(n, KP_OK);
    std::vector<double> pms(n, 0);
    // host preprocessing in parallel over the graphs; GPU preprocessing
    // (KP_FLAG_GPU_PREP of p[0]) one graph after another
    const bool on_gpu = (p[0].flags & KP_FLAG_GPU_PREP) != 0;
#pragma omp parallel for schedule(dynamic, 1) if (!on_gpu)
    for (size_t i = 0; i < n; i++) {
        const double a = now_ms();
        prc[i] = on_gpu ? gpu_preprocess(g[i], p[i].k, p[i].q, p[i].truss, &preps[i])
                        : preprocess(g[i], p[i].k, p[i].q, p[i].truss, &preps[i]);
        pms[i] = now_ms() - a;
    }
    for (size_t i = 0; i < n; i++)
        if (prc[i]) return prc[i];
    const double t1 = now_ms();
    DevBuf gbuf;
    std::vector<GraphView> views;
    std::vector<const Prep *> pp(n);
    for (size_t i = 0; i < n; i++) pp[i] = &preps[i];
    if ((rc = upload_graphs(pp, gbuf, views, E.stream))) return rc;
    const double t2 = now_ms();
    if ((rc = E.reserve_jobs(n))) return rc;
    std::vector<JobIn> jobs(n);
    for (size_t i = 0; i < n; i++) {
        E.set_job(uint32_t(i), views[i], p[i].k, p[i].q);
        jobs[i] = JobIn{uint32_t(i), views[i], p[i].k, p[i].q, 0, 0};
        root_window(views[i].pn, p[i], jobs[i].pos_begin, jobs[i].pos_end);
    }
    if ((rc = E.sync_jobs())) return rc;
    std::vector<kp_result *> rv(n);
    for (size_t i = 0; i < n; i++) rv[i] = &r[i];
    float dms = 0, sms = 0;
    if ((rc = run_pool(E, jobs, p[0], rv, dms, sms, kMonoBudget))) return rc;
    for (size_t i = 0; i < n; i++) {
        r[i].preprocess_ms = pms[i];
        r[i].pn_first = preps[i].pn_first;
        r[i].upload_ms = t2 - t1;
        r[i].decompose_ms = dms;
        r[i].search_ms = sms;
        r[i].wall_ms = (t1 - t0) + dms + sms;
    }
    return KP_OK;
}

} // namespace kp
