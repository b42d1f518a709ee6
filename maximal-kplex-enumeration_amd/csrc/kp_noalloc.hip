// kp_noalloc.hip -- kp_enumerate_noalloc: the plain enumeration with no
// device allocation (no reference counterpart).
#include <hip/hip_runtime.h>

#include "kp_device.h"
#include "kp_engine.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

using namespace kpd;

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) {                                                        \
            kp::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
            return KP_EHIP;                                                            \
        }                                                                              \
    } while (0)

// ----------------------------------------------------------------------
// Every device buffer (the graph copy, the job and accumulator records, the
// per-wave workspaces, arenas and stacks) is carved out of one file-scope
// array, `scratch`, of KP_NOALLOC_SCRATCH_BYTES.  Each GPU holds its own copy
// from the library's first use on it until the process ends.  Host code never
// passes the array's name to a kernel: at the first call on a device, with
// that device current, hipGetSymbolAddress gives the array's address on that
// device, which is kept per device and used everywhere below.  Calls on one
// GPU share the array, so calls that overlap on one GPU race; every call
// zeroes what must start at zero.
// ----------------------------------------------------------------------
static __device__ __align__(256) uint8_t scratch[KP_NOALLOC_SCRATCH_BYTES];

namespace kp {

namespace {

std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[64] = {};

int scratch_on(int device, uint8_t **out)
{
    if (device < 0 || device >= 64) return KP_ENODEV;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
            return KP_EHIP;
        }
        g_scratch_ptr[device] = static_cast<uint8_t *>(p);
    }
    *out = g_scratch_ptr[device];
    return KP_OK;
}

// bump allocator over the scratch array (offsets only; nothing is allocated)
struct Carve {
    uint8_t *base;
    size_t used = 0, cap;
    template <class T> T *take(size_t count)
    {
        const size_t at = (used + 255) & ~size_t(255);
        used = at + sizeof(T) * std::max<size_t>(count, 1);
        return reinterpret_cast<T *>(base + at);
    }
};

} // namespace

int dev_enumerate_noalloc(const kp_graph *g, const kp_params *p, kp_result *r)
{
    std::memset(r, 0, sizeof(*r));
    int count = 0, dev = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        set_error("no HIP device visible: the HIP backend is required");
        return KP_ENODEV;
    }
    HIPCHK(hipGetDevice(&dev));
    uint8_t *base = nullptr;
    int rc = scratch_on(dev, &base);
    if (rc) return rc;
    int cus = 0;
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    uint64_t ticks_per_ms = 100000;
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
            ticks_per_ms = uint64_t(khz);
    }
    const double t0 = now_ms();
    Prep prep;
    if ((rc = preprocess(g, p->k, p->q, p->truss, &prep))) return rc;
    const double prep_ms = now_ms() - t0;
    const kp_graph &G = prep.g;
    const uint32_t pn = G.n;
    int64_t pb = 0, pe = 0;
    root_window(pn, *p, pb, pe);
    const size_t S = size_t(pe - pb);
    uint32_t maxdeg = prep.maxdeg, d1 = prep.d1max;
    if (!prep.sizes_known) {
        maxdeg = d1 = 0;
        for (uint32_t u = 0; u < pn; u++) {
            maxdeg = std::max(maxdeg, G.deg[u]);
            uint32_t later = 0;
            for (uint64_t e = G.off[u]; e < G.off[u + 1]; e++) later += prep.dpos[G.nbr[e]] > prep.dpos[u];
            d1 = std::max(d1, later);
        }
    }
    const uint32_t pn1 = std::max<uint32_t>(pn, 1);
    const uint32_t lcap = std::min<uint32_t>(pn1, std::max<uint32_t>(2048, 2 * maxdeg + 1024));
    const uint32_t d1cap = std::max<uint32_t>(d1, 1), w1cap = (d1cap + 63) / 64;
    const uint32_t stride = pn1;
    uint32_t ring4 = 0, back4 = 0;
    root_stack(4, ring4, back4);
    const uint64_t arena_words = 3ull * 256 * 4 + uint64_t(lcap) * 4 + 128;
    // layout for `nw` waves; the widest one that fits the array is used
    auto layout = [&](uint32_t nw, Carve &c, RootKernArgs &ra, uint8_t **gsec) {
        *gsec = c.take<uint8_t>(8 * (size_t(pn) + 1) + 255 + 4 * G.m + 255 + 4 * 4 * (size_t(pn) + 64));
        ra.d.jobs = c.take<JobDev>(1);
        ra.s.acc = c.take<JobAcc>(1);
        ra.s.sched = c.take<Sched>(1);
        ra.ndeferred = c.take<uint32_t>(S + 1);
        ra.deferred = ra.ndeferred + 1;
        ra.d.slot_job = c.take<uint32_t>(S);
        ra.d.slot_pos = c.take<uint32_t>(S);
        ra.d.ws_cnt = c.take<uint32_t>(size_t(nw) * stride);
        ra.d.blk = c.take<uint32_t>(size_t(nw) * KP_MAX_BLOCK);
        ra.d.lft = c.take<uint32_t>(size_t(nw) * lcap);
        ra.d.info = c.take<RootInfo>(nw);
        ra.b.wr_extra = c.take<uint64_t>(nw);
        ra.d.ws_b1c = c.take<uint32_t>(size_t(nw) * d1cap);
        ra.d.ws_a1 = c.take<uint64_t>(size_t(nw) * d1cap * w1cap);
        ra.arena = c.take<uint64_t>(size_t(nw) * arena_words);
        ra.desc = c.take<RootDesc>(nw);
        ra.s.gstack = c.take<uint64_t>(size_t(nw) * back4 * FrameWords(4));
    };
    uint32_t nw = uint32_t(std::min<size_t>(std::max<size_t>(S, 1), size_t(cus) * 4));
    RootKernArgs ra{};
    Carve c{base, 0, size_t(KP_NOALLOC_SCRATCH_BYTES)};
    uint8_t *gsec = nullptr;
    for (;;) {
        c.used = 0;
        layout(nw, c, ra, &gsec);
        if (c.used <= c.cap) break;
        if (nw == 1) {
            set_error("the preprocessed graph (%u vertices, %llu entries) needs %zu bytes of device memory, more than "
                      "the %u bytes of the no-allocation scratch array", pn, (unsigned long long)G.m, c.used,
                      unsigned(KP_NOALLOC_SCRATCH_BYTES));
            return KP_ENOMEM;
        }
        nw = std::max<uint32_t>(1, nw / 2);
    }
    // the preprocessed graph, packed as upload_graphs does, in one copy
    std::vector<uint8_t> host;
    GraphView gv;
    {
        size_t o = 0;
        auto put = [&](const void *src, size_t bytes) -> const uint8_t * {
            const size_t at = o;
            host.resize(at + ((bytes + 255) & ~size_t(255)));
            if (bytes) std::memcpy(host.data() + at, src, bytes);
            o = host.size();
            return gsec + at;
        };
        gv.off = reinterpret_cast<const uint64_t *>(put(G.off, 8 * (size_t(pn) + 1)));
        gv.nbr = reinterpret_cast<const uint32_t *>(put(G.nbr, 4 * G.m));
        gv.deg = reinterpret_cast<const uint32_t *>(put(G.deg, 4 * size_t(pn)));
        gv.dpos = reinterpret_cast<const int32_t *>(put(prep.dpos.data(), 4 * size_t(pn)));
        gv.dseq = reinterpret_cast<const int32_t *>(put(prep.dseq.data(), 4 * size_t(pn)));
        gv.orig = reinterpret_cast<const uint32_t *>(put(prep.orig.data(), 4 * size_t(pn)));
    }
    const JobDev jd{gv.off, gv.nbr, gv.deg, gv.orig, gv.dpos, gv.dseq, pn, p->k, p->q, 0};
    std::vector<uint32_t> sj(S, 0), sp(S);
    for (size_t t = 0; t < S; t++) sp[t] = uint32_t(pb + int64_t(t));
    const hipStream_t st = nullptr;   // the null stream: nothing to create
    const double t1 = now_ms();
    if (!host.empty()) HIPCHK(hipMemcpyAsync(gsec, host.data(), host.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(const_cast<JobDev *>(ra.d.jobs), &jd, sizeof jd, hipMemcpyHostToDevice, st));
    auto zero = [&](void *ptr, size_t bytes) { return fill_words(st, ptr, (bytes + 7) / 8, 0); };
    if ((rc = zero(ra.s.acc, sizeof(JobAcc))) || (rc = zero(ra.d.ws_cnt, size_t(nw) * stride * 4))) return rc;
    ra.d.lcap = ra.b.lcap = lcap;
    ra.d.cnt_stride = ra.b.cnt_stride = stride;
    ra.d.d1cap = d1cap;
    ra.d.w1cap = w1cap;
    ra.b.jobs = ra.d.jobs;
    ra.b.blk = ra.d.blk;
    ra.b.lft = ra.d.lft;
    ra.arena_words = arena_words;
    ra.budget = 0;
    ra.s.desc = ra.desc;
    if (p->time_limit_s > 0) ra.s.limit_ticks = uint64_t(p->time_limit_s * 1e3 * double(ticks_per_ms));
    for (int W : {1, 2, 3, 4}) {
        if (sj.empty()) break;
        HIPCHK(hipMemcpyAsync(const_cast<uint32_t *>(ra.d.slot_job), sj.data(), 4 * sj.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(const_cast<uint32_t *>(ra.d.slot_pos), sp.data(), 4 * sj.size(), hipMemcpyHostToDevice, st));
        if ((rc = zero(ra.ndeferred, 8)) || (rc = zero(ra.s.sched, sizeof(Sched)))) return rc;
        const uint32_t nwp = uint32_t(std::min<size_t>(nw, sj.size()));
        ra.d.nslots = uint32_t(sj.size());
        ra.d.nwaves = ra.s.nwaves = nwp;
        root_stack(W, ra.s.ring_cap, ra.s.back_cap);
        if ((rc = launch_roots_w(W, st, ra, nwp))) return rc;
        uint32_t nd = 0;
        int32_t kerr = 0;
        HIPCHK(hipMemcpyAsync(&nd, ra.ndeferred, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(&kerr, &ra.s.sched->error, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (kerr) {
            set_error("per-root kernel (W=%d) reported error %d", W, kerr);
            return kerr;
        }
        std::vector<uint32_t> def(nd);
        if (nd) {
            HIPCHK(hipMemcpyAsync(def.data(), ra.deferred, 4 * nd, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
        }
        std::vector<uint32_t> nj, np;
        for (uint32_t s : def)
            if (s < sj.size()) {
                nj.push_back(sj[s]);
                np.push_back(sp[s]);
            }
        sj.swap(nj);
        sp.swap(np);
    }
    if (!sj.empty()) {
        set_error("the root at degeneracy position %u has a block of more than 256 vertices, beyond the "
                  "no-allocation path (use kp_enumerate)", sp[0]);
        return KP_ERANGE;
    }
    JobAcc A{};
    HIPCHK(hipMemcpyAsync(&A, ra.s.acc, sizeof A, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const double t2 = now_ms();
    r->count = A.count;
    r->hash_a = A.ha;
    r->hash_b = A.hb;
    r->tasks = A.tasks;
    r->valid_roots = A.roots;
    r->search_alg_bytes = A.alg;
    r->alg_bytes = A.dalg + A.alg;
    r->adj_bytes = A.adj;
    r->max_n = A.max_n;
    r->max_l = A.max_l;
    r->pn = pn;
    r->pm = G.m;
    r->pn_first = prep.pn_first;
    r->preprocess_ms = prep_ms;
    r->search_ms = t2 - t1;   // upload, decomposition, build and search, end to end
    r->wall_ms = prep_ms + (t2 - t1);
    return KP_OK;
}

} // namespace kp
