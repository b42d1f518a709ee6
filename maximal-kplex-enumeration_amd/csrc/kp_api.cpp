// kp_api.cpp -- the extern "C" entry points of libkplex (declared in include/kplex.h).
#include "kp_device.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

using namespace kp;

namespace {

int check_params(const kp_graph *g, const kp_params *p)
{
    if (!g || !p) return KP_EINVAL;
    // The reference compares q-2k against unsigned counters (src/device_funcs.cu:148,
    // :163, :189), so any q < 2k rejects every root block; it also reads the first
    // k plex members in its maximality test.  Such parameters are refused here.
    if (p->k < 1 || p->q < 2 * p->k || p->q > int(KP_MAX_BLOCK)) {
        set_error("invalid parameters k=%d q=%d (need k >= 1 and 2k <= q <= %d)", p->k, p->q, KP_MAX_BLOCK);
        return KP_EINVAL;
    }
    if (p->root_begin < 0 || p->root_end < 0 || (p->root_end && p->root_end < p->root_begin)) {
        set_error("invalid root window [%lld, %lld)", (long long)p->root_begin, (long long)p->root_end);
        return KP_EINVAL;
    }
    // Cheap structural check, always on: every id indexes a vertex.  (Sorted,
    // symmetric, loop-free lists are assumed as in the reference; the full
    // check is KP_FLAG_VALIDATE.)
    if ((g->n && (!g->off || !g->deg)) || (g->m && !g->nbr) || (g->n && g->off[g->n] != g->m)) {
        set_error("malformed CSR");
        return KP_EINVAL;
    }
    for (uint64_t e = 0; e < g->m; e++)
        if (g->nbr[e] >= g->n) {
            set_error("neighbour id %u out of range (n=%u)", g->nbr[e], g->n);
            return KP_EINVAL;
        }
    if (p->flags & KP_FLAG_VALIDATE) {
        int rc = kp_graph_validate(g);
        if (rc) return rc;
    }
    return KP_OK;
}

} // namespace

extern "C" int kp_preprocess(const kp_graph *g, const int k, const int q, const int truss, kp_graph *out,
                             uint32_t *orig, int32_t *dseq)
{
    if (!g || !out || !orig || !dseq) return KP_EINVAL;
    kp_params p{};
    p.k = k;
    p.q = q;
    p.truss = truss;
    int rc = check_params(g, &p);
    if (rc) return rc;
    Prep pr;
    rc = preprocess(g, k, q, truss, &pr);
    if (rc) return rc;
    *out = pr.g;
    pr.g = kp_graph{};
    std::copy(pr.orig.begin(), pr.orig.end(), orig);
    std::copy(pr.dseq.begin(), pr.dseq.end(), dseq);
    return KP_OK;
}

extern "C" int kp_preprocess_gpu(const kp_graph *g, const int k, const int q, const int truss, kp_graph *out,
                                 uint32_t *orig, int32_t *dseq)
{
    if (!g || !out || !orig || !dseq) return KP_EINVAL;
    kp_params p{};
    p.k = k;
    p.q = q;
    p.truss = truss;
    int rc = check_params(g, &p);
    if (rc) return rc;
    Prep pr;
    rc = gpu_preprocess(g, k, q, truss, &pr);
    if (rc) return rc;
    *out = pr.g;
    pr.g = kp_graph{};
    std::copy(pr.orig.begin(), pr.orig.end(), orig);
    std::copy(pr.dseq.begin(), pr.dseq.end(), dseq);
    return KP_OK;
}

extern "C" int kp_prepare(const kp_graph *g, const kp_params *p, kp_prepared **out)
{
    if (!out) return KP_EINVAL;
    *out = nullptr;
    int rc = check_params(g, p);
    if (rc) return rc;
    return dev_prepare(g, p, out);
}

extern "C" int kp_run(kp_prepared *h, kp_result *r)
{
    if (!h || !r) return KP_EINVAL;
    return dev_run(h, r);
}

extern "C" uint64_t kp_prepared_bytes(const kp_prepared *h) { return dev_bytes(h); }

extern "C" int kp_set_window(kp_prepared *h, int64_t root_begin, int64_t root_end)
{
    if (!h || root_begin < 0 || root_end < 0 || (root_end && root_end < root_begin)) {
        set_error("invalid root window [%lld, %lld)", (long long)root_begin, (long long)root_end);
        return KP_EINVAL;
    }
    return dev_set_window(h, root_begin, root_end);
}

extern "C" int kp_autotune(kp_prepared *h, kp_tune *out)
{
    if (!h || !out) return KP_EINVAL;
    return dev_autotune(h, out);
}

extern "C" void kp_set_tune(const kp_tune *t) { dev_set_tune(t); }

extern "C" int kp_root_subgraph(kp_prepared *h, int64_t pos, uint32_t *block, uint32_t *n, uint32_t *hop,
                                uint32_t *left, uint32_t *l, uint32_t left_cap, uint8_t *pairs, size_t pairs_cap)
{
    if (!h || !block || !n || !hop || !left || !l) return KP_EINVAL;
    return dev_root_subgraph(h, pos, block, n, hop, left, l, left_cap, pairs, pairs_cap);
}

extern "C" void kp_prepared_free(kp_prepared *h) { dev_free(h); }

extern "C" int kp_enumerate(const kp_graph *g, const kp_params *p, kp_result *r)
{
    if (!r) return KP_EINVAL;
    std::memset(r, 0, sizeof(*r));
    int rc = check_params(g, p);
    if (rc) return rc;
    return dev_enumerate(g, p, r);
}

extern "C" int kp_enumerate_noalloc(const kp_graph *g, const kp_params *p, kp_result *r)
{
    if (!r) return KP_EINVAL;
    std::memset(r, 0, sizeof(*r));
    int rc = check_params(g, p);
    if (rc) return rc;
    if ((p->flags & KP_FLAG_LIST) || p->pause_every) {
        set_error("kp_enumerate_noalloc returns counts and hashes without pauses (no KP_FLAG_LIST, no pause_every)");
        return KP_EINVAL;
    }
    return dev_enumerate_noalloc(g, p, r);
}

extern "C" int kp_enumerate_batch(const kp_graph *const *g, const kp_params *p, size_t n, kp_result *r)
{
    if (n && (!g || !p || !r)) return KP_EINVAL;
    for (size_t i = 0; i < n; i++) {
        int rc = check_params(g[i], &p[i]);
        if (rc) {
            const std::string why = last_error();
            set_error("graph %zu: %s", i, why.c_str());
            return rc;
        }
    }
    return dev_batch(g, p, n, r);
}

extern "C" void kp_result_free(kp_result *r)
{
    if (!r) return;
    std::free(r->list_off);
    std::free(r->list_ids);
    r->list_off = nullptr;
    r->list_ids = nullptr;
    r->nlisted = 0;
}

extern "C" const char *kp_strerror(int code)
{
    switch (code) {
    case KP_OK: return "ok";
    case KP_ENOENT: return "file not found";
    case KP_EIO: return "short read or write";
    case KP_ENOMEM: return "out of memory";
    case KP_ENODEV: return "no HIP device";
    case KP_EINVAL: return "invalid argument";
    case KP_ERANGE: return "root block larger than KP_MAX_BLOCK";
    case KP_EOVERFLOW: return "internal capacity exceeded";
    case KP_ETIMEDOUT: return "time limit reached";
    case KP_EBUSY: return "admission refused";
    case KP_EHIP: return "HIP runtime error";
    }
    return "unknown error";
}

extern "C" const char *kp_last_error(void) { return kp::last_error(); }

extern "C" int kp_device_info(char *buf, size_t len)
{
    if (!buf || !len) return KP_EINVAL;
    return dev_info(buf, len);
}
