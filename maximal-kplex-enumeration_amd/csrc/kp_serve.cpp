// kp_serve.cpp -- batched and served front ends (no reference counterpart:
// the reference handles one graph per process).
#include "kp_device.h"

#include <cstring>
#include <deque>
#include <map>
#include <tuple>
#include <vector>

using namespace kp;

struct kp_server {
    kp_server_config cfg{};
    std::vector<kp_graph> graphs;
    struct Query {
        uint64_t ticket;
        uint32_t graph;
        kp_params p;
    };
    std::deque<Query> pending;
    uint64_t next_ticket = 1;
    std::map<std::tuple<uint32_t, int, int, int>, kp_prepared *> cache;
    ~kp_server()
    {
        for (auto &kv : cache) dev_free(kv.second);
        for (auto &g : graphs) kp_graph_free(&g);
    }
};

extern "C" int kp_server_create(const kp_server_config *cfg, kp_server **out)
{
    if (!out) return KP_EINVAL;
    auto *s = new (std::nothrow) kp_server();
    if (!s) return KP_ENOMEM;
    if (cfg) s->cfg = *cfg;
    if (!s->cfg.max_inflight) s->cfg.max_inflight = 64;
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

extern "C" int kp_server_submit(kp_server *s, uint32_t graph_id, const kp_params *p, uint64_t *ticket)
{
    if (!s || !p || !ticket || graph_id >= s->graphs.size()) return KP_EINVAL;
    if (p->k < 1 || p->q < 2 * p->k) return KP_EINVAL;
    *ticket = s->next_ticket++;
    s->pending.push_back({*ticket, graph_id, *p});
    return KP_OK;
}

extern "C" int kp_server_poll(kp_server *s, uint64_t *tickets, kp_result *results, size_t cap, size_t *ndone)
{
    if (!s || !ndone || (cap && (!tickets || !results))) return KP_EINVAL;
    *ndone = 0;
    while (*ndone < cap && !s->pending.empty()) {
        auto q = s->pending.front();
        s->pending.pop_front();
        auto key = std::make_tuple(q.graph, q.p.k, q.p.q, q.p.truss);
        kp_prepared *h = nullptr;
        auto it = s->cache.find(key);
        if (it == s->cache.end()) {
            int rc = dev_prepare(&s->graphs[q.graph], &q.p, &h);
            if (rc) return rc;
            s->cache[key] = h;
        } else {
            h = it->second;
        }
        int rc = dev_run(h, &results[*ndone]);
        if (rc) return rc;
        tickets[*ndone] = q.ticket;
        (*ndone)++;
    }
    return KP_OK;
}

extern "C" int kp_server_destroy(kp_server *s)
{
    delete s;
    return KP_OK;
}
