// kp_graph.cpp -- binary CSR graph I/O and host graph utilities.
//
// File format (inc/graph.h:45-62, write order inc/graphIO.h:19-25):
//   u32 n, u32 m, u32 deg[n], u32 nbrs[m]   (m = directed entries)
#include "kp_internal.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace kp {

int graph_alloc(kp_graph *g, uint32_t n, uint64_t m)
{
    g->n = n;
    g->m = m;
    g->off = static_cast<uint64_t *>(std::malloc(sizeof(uint64_t) * (size_t(n) + 1)));
    g->deg = static_cast<uint32_t *>(std::malloc(sizeof(uint32_t) * (n ? n : 1)));
    g->nbr = static_cast<uint32_t *>(std::malloc(sizeof(uint32_t) * (m ? m : 1)));
    if (!g->off || !g->deg || !g->nbr) {
        kp_graph_free(g);
        return KP_ENOMEM;
    }
    return KP_OK;
}

// Builds a simple undirected CSR from an edge list (u, v) pairs: both
// directions, sorted lists, duplicates and self loops dropped.
int graph_from_edges(uint32_t n, std::vector<uint64_t> &keys, kp_graph *out)
{
    // keys hold (u << 32 | v) for u != v, one per direction
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    int rc = graph_alloc(out, n, keys.size());
    if (rc) return rc;
    std::fill(out->deg, out->deg + n, 0u);
    for (uint64_t key : keys) out->deg[key >> 32]++;
    out->off[0] = 0;
    for (uint32_t i = 0; i < n; i++) out->off[i + 1] = out->off[i] + out->deg[i];
    for (size_t i = 0; i < keys.size(); i++) out->nbr[i] = uint32_t(keys[i] & 0xffffffffu);
    return KP_OK;
}

} // namespace kp

using namespace kp;

extern "C" void kp_graph_free(kp_graph *g)
{
    if (!g) return;
    std::free(g->off);
    std::free(g->nbr);
    std::free(g->deg);
    std::memset(g, 0, sizeof(*g));
}

extern "C" int kp_graph_load_bin(const char *path, kp_graph *g)
{
    if (!path || !g) return KP_EINVAL;
    std::memset(g, 0, sizeof(*g));
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        set_error("cannot open %s", path);
        return KP_ENOENT;
    }
    uint32_t hdr[2];
    int rc = KP_OK;
    if (std::fread(hdr, 4, 2, f) != 2) rc = KP_EIO;
    if (!rc) rc = graph_alloc(g, hdr[0], hdr[1]);
    if (!rc && std::fread(g->deg, 4, g->n, f) != g->n) rc = KP_EIO;
    if (!rc) {
        g->off[0] = 0;
        for (uint32_t i = 0; i < g->n; i++) g->off[i + 1] = g->off[i] + g->deg[i];
        if (g->off[g->n] != g->m) {
            set_error("%s: degrees sum to %llu, header says m=%llu", path,
                      (unsigned long long)g->off[g->n], (unsigned long long)g->m);
            rc = KP_EINVAL;
        }
    }
    if (!rc && std::fread(g->nbr, 4, g->m, f) != g->m) rc = KP_EIO;
    std::fclose(f);
    if (rc) {
        if (rc == KP_EIO) set_error("%s: short read", path);
        kp_graph_free(g);
    }
    return rc;
}

extern "C" int kp_graph_save_bin(const kp_graph *g, const char *path)
{
    if (!g || !path) return KP_EINVAL;
    if (g->m > 0xffffffffull) return KP_EINVAL;   // the format stores m as u32
    FILE *f = std::fopen(path, "wb");
    if (!f) return KP_ENOENT;
    uint32_t hdr[2] = {g->n, uint32_t(g->m)};
    int rc = KP_OK;
    if (std::fwrite(hdr, 4, 2, f) != 2) rc = KP_EIO;
    if (!rc && std::fwrite(g->deg, 4, g->n, f) != g->n) rc = KP_EIO;
    if (!rc && std::fwrite(g->nbr, 4, g->m, f) != g->m) rc = KP_EIO;
    if (std::fclose(f)) rc = KP_EIO;
    return rc;
}

extern "C" int kp_graph_from_csr(uint32_t n, const uint32_t *deg, const uint32_t *nbr, kp_graph *g)
{
    if (!g || (n && !deg)) return KP_EINVAL;
    std::memset(g, 0, sizeof(*g));
    uint64_t m = 0;
    for (uint32_t i = 0; i < n; i++) m += deg[i];
    if (m && !nbr) return KP_EINVAL;
    int rc = graph_alloc(g, n, m);
    if (rc) return rc;
    g->off[0] = 0;
    for (uint32_t i = 0; i < n; i++) {
        g->deg[i] = deg[i];
        g->off[i + 1] = g->off[i] + deg[i];
    }
    if (m) std::memcpy(g->nbr, nbr, m * 4);
    return KP_OK;
}

extern "C" int kp_graph_validate(const kp_graph *g)
{
    if (!g || (g->n && (!g->off || !g->deg)) || (g->m && !g->nbr)) return KP_EINVAL;
    if (g->off[0] != 0 || g->off[g->n] != g->m) return KP_EINVAL;
    for (uint32_t u = 0; u < g->n; u++) {
        if (g->off[u + 1] - g->off[u] != g->deg[u]) {
            set_error("vertex %u: degree %u does not match offsets", u, g->deg[u]);
            return KP_EINVAL;
        }
        for (uint64_t p = g->off[u]; p < g->off[u + 1]; p++) {
            uint32_t v = g->nbr[p];
            if (v >= g->n || v == u || (p > g->off[u] && g->nbr[p - 1] >= v)) {
                set_error("vertex %u: neighbour list not sorted/simple at entry %llu", u,
                          (unsigned long long)(p - g->off[u]));
                return KP_EINVAL;
            }
            const uint32_t *b = g->nbr + g->off[v], *e = g->nbr + g->off[v + 1];
            if (!std::binary_search(b, e, u)) {
                set_error("edge %u-%u is not symmetric", u, v);
                return KP_EINVAL;
            }
        }
    }
    return KP_OK;
}

extern "C" int kp_graph_relabel(const kp_graph *g, uint64_t seed, kp_graph *out, uint32_t *perm_out)
{
    if (!g || !out) return KP_EINVAL;
    std::memset(out, 0, sizeof(*out));
    std::vector<uint32_t> perm(g->n);
    for (uint32_t i = 0; i < g->n; i++) perm[i] = i;
    Rng rng(seed);
    for (uint32_t i = g->n; i > 1; i--) std::swap(perm[i - 1], perm[rng.below(i)]);
    std::vector<uint64_t> keys;
    keys.reserve(g->m);
    for (uint32_t u = 0; u < g->n; u++)
        for (uint64_t p = g->off[u]; p < g->off[u + 1]; p++)
            keys.push_back((uint64_t(perm[u]) << 32) | perm[g->nbr[p]]);
    int rc = graph_from_edges(g->n, keys, out);
    if (!rc && perm_out) std::memcpy(perm_out, perm.data(), sizeof(uint32_t) * g->n);
    return rc;
}
