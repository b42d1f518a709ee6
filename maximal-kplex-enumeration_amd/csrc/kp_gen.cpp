// kp_gen.cpp -- deterministic synthetic graph generators (data layer).
//
// Random numbers come from a SplitMix64 stream (Steele, Lea & Flood 2014).
// Chung-Lu graphs are sampled with the O(n + m) skipping method of Miller &
// Hagberg, "Efficient generation of networks with given expected degrees"
// (WAW 2011): vertices sorted by descending weight, edge (u, v) kept with
// probability min(1, w_u w_v / S), geometric skips over runs of rejects.
#include "kp_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace kp;

namespace {

void add_undirected(std::vector<uint64_t> &keys, uint32_t u, uint32_t v)
{
    if (u == v) return;
    keys.push_back((uint64_t(u) << 32) | v);
    keys.push_back((uint64_t(v) << 32) | u);
}

// Weights w_i = max_deg * (i0 / (i + i0))^(1/(gamma-1)), i = 0..n-1; i0 is
// found by bisection so that sum_i w_i = 2 m_target.
std::vector<double> power_law_weights(uint32_t n, double gamma, uint64_t m_target, double max_deg)
{
    const double a = 1.0 / (gamma - 1.0);
    auto total = [&](double i0) {
        double s = 0;
        for (uint32_t i = 0; i < n; i++) s += max_deg * std::pow(i0 / (i + i0), a);
        return s;
    };
    const double want = 2.0 * double(m_target);
    double lo = 1e-9, hi = 1e12;
    for (int it = 0; it < 200; it++) {
        double mid = std::sqrt(lo * hi);
        if (total(mid) < want) lo = mid; else hi = mid;
    }
    std::vector<double> w(n);
    for (uint32_t i = 0; i < n; i++) w[i] = max_deg * std::pow(hi / (i + hi), a);
    return w;
}

} // namespace

extern "C" int kp_gen_er_planted(uint32_t n, double p, uint32_t groups, uint32_t gsize,
                                 uint32_t max_removed, uint64_t seed, kp_graph *out)
{
    if (!out || p < 0 || p > 1 || uint64_t(groups) * gsize > n) return KP_EINVAL;
    std::memset(out, 0, sizeof(*out));
    Rng rng(seed);
    std::vector<uint64_t> keys;
    // G(n, p) with geometric skipping over the n(n-1)/2 pairs
    if (p > 0) {
        const double lq = std::log1p(-p);
        long long v = 1, w = -1;
        while (v < (long long)n) {
            double r = rng.uniform();
            w += 1 + (p >= 1 ? 0 : (long long)std::floor(std::log1p(-r) / lq));
            while (w >= v && v < (long long)n) { w -= v; v++; }
            if (v < (long long)n) add_undirected(keys, uint32_t(v), uint32_t(w));
        }
    }
    // disjoint groups on a seeded random subset of vertices
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; i++) perm[i] = i;
    for (uint32_t i = 0; i < uint32_t(std::min<uint64_t>(n, uint64_t(groups) * gsize)); i++)
        std::swap(perm[i], perm[i + rng.below(n - i)]);
    std::vector<uint8_t> in_group(n, 0);
    std::vector<uint64_t> removed;
    for (uint32_t gi = 0; gi < groups; gi++) {
        const uint32_t *gv = perm.data() + size_t(gi) * gsize;
        for (uint32_t a = 0; a < gsize; a++) in_group[gv[a]] = uint8_t(1);
        std::vector<uint32_t> rem(gsize, 0);
        std::vector<uint8_t> gone(size_t(gsize) * gsize, 0);
        for (uint32_t a = 0; a < gsize; a++)
            for (uint32_t t = 0; t < max_removed; t++) {
                uint32_t b = rng.below(gsize);
                if (b == a || rem[a] >= max_removed || rem[b] >= max_removed || gone[size_t(a) * gsize + b]) continue;
                gone[size_t(a) * gsize + b] = gone[size_t(b) * gsize + a] = 1;
                rem[a]++;
                rem[b]++;
            }
        for (uint32_t a = 0; a < gsize; a++)
            for (uint32_t b = a + 1; b < gsize; b++) {
                if (gone[size_t(a) * gsize + b]) {
                    removed.push_back((uint64_t(gv[a]) << 32) | gv[b]);
                    removed.push_back((uint64_t(gv[b]) << 32) | gv[a]);
                } else {
                    add_undirected(keys, gv[a], gv[b]);
                }
            }
    }
    // inside a group the planted pattern wins over the random edges
    if (!removed.empty()) {
        std::sort(removed.begin(), removed.end());
        keys.erase(std::remove_if(keys.begin(), keys.end(),
                                  [&](uint64_t key) { return std::binary_search(removed.begin(), removed.end(), key); }),
                   keys.end());
    }
    return graph_from_edges(n, keys, out);
}

extern "C" int kp_gen_chung_lu(uint32_t n, double gamma, uint64_t m_target, double max_deg,
                               uint32_t ncomm, uint32_t cmin, uint32_t cmax, double dens,
                               uint64_t seed, kp_graph *out)
{
    if (!out || n < 2 || gamma <= 1.0 || max_deg <= 0 || cmin > cmax || dens < 0 || dens > 1) return KP_EINVAL;
    std::memset(out, 0, sizeof(*out));
    Rng rng(seed);
    std::vector<double> w = power_law_weights(n, gamma, m_target, max_deg);
    double S = 0;
    for (double x : w) S += x;
    // random placement of the weights over vertex ids
    std::vector<uint32_t> id(n);
    for (uint32_t i = 0; i < n; i++) id[i] = i;
    for (uint32_t i = n; i > 1; i--) std::swap(id[i - 1], id[rng.below(i)]);
    std::vector<uint64_t> keys;
    keys.reserve(size_t(2.2 * m_target) + 16);
    for (uint32_t u = 0; u + 1 < n; u++) {
        uint32_t v = u + 1;
        double p = std::min(w[u] * w[v] / S, 1.0);
        while (v < n && p > 0) {
            if (p < 1) {
                double r = rng.uniform();
                v += uint32_t(std::floor(std::log1p(-r) / std::log1p(-p)));
            }
            if (v < n) {
                double qv = std::min(w[u] * w[v] / S, 1.0);
                if (rng.uniform() < qv / p) add_undirected(keys, id[u], id[v]);
                p = qv;
                v++;
            }
        }
    }
    // planted communities on disjoint random vertex subsets
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; i++) perm[i] = i;
    uint32_t used = 0;
    for (uint32_t c = 0; c < ncomm; c++) {
        uint32_t sz = cmin + rng.below(cmax - cmin + 1);
        if (used + sz > n) break;
        for (uint32_t i = used; i < used + sz; i++) std::swap(perm[i], perm[i + rng.below(n - i)]);
        for (uint32_t a = used; a < used + sz; a++)
            for (uint32_t b = a + 1; b < used + sz; b++)
                if (rng.uniform() < dens) add_undirected(keys, perm[a], perm[b]);
        used += sz;
    }
    return graph_from_edges(n, keys, out);
}
