// kp_internal.h -- shared declarations of the host side of libkplex.
#pragma once
#include "../../include/kplex.h"

#include <cstdint>
#include <vector>

namespace kp {

// SplitMix64 stream (Steele, Lea & Flood 2014).
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    static uint64_t mix(uint64_t z)
    {
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    uint64_t next() { return mix(s += 0x9E3779B97F4A7C15ULL); }
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint32_t below(uint32_t n) { return n ? uint32_t((unsigned __int128)next() * n >> 64) : 0; }
};

void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

int graph_alloc(kp_graph *g, uint32_t n, uint64_t m);
int graph_from_edges(uint32_t n, std::vector<uint64_t> &keys, kp_graph *out);

// Host preprocessing result: the (q-k)-core (after the optional truss and
// second core), its new->original id map and the degeneracy order.
struct Prep {
    kp_graph g{};
    std::vector<uint32_t> orig;
    std::vector<int32_t> dseq, dpos;
    uint32_t pn_first = 0;   // vertices after the first core (before the truss)
    double core_ms = 0, support_ms = 0, truss_ms = 0, degeneracy_ms = 0;   // host time per stage
    uint32_t maxdeg = 0, d1max = 0;   // largest degree / largest later-neighbour count (when sizes_known)
    bool sizes_known = false;
    Prep() = default;
    Prep(const Prep &) = delete;
    Prep &operator=(const Prep &) = delete;
    ~Prep() { kp_graph_free(&g); }
};
int preprocess(const kp_graph *g, int k, int q, int truss, Prep *out);
// the stages of preprocess, shared with the GPU preprocessing (kp_gprep.hip)
int core_peel(const kp_graph *g, int bd, kp_graph *out, std::vector<uint32_t> &old_of_new);
void truss_supports(const kp_graph *g, std::vector<int32_t> &tri);
int truss_peel(kp_graph *g, int q, int k, std::vector<int32_t> &tri);
// max_later (optional) receives the largest number of later neighbours of a vertex
void degeneracy_order(const kp_graph *g, std::vector<int32_t> &dseq, std::vector<int32_t> &dpos,
                      uint32_t *max_later = nullptr);
// after the first core (out->g, old ids o1): the optional truss, with the
// supports of out->g's entries in *tri (counted here when tri is null), the
// second core, the degeneracy order
int finish_preprocess(int k, int q, int truss, std::vector<uint32_t> &o1, std::vector<int32_t> *tri, Prep *out);
// the same result as preprocess, with the first core, the relabel and the
// truss supports computed on the current HIP device (kp_gprep.hip)
int gpu_preprocess(const kp_graph *g, int k, int q, int truss, Prep *out);

double now_ms();

// hash constants of the result hash (documented in kplex.h)
constexpr uint64_t HASH_SA = 0x243F6A8885A308D3ULL;
constexpr uint64_t HASH_LA = 0x13198A2E03707344ULL;
constexpr uint64_t HASH_SB = 0xA4093822299F31D0ULL;
constexpr uint64_t HASH_LB = 0x082EFA98EC4E6C89ULL;

} // namespace kp
