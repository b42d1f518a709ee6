// kp_device.hip -- MI355X (gfx950) kernels and device driver of libkplex.
//
// Pipeline per (graph, k, q):
//   decompose_kernel   one wave64 per root (grid-stride): later/earlier split of
//                      N(r), (q-2k)-core of the later 1-hop on a 64-bit bitset
//                      adjacency, left filter and 2-hop candidates counted with a
//                      per-wave counter array.   [reference: decompose,
//                      src/device_funcs.cu:28-298]
//   build_kernel<WC>   one wave per valid root: bitset adjacency rows, the two
//                      pair-class bitsets (class >= UNLINK2EQUAL, class >=
//                      UNLINK2MORE), left rows, original ids.   [calculateDegrees,
//                      fillNeighbors, buildCommonMtx :300-460, :1727-1853]
//   search_kernel<W>   persistent: every wave64 is a worker with a DFS stack of
//                      (P, C, C2, X) bitset frames; k-search frames [kSearch
//                      :1855-2444] and Branch frames [BNB :1516-1724] share the
//                      stack; idle waves take roots from an atomic cursor or
//                      frames donated by busy waves through a bounded MPMC queue
//                      in HBM.
//
// Wave layout: lane L of a wave owns the block vertices L, 64+L, ... (word j of
// every bitset holds vertices 64j..64j+63, bit L = vertex 64j+L).  Sets are
// wave-uniform (SGPRs); per-vertex degrees are per-lane popcounts of the owned
// vertices' adjacency rows, which stay in VGPRs while a wave works on a root.
#include <hip/hip_runtime.h>

#include "kp_device.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

namespace kpd {

constexpr uint32_t NCAP = KP_MAX_BLOCK;          // largest block
constexpr int WPB = 4;                           // waves per block (256 threads)
constexpr uint32_t EXCL = 0x80000000u;           // counter-array flags
constexpr uint32_t CLAIM = 0x40000000u;
constexpr uint32_t CMASK = 0x3FFFFFFFu;
constexpr uint32_t SPIN_LIMIT = 1u << 24;

struct RootInfo {          // decomposition result per root position
    int32_t status;        // 1 valid, 0 not, <0 error
    uint32_t n, h, l;
    uint64_t wr;           // CSR entries read: d(r) + sum over later neighbours, then + sum over B and L
};

struct RootDesc {          // one valid root in the subgraph arena
    uint64_t off;          // u64-word offset in the arena
    uint32_t n, h, l, pos;
};

struct Sched {             // persistent-kernel scheduler state (one per launch)
    uint32_t next_root;
    uint32_t active;
    uint32_t q_count;      // published frames over all sub-queues
    uint32_t hungry;       // waves currently looking for work
    uint64_t q_head, q_tail;   // (unused; kept for layout)
    int32_t error;
    uint32_t exited;
    uint64_t count, ha, hb, tasks, alg;
    uint64_t list_used;
    uint64_t donated;
    uint64_t frames, spills, pops;
    uint64_t busy_ticks, idle_ticks;   // wall-clock ticks summed over waves
};

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

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ uint64_t lt_mask() { return (1ull << __lane_id()) - 1ull; }
__device__ __forceinline__ uint32_t rfl(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t rfl64(uint64_t x)
{
    return (uint64_t(rfl(uint32_t(x >> 32))) << 32) | rfl(uint32_t(x));
}
// Wave64 reductions: four DPP steps (quad_perm [1,0,3,2], [2,3,0,1],
// row_half_mirror, row_mirror) reduce each 16-lane row, then the four row
// results are read out with v_readlane.  The result is wave-uniform.
#define KP_DPP(v, ctrl) uint32_t(__builtin_amdgcn_update_dpp(0, int(v), ctrl, 0xF, 0xF, false))
__device__ __forceinline__ uint32_t wmin32(uint32_t v)
{
    v = min(v, KP_DPP(v, 0xB1));
    v = min(v, KP_DPP(v, 0x4E));
    v = min(v, KP_DPP(v, 0x141));
    v = min(v, KP_DPP(v, 0x140));
    const uint32_t a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
    const uint32_t c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
    return min(min(a, b), min(c, d));
}
__device__ __forceinline__ uint64_t dpp64(uint64_t v, int ctrl_sel)
{
    uint32_t lo = uint32_t(v), hi = uint32_t(v >> 32);
    switch (ctrl_sel) {
    case 0: lo = KP_DPP(lo, 0xB1); hi = KP_DPP(hi, 0xB1); break;
    case 1: lo = KP_DPP(lo, 0x4E); hi = KP_DPP(hi, 0x4E); break;
    case 2: lo = KP_DPP(lo, 0x141); hi = KP_DPP(hi, 0x141); break;
    default: lo = KP_DPP(lo, 0x140); hi = KP_DPP(hi, 0x140); break;
    }
    return (uint64_t(hi) << 32) | lo;
}
__device__ __forceinline__ uint64_t readlane64(uint64_t x, int l)
{
    const uint32_t lo = __builtin_amdgcn_readlane(int(uint32_t(x)), l);
    const uint32_t hi = __builtin_amdgcn_readlane(int(uint32_t(x >> 32)), l);
    return (uint64_t(hi) << 32) | lo;
}
__device__ __forceinline__ uint64_t wsum64(uint64_t v)
{
    v += dpp64(v, 0);
    v += dpp64(v, 1);
    v += dpp64(v, 2);
    v += dpp64(v, 3);
    return readlane64(v, 0) + readlane64(v, 16) + readlane64(v, 32) + readlane64(v, 48);
}
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint32_t ald(uint32_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void ast(uint32_t *p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void aadd(uint32_t *p, uint32_t v) { __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ uint32_t aor(uint32_t *p, uint32_t v) { return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void fence_agent() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent"); }
__device__ __forceinline__ void fence_wg() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

// ============================================================ decompose
struct DecompArgs {
    const uint64_t *off;
    const uint32_t *nbr, *deg;
    const int32_t *dpos, *dseq;
    uint32_t pn;
    int k, q;
    int64_t pos_begin, nroots;   // root positions [pos_begin, pos_begin + nroots)
    uint32_t *blk, *lft;
    uint32_t lcap;
    RootInfo *info;
    uint32_t *ws_cnt, *ws_b1c;
    uint64_t *ws_a1;
    uint32_t d1cap, w1cap, nwaves;
};

// Root r = dseq[pos] (src/device_funcs.cu:28-298), restated on bitsets:
//  B1 = (q-2k)-core of G[later neighbours]   (fixpoint :139-171; the core is
//       unique, so any peeling order gives the reference's set)
//  L1 = earlier neighbours with >= q-2k+1 neighbours in B1          (:177-209)
//  B2 = later non-neighbours with >= q-2k+2 neighbours in B1        (:211-273)
//  L2 = earlier non-neighbours with >= q-2k+3 neighbours in B1
// Canonical layout: B = [r, B1 asc, B2 asc], L = [L1 asc, L2].
__global__ __launch_bounds__(256) void decompose_kernel(DecompArgs a)
{
    __shared__ uint64_t s_row[WPB][64];
    __shared__ uint64_t s_alive[WPB][64];
    __shared__ uint32_t s_b2[WPB][NCAP];
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    const uint32_t gw = blockIdx.x * WPB + wv;
    if (gw >= a.nwaves) return;
    uint32_t *cnt = a.ws_cnt + size_t(gw) * a.pn;
    uint32_t *b1c = a.ws_b1c + size_t(gw) * a.d1cap;
    uint64_t *A1 = a.ws_a1 + size_t(gw) * a.d1cap * a.w1cap;
    uint64_t *row = s_row[wv], *alive = s_alive[wv];
    uint32_t *sb2 = s_b2[wv];
    const int lb2k = a.q - 2 * a.k, bd = a.q - a.k;

    for (int64_t slot = gw; slot < a.nroots; slot += a.nwaves) {
        const int64_t pos = a.pos_begin + slot;
        const uint32_t r = uint32_t(a.dseq[pos]);
        const uint64_t ro = a.off[r];
        const uint32_t dr = a.deg[r];
        uint32_t *B = a.blk + size_t(slot) * NCAP;
        uint32_t *L = a.lft + size_t(slot) * a.lcap;
        RootInfo info{0, 0, 0, 0, dr};

        // 1) split N(r) keeping list (ascending id) order
        uint32_t m1 = 0, ml = 0;
        uint64_t wsum = 0;
        for (uint32_t base = 0; base < dr; base += 64) {
            const uint32_t idx = base + ln;
            const bool v = idx < dr;
            const uint32_t w = v ? a.nbr[ro + idx] : 0u;
            const bool later = v && a.dpos[w] > int32_t(pos);
            const bool earlier = v && !later;
            const uint64_t mL = __ballot(later), mE = __ballot(earlier);
            if (later) {
                const uint32_t t = m1 + __popcll(mL & lt_mask());
                if (t < a.d1cap) b1c[t] = w;
                wsum += a.deg[w];
            }
            if (earlier) {
                const uint32_t t = ml + __popcll(mE & lt_mask());
                if (t < a.lcap) L[t] = w;
            }
            m1 += __popcll(mL);
            ml += __popcll(mE);
        }
        info.wr += rfl64(wsum64(wsum));
        if (int(m1) < bd) {   // |B| - 1 < q - k before any filtering
            if (ln == 0) a.info[slot] = info;
            continue;
        }
        if (m1 > a.d1cap || m1 > 64u * 64u) {
            info.status = -1;
            if (ln == 0) a.info[slot] = info;
            continue;
        }
        const uint32_t w1 = (m1 + 63) >> 6;

        // 2) bitset adjacency among the later neighbours (index map in cnt[])
        for (uint32_t t = ln; t < m1; t += 64) ast(&cnt[b1c[t]], t + 1);
        fence_agent();
        for (uint32_t i = 0; i < m1; i++) {
            const uint32_t u = b1c[i];
            const uint32_t du = a.deg[u];
            if (int(du) < lb2k) continue;          // cannot be in the core: row never read
            const uint64_t uo = a.off[u];
            for (uint32_t t = ln; t < w1; t += 64) row[t] = 0;
            fence_wg();
            for (uint32_t base = 0; base < du; base += 64) {
                const uint32_t idx = base + ln;
                if (idx < du) {
                    const uint32_t c = ald(&cnt[a.nbr[uo + idx]]);
                    if (c) {
                        const uint32_t j = c - 1;
                        atomicOr((unsigned long long *)&row[j >> 6], 1ull << (j & 63));
                    }
                }
            }
            fence_wg();
            for (uint32_t t = ln; t < w1; t += 64) A1[size_t(i) * a.w1cap + t] = row[t];
        }
        for (uint32_t t = ln; t < m1; t += 64) ast(&cnt[b1c[t]], 0);
        for (uint32_t t = 0; t < w1; t++) {
            const uint32_t i = 64 * t + ln;
            const bool al = i < m1 && int(a.deg[b1c[i]]) >= lb2k;
            const uint64_t m = __ballot(al);
            if (ln == 0) alive[t] = m;
        }
        fence_agent();

        // 3) (q-2k)-core of the later neighbourhood
        for (;;) {
            bool changed = false;
            for (uint32_t t = 0; t < w1; t++) {
                const uint64_t am = alive[t];
                const bool mine = (am >> ln) & 1;
                const uint32_t i = 64 * t + ln;
                int c = 0;
                if (mine)
                    for (uint32_t w = 0; w < w1; w++) c += __popcll(A1[size_t(i) * a.w1cap + w] & alive[w]);
                const uint64_t dead = __ballot(mine && c < lb2k);
                if (dead) {
                    if (ln == 0) alive[t] = am & ~dead;
                    changed = true;
                    fence_wg();
                }
            }
            if (!changed) break;
        }
        uint32_t nb1 = 0;
        for (uint32_t t = 0; t < w1; t++) nb1 += __popcll(alive[t]);
        if (int(nb1) < bd) {
            if (ln == 0) a.info[slot] = info;
            continue;
        }
        const uint32_t hop = nb1 + 1;
        if (hop > NCAP) {
            info.status = -1;
            if (ln == 0) a.info[slot] = info;
            continue;
        }
        if (ln == 0) B[0] = r;
        {
            uint32_t wp = 1;
            for (uint32_t t = 0; t < w1; t++) {
                const uint64_t am = alive[t];
                if ((am >> ln) & 1) B[wp + __popcll(am & lt_mask())] = b1c[64 * t + ln];
                wp += __popcll(am);
            }
        }
        fence_agent();

        // 4) mark r and N(r); 5) count |N(w) ∩ B1| for every w adjacent to B1
        if (ln == 0) aor(&cnt[r], EXCL);
        for (uint32_t t = ln; t < dr; t += 64) aor(&cnt[a.nbr[ro + t]], EXCL);
        for (uint32_t b = 1; b < hop; b++) {
            const uint32_t u = B[b];
            const uint64_t uo = a.off[u];
            const uint32_t du = a.deg[u];
            for (uint32_t t = ln; t < du; t += 64) aadd(&cnt[a.nbr[uo + t]], 1);
        }
        fence_agent();

        // 6) left filter, in place (list order kept)
        const uint32_t mlc = min(ml, a.lcap);
        uint32_t nl1 = 0;
        for (uint32_t base = 0; base < mlc; base += 64) {
            const uint32_t idx = base + ln;
            const bool v = idx < mlc;
            const uint32_t u = v ? L[idx] : 0u;
            const bool keep = v && int(ald(&cnt[u]) & CMASK) >= lb2k + 1;
            const uint64_t mk = __ballot(keep);
            if (keep) L[nl1 + __popcll(mk & lt_mask())] = u;
            nl1 += __popcll(mk);
        }

        // 7) 2-hop: first lane to see a passing w claims it
        uint32_t nb2 = 0, nl2 = 0;
        for (uint32_t b = 1; b < hop; b++) {
            const uint32_t u = B[b];
            const uint64_t uo = a.off[u];
            const uint32_t du = a.deg[u];
            for (uint32_t base = 0; base < du; base += 64) {
                const uint32_t idx = base + ln;
                bool own = false, later = false;
                uint32_t w = 0;
                if (idx < du) {
                    w = a.nbr[uo + idx];
                    const uint32_t c = ald(&cnt[w]);
                    if (!(c & (EXCL | CLAIM))) {
                        later = a.dpos[w] > int32_t(pos);
                        const int thr = lb2k + (later ? 2 : 3);
                        if (int(c & CMASK) >= thr) own = !(aor(&cnt[w], CLAIM) & CLAIM);
                    }
                }
                const uint64_t mLa = __ballot(own && later), mEa = __ballot(own && !later);
                if (own && later) {
                    const uint32_t p = hop + nb2 + __popcll(mLa & lt_mask());
                    if (p < NCAP) B[p] = w;
                }
                if (own && !later) {
                    const uint32_t p = nl1 + nl2 + __popcll(mEa & lt_mask());
                    if (p < a.lcap) L[p] = w;
                }
                nb2 += __popcll(mLa);
                nl2 += __popcll(mEa);
            }
        }
        // 8) restore the counter array to zero
        fence_agent();
        if (ln == 0) ast(&cnt[r], 0);
        for (uint32_t t = ln; t < dr; t += 64) ast(&cnt[a.nbr[ro + t]], 0);
        for (uint32_t b = 1; b < hop; b++) {
            const uint32_t u = B[b];
            const uint64_t uo = a.off[u];
            const uint32_t du = a.deg[u];
            for (uint32_t t = ln; t < du; t += 64) ast(&cnt[a.nbr[uo + t]], 0);
        }
        fence_agent();
        const uint32_t n = hop + nb2, l = nl1 + nl2;
        info.n = n;
        info.h = hop;
        info.l = l;
        if (n > NCAP) info.status = -1;
        else if (ml > a.lcap || l > a.lcap) info.status = -2;
        else info.status = int(n) >= a.q ? 1 : 0;
        // 9) 2-hop later vertices in ascending id order (rank sort in LDS)
        if (info.status == 1 && nb2 > 1) {
            for (uint32_t t = ln; t < nb2; t += 64) sb2[t] = B[hop + t];
            fence_wg();
            for (uint32_t t = ln; t < nb2; t += 64) {
                const uint32_t x = sb2[t];
                uint32_t rank = 0;
                for (uint32_t j = 0; j < nb2; j++) rank += sb2[j] < x;
                B[hop + rank] = x;
            }
            fence_agent();
        }
        if (ln == 0) a.info[slot] = info;
    }
}

// ============================================================ build
struct BuildArgs {
    const uint64_t *off;
    const uint32_t *nbr, *deg, *orig;
    const uint32_t *blk, *lft;
    uint32_t lcap;
    const RootDesc *desc;
    uint32_t ndesc;
    uint64_t *arena;
    uint32_t *ws_cnt;
    uint32_t pn, nwaves;
    int k, q;
    uint64_t *wr_extra;   // per desc: sum of degrees over B and L
};

// Per root: adjacency rows adj[i] (bit j = B[j] adjacent to B[i]), class rows
// cls2[i] (bit j: class(i,j) >= UNLINK2EQUAL) and cls4[i] (>= UNLINK2MORE)
// with the thresholds of buildCommonMtx (src/device_funcs.cu:1768-1773) over
// common = |N(i) ∩ N(j) ∩ B[0,h)|; left rows ladj[u] (bit i = L[u] adjacent
// to B[i]); original ids.  Row stride is the class width WC.
template <int WC>
__global__ __launch_bounds__(256) void build_kernel(BuildArgs a)
{
    __shared__ uint64_t s_row[WPB][WC];
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    const uint32_t gw = blockIdx.x * WPB + wv;
    if (gw >= a.nwaves) return;
    uint32_t *cnt = a.ws_cnt + size_t(gw) * a.pn;
    uint64_t *row = s_row[wv];
    const int k = a.k, q = a.q;
    const int km2 = max(k - 2, 0), km3 = max(k - 3, 0);
    const int thPP1 = q - k - 2 * km2, thPP2 = q - k - 2 * km3;
    const int thPC1 = q - 2 * k - km2, thPC2 = q - k - 1 - km2 - km3;
    const int thCC1 = q - 2 * k - (k - 1), thCC2 = q - 2 * k + 2 - (k - 1);

    for (uint32_t di = gw; di < a.ndesc; di += a.nwaves) {
        const RootDesc d = a.desc[di];
        const uint32_t n = d.n, h = d.h, l = d.l;
        const uint32_t *B = a.blk + size_t(d.pos) * NCAP;
        const uint32_t *L = a.lft + size_t(d.pos) * a.lcap;
        uint64_t *adj = a.arena + d.off;
        uint64_t *c2 = adj + size_t(n) * WC;
        uint64_t *c4 = c2 + size_t(n) * WC;
        uint64_t *ladj = c4 + size_t(n) * WC;
        uint32_t *origp = reinterpret_cast<uint32_t *>(ladj + size_t(l) * WC);

        for (uint32_t j = ln; j < n; j += 64) ast(&cnt[B[j]], j + 1);
        for (uint32_t j = ln; j < l; j += 64) ast(&cnt[L[j]], CLAIM | (j + 1));
        fence_agent();
        uint64_t wsum = 0;
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t u = B[i];
            const uint64_t uo = a.off[u];
            const uint32_t du = a.deg[u];
            if (ln < WC) row[ln] = 0;
            fence_wg();
            for (uint32_t base = 0; base < du; base += 64) {
                const uint32_t idx = base + ln;
                if (idx < du) {
                    const uint32_t c = ald(&cnt[a.nbr[uo + idx]]);
                    if (c & CLAIM) {
                        const uint32_t j = (c & CMASK) - 1;
                        atomicOr((unsigned long long *)&ladj[size_t(j) * WC + (i >> 6)], 1ull << (i & 63));
                    } else if (c) {
                        const uint32_t j = c - 1;
                        atomicOr((unsigned long long *)&row[j >> 6], 1ull << (j & 63));
                    }
                }
            }
            fence_wg();
            if (ln < WC) adj[size_t(i) * WC + ln] = row[ln];
            if (ln == 0) wsum += du;
        }
        for (uint32_t j = ln; j < l; j += 64) wsum += a.deg[L[j]];
        for (uint32_t j = ln; j < n; j += 64) origp[j] = a.orig[B[j]];
        wsum = wsum64(wsum);
        if (ln == 0) a.wr_extra[di] = wsum;
        fence_agent();

        uint64_t H[WC];
#pragma unroll
        for (int w = 0; w < WC; w++) {
            const int lo = 64 * w;
            H[w] = int(h) >= lo + 64 ? ~0ull : (int(h) <= lo ? 0ull : ((1ull << (h - lo)) - 1ull));
        }
#pragma unroll
        for (int t = 0; t < WC; t++) {
            if (64u * t >= n) break;
            const uint32_t i = 64 * t + ln;
            const bool act = i < n;
            uint64_t Ri[WC], r2[WC], r4[WC];
#pragma unroll
            for (int w = 0; w < WC; w++) {
                Ri[w] = act ? adj[size_t(i) * WC + w] : 0ull;
                r2[w] = 0;
                r4[w] = 0;
            }
#pragma unroll
            for (int jw = 0; jw < WC; jw++) {
                for (uint32_t b = 0; b < 64; b++) {
                    const uint32_t j = 64 * jw + b;
                    if (j >= n) break;
                    int common = 0;
#pragma unroll
                    for (int w = 0; w < WC; w++) common += __popcll(Ri[w] & adj[size_t(j) * WC + w] & H[w]);
                    const bool linked = (Ri[jw] >> b) & 1;
                    const bool ih = i < h, jh = j < h;
                    int t1, t2;
                    bool raw = false;
                    if (ih && jh) { t1 = thCC1; t2 = thCC2; }
                    else if (ih || jh) { t1 = thPC1; t2 = thPC2; }
                    else { t1 = thPP1; t2 = thPP2; raw = (k == 2); }
                    int cls;
                    if (raw) cls = linked ? 1 : 0;
                    else if (linked) cls = common > t1 ? 5 : (common == t1 ? 3 : 1);
                    else cls = common > t2 ? 4 : (common == t2 ? 2 : 0);
                    if (j == i) cls = 0;
                    if (cls >= 2) r2[jw] |= 1ull << b;
                    if (cls >= 4) r4[jw] |= 1ull << b;
                }
            }
            if (act) {
#pragma unroll
                for (int w = 0; w < WC; w++) {
                    c2[size_t(i) * WC + w] = r2[w];
                    c4[size_t(i) * WC + w] = r4[w];
                }
            }
        }
        for (uint32_t j = ln; j < n; j += 64) ast(&cnt[B[j]], 0);
        for (uint32_t j = ln; j < l; j += 64) ast(&cnt[L[j]], 0);
        fence_agent();
    }
}

// ============================================================ search
struct SearchArgs {
    const RootDesc *desc;
    uint32_t ndesc;
    const uint64_t *arena;
    int k, q;
    uint64_t *qframes, *qseq;  // NSUBQ * qcap slots
    uint32_t qcap;             // slots per shard
    SubQ *subq;
    Sched *sched;
    uint32_t *list;
    uint64_t list_cap;
    uint64_t *list_used;     // shared by all size-class launches
    uint32_t nwaves;
    uint32_t ring_cap;       // frames per wave in the LDS ring
    uint32_t back_cap;       // frames per wave in the HBM backing stack
    uint64_t *gstack;        // nwaves * back_cap frames
    uint32_t donate_period, donate_target;
    uint64_t limit_ticks;    // 0 = no limit; else wall-clock ticks per launch
};

enum : uint32_t { FT_BRANCH = 0, FT_KSEARCH = 1 };

template <int W>
struct Frame {
    uint64_t P[W], C[W], D[W], X[W];
    uint32_t root, type, r;
};

// frame slot: [meta, P[W], C[W], X[W], D[W], pad] -- an even number of
// 64-bit words so that slots stay 16-byte aligned
template <int W>
struct FrameLayout {
    static constexpr int FW = (4 * W + 2) & ~1;
};

template <int W>
__device__ __forceinline__ uint64_t rdl64(uint64_t x, uint32_t l)
{
    const uint32_t lo = __builtin_amdgcn_readlane(int(uint32_t(x)), int(l));
    const uint32_t hi = __builtin_amdgcn_readlane(int(uint32_t(x >> 32)), int(l));
    return (uint64_t(hi) << 32) | lo;
}

// LW: left-vertex rows held in registers (left vertices 64t+lane, t < LW)
template <int W, int LW>
struct Worker {
    static constexpr int FW = FrameLayout<W>::FW;
    const SearchArgs &a;
    Sched *S;
    uint64_t *ring, *back;
    uint32_t ln;
    int k, q;
    // two-level stack (uniform): the newest frames in the LDS ring, older ones in HBM
    uint32_t rb = 0, rn = 0, bb = 0, bn = 0;
    // current root (uniform)
    uint32_t cur = 0xffffffffu, n = 0, h = 0, l = 0;
    const uint64_t *LA = nullptr;
    // rows of the vertices this lane owns (64j + lane), kept in VGPRs
    uint64_t adjR[W][W], c2R[W][W], c4R[W][W];
    uint64_t lftR[LW > 0 ? LW : 1][W];
    uint32_t origR[W];
    // accumulators (uniform)
    uint64_t cnt = 0, ha = 0, hb = 0, tasks = 0, alg = 0, frames = 0, spills = 0, pops = 0;
    int err = 0;
    uint64_t t_start = 0;

    __device__ Worker(const SearchArgs &args, uint64_t *ring_, uint64_t *back_)
        : a(args), S(args.sched), ring(ring_), back(back_)
    {
        ln = __lane_id();
        k = a.k;
        q = a.q;
        t_start = wall_clock64();
    }
    __device__ __forceinline__ bool past_deadline() const
    {
        return a.limit_ticks && wall_clock64() - t_start > a.limit_ticks;
    }
    __device__ __forceinline__ void set_error(int code)
    {
        if (ln == 0) atomicCAS(&S->error, 0, code);
        err = code;
    }

    // ---- frame slots (only lane 0 touches them; values broadcast with readfirstlane) ----
    __device__ __forceinline__ void store_slot(uint64_t *dst, uint64_t meta, const uint64_t *P, const uint64_t *C,
                                               const uint64_t *X, const uint64_t *D, bool withD)
    {
        if (ln == 0) {
            dst[0] = meta;
#pragma unroll
            for (int w = 0; w < W; w++) {
                dst[1 + w] = P[w];
                dst[1 + W + w] = C[w];
                dst[1 + 2 * W + w] = X[w];
                if (withD) dst[1 + 3 * W + w] = D[w];
            }
        }
    }
    __device__ __forceinline__ void load_slot(const uint64_t *src, Frame<W> &f)
    {
        uint64_t meta = 0;
        uint64_t t[4 * W];
        if (ln == 0) {
            meta = src[0];
#pragma unroll
            for (int i = 0; i < 3 * W; i++) t[i] = src[1 + i];
        }
        meta = rfl64(meta);
        f.root = uint32_t(meta);
        f.type = uint32_t(meta >> 32) & 0xff;
        f.r = uint32_t(meta >> 40) & 0xff;
        if (f.type == FT_KSEARCH && ln == 0) {
#pragma unroll
            for (int i = 0; i < W; i++) t[3 * W + i] = src[1 + 3 * W + i];
        }
#pragma unroll
        for (int w = 0; w < W; w++) {
            f.P[w] = rfl64(t[w]);
            f.C[w] = rfl64(t[W + w]);
            f.X[w] = rfl64(t[2 * W + w]);
            f.D[w] = f.type == FT_KSEARCH ? rfl64(t[3 * W + w]) : 0ull;
        }
    }
    static __device__ __forceinline__ uint64_t make_meta(uint32_t root, uint32_t type, uint32_t r)
    {
        return uint64_t(root) | (uint64_t(type) << 32) | (uint64_t(r) << 40);
    }
    __device__ __forceinline__ void copy_slot(uint64_t *dst, const uint64_t *src)
    {
        if (ln == 0) {
#pragma unroll
            for (int i = 0; i < FW; i++) dst[i] = src[i];
        }
    }
    __device__ __forceinline__ uint64_t *ring_slot(uint32_t i) const
    {
        uint32_t s = rb + i;
        if (s >= a.ring_cap) s -= a.ring_cap;
        return ring + size_t(s) * FW;
    }
    __device__ __forceinline__ uint64_t *back_slot(uint32_t i) const
    {
        uint32_t s = bb + i;
        if (s >= a.back_cap) s -= a.back_cap;
        return back + size_t(s) * FW;
    }
    __device__ __forceinline__ uint32_t depth() const { return rn + bn; }

    __device__ bool push(uint64_t meta, const uint64_t *P, const uint64_t *C, const uint64_t *X, const uint64_t *D,
                         bool withD)
    {
        if (rn == a.ring_cap) {   // spill the oldest ring frame to the HBM backing stack
            if (bn == a.back_cap) {
                set_error(KP_EOVERFLOW);
                return false;
            }
            copy_slot(back_slot(bn), ring_slot(0));
            bn++;
            rb = (rb + 1 == a.ring_cap) ? 0 : rb + 1;
            rn--;
            spills++;
        }
        store_slot(ring_slot(rn), meta, P, C, X, D, withD);
        rn++;
        return true;
    }
    __device__ __forceinline__ bool push(const Frame<W> &f)
    {
        return push(make_meta(f.root, f.type, f.r), f.P, f.C, f.X, f.D, f.type == FT_KSEARCH);
    }
    __device__ void pop_top(Frame<W> &f)
    {
        if (rn) {
            rn--;
            load_slot(ring_slot(rn), f);
        } else {
            bn--;
            load_slot(back_slot(bn), f);
        }
    }
    // the oldest frame: the largest pending subtree
    __device__ void pop_bottom(Frame<W> &f)
    {
        if (bn > 0) {
            load_slot(back_slot(0), f);
            bb = (bb + 1 == a.back_cap) ? 0 : bb + 1;
            bn--;
        } else {
            load_slot(ring_slot(0), f);
            rb = (rb + 1 == a.ring_cap) ? 0 : rb + 1;
            rn--;
        }
    }

    // ---- sharded MPMC frame queue ----
    __device__ __forceinline__ uint32_t home_shard() const { return (blockIdx.x * 4u + (threadIdx.x >> 6)) % NSUBQ; }

    // Reserves a free slot in some shard (home shard first).  Returns the
    // global slot index, or ~0 when every shard is full.
    __device__ size_t q_reserve(uint32_t &shard_out, uint64_t &ticket_out)
    {
        uint32_t ok = 0, shard = 0;
        uint64_t t = 0;
        if (ln == 0) {
            const uint32_t home = home_shard();
            for (uint32_t p = 0; p < NSUBQ && !ok; p++) {
                const uint32_t j = (home + p) % NSUBQ;
                const uint32_t s = atomicAdd(&a.subq[j].size, 1u);
                if (s >= a.qcap) {
                    atomicSub(&a.subq[j].size, 1u);
                } else {
                    ok = 1;
                    shard = j;
                    t = atomicAdd((unsigned long long *)&a.subq[j].tail, 1ull);
                }
            }
        }
        if (!rfl(ok)) return ~size_t(0);
        shard = rfl(shard);
        t = rfl64(t);
        const size_t slot = size_t(shard) * a.qcap + uint32_t(t % a.qcap);
        uint32_t bad = 0;
        if (ln == 0) {
            uint32_t spins = 0;
            while (__hip_atomic_load(&a.qseq[slot], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != t) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > SPIN_LIMIT) { bad = 1; break; }
            }
        }
        if (rfl(bad)) {
            set_error(KP_ETIMEDOUT);
            return ~size_t(0);
        }
        shard_out = shard;
        ticket_out = t;
        return slot;
    }
    __device__ void q_publish(size_t slot, uint32_t shard, uint64_t t, const Frame<W> &f)
    {
        store_slot(a.qframes + slot * FW, make_meta(f.root, f.type, f.r), f.P, f.C, f.X, f.D, f.type == FT_KSEARCH);
        if (ln == 0) {
            __hip_atomic_store(&a.qseq[slot], t + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&a.subq[shard].count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&S->q_count, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    // The caller has already counted itself in S->active.
    __device__ bool q_pop(Frame<W> &f)
    {
        uint32_t ok = 0, shard = 0;
        uint64_t hd = 0;
        if (ln == 0) {
            const uint32_t home = home_shard();
            for (uint32_t p = 0; p < NSUBQ && !ok; p++) {
                const uint32_t j = (home + p) % NSUBQ;
                uint32_t c = __hip_atomic_load(&a.subq[j].count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                while (c > 0) {
                    if (__hip_atomic_compare_exchange_strong(&a.subq[j].count, &c, c - 1, __ATOMIC_ACQUIRE,
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        ok = 1;
                        shard = j;
                        break;
                    }
                }
            }
            if (ok) hd = atomicAdd((unsigned long long *)&a.subq[shard].head, 1ull);
        }
        if (!rfl(ok)) return false;
        shard = rfl(shard);
        hd = rfl64(hd);
        const size_t slot = size_t(shard) * a.qcap + uint32_t(hd % a.qcap);
        uint32_t bad = 0;
        if (ln == 0) {
            uint32_t spins = 0;
            while (__hip_atomic_load(&a.qseq[slot], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != hd + 1) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > SPIN_LIMIT) { bad = 1; break; }
            }
        }
        if (rfl(bad)) {
            set_error(KP_ETIMEDOUT);
            return false;
        }
        load_slot(a.qframes + slot * FW, f);
        if (ln == 0) {
            __hip_atomic_store(&a.qseq[slot], hd + a.qcap, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            atomicSub(&a.subq[shard].size, 1u);
            __hip_atomic_fetch_sub(&S->q_count, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
        }
        pops++;
        return true;
    }

    // ---- root cache: this lane's rows of the current root ----
    __device__ void load_root(uint32_t ridx)
    {
        const RootDesc *d = a.desc + ridx;
        n = rfl(d->n);
        h = rfl(d->h);
        l = rfl(d->l);
        const uint64_t off = rfl64(d->off);
        const uint64_t *A = a.arena + off;
        const uint64_t *C2 = A + size_t(n) * W;
        const uint64_t *C4 = C2 + size_t(n) * W;
        LA = C4 + size_t(n) * W;
        const uint32_t *og = reinterpret_cast<const uint32_t *>(LA + size_t(l) * W);
#pragma unroll
        for (int j = 0; j < W; j++) {
            const uint32_t v = 64 * j + ln;
            const bool in = v < n;
#pragma unroll
            for (int w = 0; w < W; w++) {
                adjR[j][w] = in ? A[size_t(v) * W + w] : 0ull;
                c2R[j][w] = in ? C2[size_t(v) * W + w] : 0ull;
                c4R[j][w] = in ? C4[size_t(v) * W + w] : 0ull;
            }
            origR[j] = in ? og[v] : 0u;
        }
#pragma unroll
        for (int t = 0; t < LW; t++) {
            const uint32_t u = 64 * t + ln;
#pragma unroll
            for (int w = 0; w < W; w++) lftR[t][w] = u < l ? LA[size_t(u) * W + w] : 0ull;
        }
        cur = ridx;
    }

    // row of vertex v (uniform) out of the owner lane's registers
    __device__ __forceinline__ void row_of(const uint64_t (&R)[W][W], uint32_t v, uint64_t *out) const
    {
        const uint32_t pw = v >> 6, pl = v & 63;
#pragma unroll
        for (int j = 0; j < W; j++)
            if (uint32_t(j) == pw) {
#pragma unroll
                for (int w = 0; w < W; w++) out[w] = rdl64<W>(R[j][w], pl);
            }
    }

    // ---- per-vertex helpers ----
    __device__ __forceinline__ uint32_t deg_in(int j, const uint64_t *Sx) const
    {
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < W; w++) c += __popcll(adjR[j][w] & Sx[w]);
        return c;
    }
    __device__ __forceinline__ bool has(const uint64_t *Sx, int j) const { return (Sx[j] >> ln) & 1; }
    __device__ __forceinline__ bool subset_of_adj(int j, const uint64_t *Sx) const
    {
        uint64_t bad = 0;
#pragma unroll
        for (int w = 0; w < W; w++) bad |= Sx[w] & ~adjR[j][w];
        return bad == 0;
    }
    static __device__ __forceinline__ uint32_t popc(const uint64_t *Sx)
    {
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < W; w++) c += __popcll(Sx[w]);
        return c;
    }

    // upperBound (:1002-1040): every P vertex keeps >= q-k neighbours in Sx
    __device__ __forceinline__ bool ub_ok(const uint64_t *P, const uint64_t *Sx) const
    {
        bool fail = false;
#pragma unroll
        for (int j = 0; j < W; j++)
            if (has(P, j) && int(deg_in(j, Sx)) + k < q) fail = true;
        return __ballot(fail) == 0;
    }

    // kplex-compatibility mask w.r.t. P (isKplex2/isKplex3 :568-598)
    __device__ __forceinline__ void compat_mask(const uint64_t *P, uint32_t ps, uint64_t *OK) const
    {
        uint32_t dP[W];
        uint64_t Sat[W];
#pragma unroll
        for (int j = 0; j < W; j++) {
            dP[j] = deg_in(j, P);
            Sat[j] = __ballot(has(P, j) && int(dP[j]) + k == int(ps));
        }
#pragma unroll
        for (int j = 0; j < W; j++) OK[j] = __ballot(int(dP[j]) + k >= int(ps) + 1 && subset_of_adj(j, Sat));
    }

    // isMaximal_opt / isMaximalPC_opt (:703-1000): no left vertex u with
    // |N(u) ∩ S| + k >= |S| + 1 that is adjacent to every saturated vertex.
    __device__ bool maximal(const uint64_t *Sx, uint32_t s) const
    {
        uint64_t Sat[W];
#pragma unroll
        for (int j = 0; j < W; j++) Sat[j] = __ballot(has(Sx, j) && int(deg_in(j, Sx)) + k < int(s) + 1);
#pragma unroll
        for (int t = 0; t < LW; t++) {
            if (64u * t >= l) return true;
            int c = 0;
            uint64_t bad = 0;
#pragma unroll
            for (int w = 0; w < W; w++) {
                c += __popcll(lftR[t][w] & Sx[w]);
                bad |= Sat[w] & ~lftR[t][w];
            }
            if (__ballot(64u * t + ln < l && bad == 0 && c + k >= int(s) + 1)) return false;
        }
        for (uint32_t t = 64u * LW; t < l; t += 64) {
            const uint32_t u = t + ln;
            bool ext = false;
            if (u < l) {
                int c = 0;
                uint64_t bad = 0;
#pragma unroll
                for (int w = 0; w < W; w++) {
                    const uint64_t x = LA[size_t(u) * W + w];
                    c += __popcll(x & Sx[w]);
                    bad |= Sat[w] & ~x;
                }
                ext = bad == 0 && c + k >= int(s) + 1;
            }
            if (__ballot(ext)) return false;
        }
        return true;
    }

    __device__ void record(const uint64_t *Sx, uint32_t s)
    {
        uint64_t sa = 0, sb = 0;
#pragma unroll
        for (int j = 0; j < W; j++)
            if (has(Sx, j)) {
                sa += mix64(uint64_t(origR[j]) ^ kp::HASH_SA);
                sb += mix64(uint64_t(origR[j]) ^ kp::HASH_SB);
            }
        sa = rfl64(wsum64(sa));
        sb = rfl64(wsum64(sb));
        cnt++;
        ha += mix64(sa + uint64_t(s) * kp::HASH_LA);
        hb += mix64(sb + uint64_t(s) * kp::HASH_LB);
        if (a.list) {
            uint64_t base = 0;
            if (ln == 0) base = atomicAdd((unsigned long long *)a.list_used, (unsigned long long)(s + 1));
            base = rfl64(base);
            if (base + s + 1 > a.list_cap) {
                set_error(KP_EOVERFLOW);
                return;
            }
            if (ln == 0) a.list[base] = s;
            uint32_t o = uint32_t(base) + 1;
#pragma unroll
            for (int j = 0; j < W; j++) {
                if (has(Sx, j)) a.list[o + __popcll(Sx[j] & lt_mask())] = origR[j];
                o += __popcll(Sx[j]);
            }
        }
    }

    __device__ __forceinline__ void count_task() { tasks++; alg += 2ull * (5ull * n + 16ull); }

    // ---- Branch frame (BNB :1516-1724) ----
    // The exclude child goes on the stack; the include child, when it exists,
    // is returned in nf and processed next without a trip through the stack.
    __device__ bool branch(const Frame<W> &f, Frame<W> &nf)
    {
        const uint64_t *P = f.P, *C = f.C, *X = f.X;
        const uint32_t ps = popc(P), cs = popc(C);
        if (int(ps + cs) < q) return false;
        if (cs == 0) {
            if (popc(X) == 0 && int(ps) >= q && maximal(P, ps)) record(P, ps);
            return false;
        }
        uint64_t PC[W];
#pragma unroll
        for (int w = 0; w < W; w++) PC[w] = P[w] | C[w];
        uint32_t dG[W], dP[W];
#pragma unroll
        for (int j = 0; j < W; j++) {
            dG[j] = deg_in(j, PC);
            dP[j] = deg_in(j, P);
        }
        uint32_t key = ~0u;
#pragma unroll
        for (int j = 0; j < W; j++)
            if (has(P, j)) key = min(key, (dG[j] << 10) | (64u * j + ln));
        key = rfl(wmin32(key));
        const int minP = int(key >> 10);
        const uint32_t pp = key & 1023u;
        if (minP + k < max(q, int(ps))) return false;
        uint32_t pivot;
        if (minP + k < int(ps + cs)) {
            uint64_t adjPP[W];
            row_of(adjR, pp, adjPP);
            uint32_t k2 = ~0u;
#pragma unroll
            for (int j = 0; j < W; j++)
                if (((C[j] & ~adjPP[j]) >> ln) & 1) k2 = min(k2, (dG[j] << 20) | (dP[j] << 10) | (64u * j + ln));
            k2 = rfl(wmin32(k2));
            if (k2 == ~0u) {   // unreachable while the degrees are consistent
                set_error(KP_EINVAL);
                return false;
            }
            pivot = k2 & 1023u;
        } else {
            uint32_t k2 = ~0u;
#pragma unroll
            for (int j = 0; j < W; j++)
                if (has(C, j)) k2 = min(k2, (dG[j] << 20) | (dP[j] << 10) | (64u * j + ln));
            k2 = rfl(wmin32(k2));
            const int mn = min(minP, int(k2 >> 20));
            if (mn >= int(ps + cs) - k) {
                // P ∪ C is a k-plex (isKplexPC2 :600-622 over X, then maximality)
                const uint32_t tot = ps + cs;
                uint64_t Sat[W];
#pragma unroll
                for (int j = 0; j < W; j++) Sat[j] = __ballot(has(PC, j) && int(dG[j]) + k == int(tot));
                bool ext = false;
#pragma unroll
                for (int j = 0; j < W; j++)
                    if (has(X, j) && int(dG[j]) + k >= int(tot) + 1 && subset_of_adj(j, Sat)) ext = true;
                if (__ballot(ext)) return false;
                if (maximal(PC, tot)) record(PC, tot);
                return false;
            }
            pivot = k2 & 1023u;
        }
        // branchInCand2 (:1467-1472): exclude child, then include child
        const uint32_t pw = pivot >> 6;
        const uint64_t pb = 1ull << (pivot & 63);
        uint64_t C0[W], XE[W];
#pragma unroll
        for (int w = 0; w < W; w++) {
            C0[w] = C[w] & ~(w == int(pw) ? pb : 0ull);
            XE[w] = X[w] | (w == int(pw) ? pb : 0ull);
        }
        count_task();
        const uint64_t meta = make_meta(f.root, FT_BRANCH, 0);
        if (!push(meta, P, C0, XE, nullptr, false)) return false;
        // include child (:1273-1465)
        uint64_t P2[W], OK[W], c2v[W], c4v[W];
#pragma unroll
        for (int w = 0; w < W; w++) P2[w] = P[w] | (w == int(pw) ? pb : 0ull);
        compat_mask(P2, ps + 1, OK);
        row_of(c2R, pivot, c2v);
        uint64_t C2[W], PC2[W];
#pragma unroll
        for (int w = 0; w < W; w++) {
            C2[w] = C0[w] & c2v[w] & OK[w];
            PC2[w] = P2[w] | C2[w];
        }
        if (!ub_ok(P2, PC2)) return false;
        row_of(c4R, pivot, c4v);
#pragma unroll
        for (int w = 0; w < W; w++) {
            nf.P[w] = P2[w];
            nf.C[w] = C2[w];
            nf.X[w] = X[w] & c4v[w] & OK[w];
            nf.D[w] = 0;
        }
        nf.root = f.root;
        nf.type = FT_BRANCH;
        nf.r = 0;
        count_task();
        return true;
    }

    // emit a Branch task from the k-search (state-0 emit, :2005-2082) as the next frame
    __device__ __forceinline__ bool emit0(uint32_t root, const uint64_t *P, const uint64_t *C, const uint64_t *X,
                                          Frame<W> &nf)
    {
        const uint32_t ps = popc(P), cs = popc(C);
        if (int(ps + cs) < q) return false;
        if (ps > 1) {
            uint64_t PC[W];
#pragma unroll
            for (int w = 0; w < W; w++) PC[w] = P[w] | C[w];
            if (!ub_ok(P, PC)) return false;
        }
        count_task();
#pragma unroll
        for (int w = 0; w < W; w++) {
            nf.P[w] = P[w];
            nf.C[w] = C[w];
            nf.X[w] = X[w];
            nf.D[w] = 0;
        }
        nf.root = root;
        nf.type = FT_BRANCH;
        nf.r = 0;
        return true;
    }

    // ---- k-search frame: enumerate S ⊆ C2 with |S| <= budget (:1855-2444) ----
    // The exclude child goes on the stack; the include child is the next frame.
    __device__ bool ksearch(const Frame<W> &f, Frame<W> &nf)
    {
        int top = -1;
#pragma unroll
        for (int w = W - 1; w >= 0; w--)
            if (top < 0 && f.D[w]) top = 64 * w + 63 - __clzll(f.D[w]);
        if (top < 0) return emit0(f.root, f.P, f.C, f.X, nf);
        if (f.r == 0) {   // no budget: every C2 vertex is excluded
            uint64_t XD[W];
#pragma unroll
            for (int w = 0; w < W; w++) XD[w] = f.X[w] | f.D[w];
            return emit0(f.root, f.P, f.C, XD, nf);
        }
        const uint32_t u = uint32_t(top);
        const uint32_t uw = u >> 6;
        const uint64_t ub = 1ull << (u & 63);
        uint64_t D1[W], X1[W];
#pragma unroll
        for (int w = 0; w < W; w++) {
            D1[w] = f.D[w] & ~(w == int(uw) ? ub : 0ull);
            X1[w] = f.X[w] | (w == int(uw) ? ub : 0ull);
        }
        if (!push(make_meta(f.root, FT_KSEARCH, f.r), f.P, f.C, X1, D1, true)) return false;
        uint64_t c2u[W];
        row_of(c2R, u, c2u);
#pragma unroll
        for (int w = 0; w < W; w++) {
            nf.P[w] = f.P[w] | (w == int(uw) ? ub : 0ull);
            nf.C[w] = f.C[w] & c2u[w];      // updateCand13
            nf.D[w] = D1[w] & c2u[w];       // updateCand23
            nf.X[w] = f.X[w];
        }
        nf.root = f.root;
        if (f.r == 1) {
            // state-4 emit (:2255-2393): bounds, then C1 filtered by isKplex2
            const uint32_t ps = popc(nf.P), cs = popc(nf.C);
            if (int(ps + cs) < q) return false;
            if (ps > 1) {
                uint64_t PC[W];
#pragma unroll
                for (int w = 0; w < W; w++) PC[w] = nf.P[w] | nf.C[w];
                if (!ub_ok(nf.P, PC)) return false;
            }
            uint64_t OK[W];
            compat_mask(nf.P, ps, OK);
#pragma unroll
            for (int w = 0; w < W; w++) {
                nf.C[w] &= OK[w];
                nf.D[w] = 0;
            }
            nf.type = FT_BRANCH;
            nf.r = 0;
            count_task();
        } else {
            nf.type = FT_KSEARCH;
            nf.r = f.r - 1;
        }
        return true;
    }

    __device__ void push_root(uint32_t ridx)
    {
        const RootDesc *d = a.desc + ridx;
        const uint32_t rn_ = rfl(d->n), rh = rfl(d->h);
        uint64_t P[W], C[W], D[W], X[W];
#pragma unroll
        for (int w = 0; w < W; w++) {
            const int lo = 64 * w;
            auto range = [&](int b, int e) -> uint64_t {   // bits [b, e) of word w
                const int s = max(b - lo, 0), t = min(e - lo, 64);
                if (t <= s) return 0ull;
                const uint64_t hi = t == 64 ? ~0ull : ((1ull << t) - 1ull);
                return hi & ~((1ull << s) - 1ull);
            };
            P[w] = range(0, 1);
            C[w] = range(1, int(rh));
            D[w] = range(int(rh), int(rn_));
            X[w] = 0;
        }
        push(make_meta(ridx, FT_KSEARCH, uint32_t(k - 1)), P, C, X, D, true);
    }

    // Termination: a wave counts itself in S->active only while it holds
    // frames, and it increments S->active before taking a frame or a root
    // (seq_cst), so "no roots left, queue empty, nobody active", read in that
    // order, means no frame exists anywhere.  A wave that reads it exits; a
    // wave that still holds work never exits before its stack is empty.
    // Waves without work count themselves in S->hungry; busy waves donate
    // their oldest frames (the largest pending subtrees) while hungry waves
    // outnumber the queued frames.
    __device__ void run()
    {
        bool active_self = false, hungry_self = false, have = false;
        uint32_t since_donate = 0, since_check = 0, backoff = 1;
        uint64_t t_mark = t_start, busy_t = 0, idle_t = 0;
        const uint32_t gw = blockIdx.x * 4u + (threadIdx.x >> 6);
        Frame<W> f, nf;
        for (;;) {
            if (!have) {
                if (depth() == 0) {
                    if (active_self) {
                        const uint64_t now = wall_clock64();
                        busy_t += now - t_mark;
                        t_mark = now;
                        if (ln == 0) __hip_atomic_fetch_sub(&S->active, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                        active_self = false;
                    }
                    if (err) break;
                    bool got = false;
                    uint32_t qc = 0, nr = 0;
                    if (ln == 0) {
                        qc = __hip_atomic_load(&S->q_count, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                        nr = __hip_atomic_load(&S->next_root, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    qc = rfl(qc);
                    nr = rfl(nr);
                    if (qc > 0) {
                        if (ln == 0) __hip_atomic_fetch_add(&S->active, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                        got = q_pop(f);
                        if (got) {
                            active_self = true;
                            have = true;
                        } else if (ln == 0) {
                            __hip_atomic_fetch_sub(&S->active, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                        }
                        if (err) break;
                    }
                    if (!got && nr < a.ndesc) {
                        uint32_t rr = 0xffffffffu;
                        if (ln == 0) {
                            __hip_atomic_fetch_add(&S->active, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                            rr = __hip_atomic_fetch_add(&S->next_root, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                            if (rr >= a.ndesc)
                                __hip_atomic_fetch_sub(&S->active, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                        }
                        rr = rfl(rr);
                        if (rr < a.ndesc) {
                            active_self = true;
                            push_root(rr);
                            got = true;
                        }
                    }
                    if (!got) {
                        if (!hungry_self) {
                            if (ln == 0) atomicAdd(&S->hungry, 1u);
                            hungry_self = true;
                        }
                        uint32_t done = 0, stop = 0;
                        if (ln == 0) {
                            const uint32_t nr2 = __hip_atomic_load(&S->next_root, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                            const uint32_t qc2 = __hip_atomic_load(&S->q_count, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                            const uint32_t ac = __hip_atomic_load(&S->active, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                            done = nr2 >= a.ndesc && qc2 == 0 && ac == 0;
                            stop = __hip_atomic_load(&S->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
                            if (past_deadline()) stop = 2;
                        }
                        done = rfl(done);
                        stop = rfl(stop);
                        if (stop == 2) set_error(KP_ETIMEDOUT);
                        if (done || stop) break;
                        // jittered exponential backoff: waves do not poll in lockstep
                        const uint32_t jit = (gw * 2654435761u + backoff * 97u) >> 28;
                        for (uint32_t z = 0; z < backoff + jit; z++) __builtin_amdgcn_s_sleep(16);
                        if (backoff < 16) backoff <<= 1;
                        continue;
                    }
                    if (hungry_self) {
                        if (ln == 0) atomicSub(&S->hungry, 1u);
                        hungry_self = false;
                    }
                    backoff = 1;
                    const uint64_t now = wall_clock64();
                    idle_t += now - t_mark;
                    t_mark = now;
                }
                if (!have) pop_top(f);
            }
            frames++;
            if (f.root != cur) load_root(f.root);
            have = (f.type == FT_KSEARCH) ? ksearch(f, nf) : branch(f, nf);
            if (err) break;
            if (have) f = nf;
            if (++since_donate >= a.donate_period) {
                since_donate = 0;
                const uint32_t avail = depth() + (have ? 1u : 0u);
                if (avail >= 2 && depth() >= 1) {
                    uint32_t want = 0;
                    if (ln == 0) {
                        const uint32_t hg = __hip_atomic_load(&S->hungry, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        const uint32_t qc = __hip_atomic_load(&S->q_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        want = hg > qc ? min(hg - qc, 4u) : 0u;
                    }
                    want = rfl(want);
                    for (uint32_t z = 0; z < want && depth() >= 1 && depth() + (have ? 1u : 0u) >= 2; z++) {
                        uint32_t shard;
                        uint64_t t;
                        const size_t slot = q_reserve(shard, t);
                        if (slot == ~size_t(0)) break;   // queue full (or error): keep the frames
                        Frame<W> b;
                        pop_bottom(b);
                        q_publish(slot, shard, t, b);
                        if (ln == 0) atomicAdd((unsigned long long *)&S->donated, 1ull);
                    }
                    if (err) break;
                }
            }
            if (++since_check >= 1024) {
                since_check = 0;
                uint32_t stop = 0;
                if (ln == 0) {
                    stop = __hip_atomic_load(&S->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
                    if (past_deadline()) stop = 2;
                }
                stop = rfl(stop);
                if (stop == 2) set_error(KP_ETIMEDOUT);
                if (stop) break;
            }
        }
        if (active_self && ln == 0) __hip_atomic_fetch_sub(&S->active, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
        if (hungry_self && ln == 0) atomicSub(&S->hungry, 1u);
        {
            const uint64_t now = wall_clock64();
            if (active_self) busy_t += now - t_mark;
            else idle_t += now - t_mark;
        }
        if (ln == 0) {
            atomicAdd((unsigned long long *)&S->busy_ticks, (unsigned long long)busy_t);
            atomicAdd((unsigned long long *)&S->idle_ticks, (unsigned long long)idle_t);
            atomicAdd((unsigned long long *)&S->count, (unsigned long long)cnt);
            atomicAdd((unsigned long long *)&S->ha, (unsigned long long)ha);
            atomicAdd((unsigned long long *)&S->hb, (unsigned long long)hb);
            atomicAdd((unsigned long long *)&S->tasks, (unsigned long long)tasks);
            atomicAdd((unsigned long long *)&S->alg, (unsigned long long)alg);
            atomicAdd((unsigned long long *)&S->frames, (unsigned long long)frames);
            atomicAdd((unsigned long long *)&S->spills, (unsigned long long)spills);
            atomicAdd((unsigned long long *)&S->pops, (unsigned long long)pops);
            atomicAdd(&S->exited, 1u);
        }
    }
};

template <int W, int LW>
__global__ __launch_bounds__(256) void search_kernel(SearchArgs a)
{
    extern __shared__ uint64_t lds[];
    constexpr int FW = FrameLayout<W>::FW;
    const int wv = threadIdx.x >> 6;
    const uint32_t gw = blockIdx.x * (blockDim.x >> 6) + wv;
    if (gw >= a.nwaves) return;
    uint64_t *ring = lds + size_t(wv) * a.ring_cap * FW;
    uint64_t *back = a.gstack ? a.gstack + size_t(gw) * a.back_cap * FW : nullptr;
    Worker<W, LW> wk(a, ring, back);
    wk.run();
}

// slot j of every shard starts with sequence number j (free for ticket j)
__global__ void iota_kernel(uint64_t *seq, uint32_t n, uint32_t per_shard)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) seq[i] = i % per_shard;
}

} // namespace kpd

// ======================================================================
// host driver
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

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need)
    {
        if (need <= bytes && p) return KP_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (need == 0) need = 8;
        if (hipMalloc(&p, need) != hipSuccess) {
            kp::set_error("hipMalloc(%zu) failed", need);
            p = nullptr;
            return KP_ENOMEM;
        }
        bytes = need;
        return KP_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

constexpr int kClassW[] = {1, 2, 3, 4, 6, 8, 16};
constexpr int kNumClass = 7;

int class_of(uint32_t n)
{
    const uint32_t w = (n + 63) / 64;
    for (int c = 0; c < kNumClass; c++)
        if (uint32_t(kClassW[c]) >= w) return c;
    return -1;
}

} // namespace

struct kp_prepared {
    kp_params p{};
    kp::Prep prep;
    uint32_t pn = 0;
    uint64_t pm = 0;
    uint32_t maxdeg = 0, d1max = 0;
    double preprocess_ms = 0, upload_ms = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    int cus = 0;
    // graph in HBM
    DevBuf off, nbr, deg, dpos, dseq, orig;
    // per-run buffers (reused across runs)
    DevBuf blk, lft, info, ws_cnt, ws_b1c, ws_a1, desc, arena, wr_extra, list, list_used;
    // one scheduler / queue / backing stack per size class (their kernels run concurrently)
    DevBuf cq[kNumClass], cseq[kNumClass], csched[kNumClass], cstack[kNumClass], csubq[kNumClass];
    hipStream_t cstream[kNumClass] = {};
    uint32_t lcap = 0;
    ~kp_prepared()
    {
        DevBuf *all[] = {&off, &nbr, &deg, &dpos, &dseq, &orig, &blk, &lft, &info, &ws_cnt, &ws_b1c, &ws_a1,
                         &desc, &arena, &wr_extra, &list, &list_used};
        for (DevBuf *b : all) b->release();
        for (int c = 0; c < kNumClass; c++) {
            cq[c].release();
            cseq[c].release();
            csched[c].release();
            cstack[c].release();
            csubq[c].release();
            if (cstream[c]) (void)hipStreamDestroy(cstream[c]);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }
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

uint64_t dev_bytes(const kp_prepared *h)
{
    if (!h) return 0;
    const DevBuf *all[] = {&h->off, &h->nbr, &h->deg, &h->dpos, &h->dseq, &h->orig, &h->blk, &h->lft, &h->info,
                           &h->ws_cnt, &h->ws_b1c, &h->ws_a1, &h->desc, &h->arena, &h->wr_extra, &h->list, &h->list_used};
    uint64_t s = 0;
    for (const DevBuf *b : all) s += b->bytes;
    for (int c = 0; c < kNumClass; c++)
        s += h->cq[c].bytes + h->cseq[c].bytes + h->csched[c].bytes + h->cstack[c].bytes + h->csubq[c].bytes;
    return s;
}

void dev_free(kp_prepared *h) { delete h; }

int dev_prepare(const kp_graph *g, const kp_params *p, kp_prepared **out)
{
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        set_error("no HIP device visible: the HIP backend is required");
        return KP_ENODEV;
    }
    auto *h = new (std::nothrow) kp_prepared();
    if (!h) return KP_ENOMEM;
    h->p = *p;
    if (hipGetDevice(&h->device) != hipSuccess) { delete h; return KP_EHIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) != hipSuccess) { delete h; return KP_EHIP; }
    h->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) { delete h; return KP_EHIP; }
    for (int c = 0; c < kNumClass; c++)
        if (hipStreamCreateWithFlags(&h->cstream[c], hipStreamNonBlocking) != hipSuccess) { delete h; return KP_EHIP; }

    const double t0 = now_ms();
    int rc = preprocess(g, p->k, p->q, p->truss, &h->prep);
    if (rc) { delete h; return rc; }
    h->preprocess_ms = now_ms() - t0;
    const kp_graph &G = h->prep.g;
    h->pn = G.n;
    h->pm = G.m;
    uint32_t maxdeg = 0, d1 = 0;
    for (uint32_t v = 0; v < G.n; v++) {
        maxdeg = std::max(maxdeg, G.deg[v]);
        uint32_t later = 0;
        for (uint64_t e = G.off[v]; e < G.off[v + 1]; e++) later += h->prep.dpos[G.nbr[e]] > h->prep.dpos[v];
        d1 = std::max(d1, later);
    }
    h->maxdeg = maxdeg;
    h->d1max = d1;

    const double t1 = now_ms();
    const size_t pn = std::max<size_t>(G.n, 1);
    if ((rc = h->off.ensure(sizeof(uint64_t) * (pn + 1))) || (rc = h->nbr.ensure(sizeof(uint32_t) * std::max<uint64_t>(G.m, 1))) ||
        (rc = h->deg.ensure(4 * pn)) || (rc = h->dpos.ensure(4 * pn)) || (rc = h->dseq.ensure(4 * pn)) ||
        (rc = h->orig.ensure(4 * pn))) {
        delete h;
        return rc;
    }
    auto up = [&](DevBuf &b, const void *src, size_t bytes) -> int {
        if (!bytes) return KP_OK;
        return hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, h->stream) == hipSuccess ? KP_OK : KP_EHIP;
    };
    if ((rc = up(h->off, G.off, sizeof(uint64_t) * (size_t(G.n) + 1))) || (rc = up(h->nbr, G.nbr, 4 * G.m)) ||
        (rc = up(h->deg, G.deg, 4 * size_t(G.n))) || (rc = up(h->dpos, h->prep.dpos.data(), 4 * size_t(G.n))) ||
        (rc = up(h->dseq, h->prep.dseq.data(), 4 * size_t(G.n))) || (rc = up(h->orig, h->prep.orig.data(), 4 * size_t(G.n)))) {
        set_error("upload to HBM failed");
        delete h;
        return rc;
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess) { delete h; return KP_EHIP; }
    h->upload_ms = now_ms() - t1;
    *out = h;
    return KP_OK;
}

namespace {

template <int WC>
int launch_build(kp_prepared *h, const BuildArgs &ba, uint32_t nwg)
{
    hipLaunchKernelGGL((build_kernel<WC>), dim3(nwg), dim3(256), 0, h->stream, ba);
    HIPCHK(hipGetLastError());
    return KP_OK;
}

struct ClassLaunch {
    int W = 0;
    uint32_t nwaves = 0, ring = 0, back = 0;
    int blocks_per_cu = 0;
};

template <int W, int LW>
int launch_search(kp_prepared *h, int c, SearchArgs sa, uint32_t maxn, ClassLaunch *cl)
{
    constexpr int FW = FrameLayout<W>::FW;
    const uint32_t need = maxn + 3;   // DFS depth bound: every child shrinks C2 or C
    const uint32_t ring = std::min<uint32_t>(need, std::max<uint32_t>(16, 8192 / (FW * 8)));
    sa.ring_cap = ring;
    sa.back_cap = need > ring ? need - ring + 1 : 0;
    const size_t dyn = size_t(ring) * FW * 8 * 4;
    auto kern = search_kernel<W, LW>;
    if (dyn > 64 * 1024) HIPCHK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(dyn)));
    int nb = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 256, dyn));
    if (nb < 1) nb = 1;
    const uint32_t nwg = uint32_t(nb) * uint32_t(h->cus);
    sa.nwaves = nwg * 4;
    sa.qcap = std::max<uint32_t>(256, sa.nwaves * 4 / NSUBQ);   // slots per shard
    sa.donate_period = 8;
    sa.donate_target = std::max<uint32_t>(16, sa.nwaves / 4);
    const size_t nslots = size_t(sa.qcap) * NSUBQ;
    int rc;
    if ((rc = h->cq[c].ensure(nslots * FW * 8)) || (rc = h->cseq[c].ensure(nslots * 8)) ||
        (rc = h->csched[c].ensure(sizeof(Sched))) || (rc = h->csubq[c].ensure(sizeof(SubQ) * NSUBQ)))
        return rc;
    if (sa.back_cap && (rc = h->cstack[c].ensure(size_t(sa.nwaves) * sa.back_cap * FW * 8))) return rc;
    sa.qframes = h->cq[c].as<uint64_t>();
    sa.qseq = h->cseq[c].as<uint64_t>();
    sa.sched = h->csched[c].as<Sched>();
    sa.subq = h->csubq[c].as<SubQ>();
    sa.gstack = sa.back_cap ? h->cstack[c].as<uint64_t>() : nullptr;
    hipStream_t st = h->cstream[c];
    HIPCHK(hipMemsetAsync(sa.sched, 0, sizeof(Sched), st));
    HIPCHK(hipMemsetAsync(sa.subq, 0, sizeof(SubQ) * NSUBQ, st));
    hipLaunchKernelGGL(iota_kernel, dim3(uint32_t((nslots + 255) / 256)), dim3(256), 0, st, sa.qseq, uint32_t(nslots),
                       sa.qcap);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(kern, dim3(nwg), dim3(256), dyn, st, sa);
    HIPCHK(hipGetLastError());
    cl->W = W;
    cl->nwaves = sa.nwaves;
    cl->ring = sa.ring_cap;
    cl->back = sa.back_cap;
    cl->blocks_per_cu = nb;
    return KP_OK;
}

int run_search_class(kp_prepared *h, int c, SearchArgs sa, uint32_t maxn, ClassLaunch *cl)
{
    switch (kClassW[c]) {
    case 1: return launch_search<1, 2>(h, c, sa, maxn, cl);
    case 2: return launch_search<2, 2>(h, c, sa, maxn, cl);
    case 3: return launch_search<3, 2>(h, c, sa, maxn, cl);
    case 4: return launch_search<4, 1>(h, c, sa, maxn, cl);
    case 6: return launch_search<6, 0>(h, c, sa, maxn, cl);
    case 8: return launch_search<8, 0>(h, c, sa, maxn, cl);
    case 16: return launch_search<16, 0>(h, c, sa, maxn, cl);
    }
    return KP_EINVAL;
}

int run_build_class(kp_prepared *h, int c, const BuildArgs &ba, uint32_t nwg)
{
    switch (kClassW[c]) {
    case 1: return launch_build<1>(h, ba, nwg);
    case 2: return launch_build<2>(h, ba, nwg);
    case 3: return launch_build<3>(h, ba, nwg);
    case 4: return launch_build<4>(h, ba, nwg);
    case 6: return launch_build<6>(h, ba, nwg);
    case 8: return launch_build<8>(h, ba, nwg);
    case 16: return launch_build<16>(h, ba, nwg);
    }
    return KP_EINVAL;
}

} // namespace

int dev_run(kp_prepared *h, kp_result *res)
{
    std::memset(res, 0, sizeof(*res));
    const kp_params &P = h->p;
    HIPCHK(hipSetDevice(h->device));
    const kp_graph &G = h->prep.g;
    res->pn = h->pn;
    res->pm = h->pm;
    res->preprocess_ms = h->preprocess_ms;
    res->upload_ms = h->upload_ms;
    // roots at degeneracy positions [0, pn - q + 2) (src/device_funcs.cu:44-45), optionally windowed
    int64_t pos_end = std::max<int64_t>(0, int64_t(h->pn) - P.q + 2);
    int64_t pos_begin = 0;
    if (P.root_end > 0 || P.root_begin > 0) {
        pos_begin = std::min<int64_t>(P.root_begin, pos_end);
        if (P.root_end > 0) pos_end = std::min<int64_t>(P.root_end, pos_end);
        pos_end = std::max(pos_end, pos_begin);
    }
    const int64_t nroots = pos_end - pos_begin;
    hipEvent_t e0, e1, e2;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventCreate(&e2));
    struct EvGuard { hipEvent_t a, b, c; ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); (void)hipEventDestroy(c); } } eg{e0, e1, e2};
    HIPCHK(hipEventRecord(e0, h->stream));
    int rc;
    std::vector<RootInfo> info(static_cast<size_t>(nroots));
    std::vector<RootDesc> descs;
    std::vector<int> dclass;
    std::vector<uint32_t> class_begin(kNumClass + 1, 0), class_maxn(kNumClass, 0);
    std::vector<uint64_t> wr_extra;
    if (nroots > 0) {
        // ---- decomposition ----
        h->lcap = std::min<uint32_t>(std::max<uint32_t>(h->pn, 1), std::max<uint32_t>(2048, 2 * h->maxdeg + 1024));
        const uint32_t d1cap = std::max<uint32_t>(h->d1max, 1);
        const uint32_t w1cap = (d1cap + 63) / 64;
        uint32_t nwaves = uint32_t(std::min<int64_t>(nroots, int64_t(h->cus) * 16));
        // keep the per-wave counter arrays under ~8 GiB
        const uint64_t cnt_bytes = uint64_t(std::max<uint32_t>(h->pn, 1)) * 4;
        nwaves = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(nwaves, (8ull << 30) / cnt_bytes)));
        const uint32_t nwg = (nwaves + WPB - 1) / WPB;
        if ((rc = h->blk.ensure(size_t(nroots) * NCAP * 4)) || (rc = h->lft.ensure(size_t(nroots) * h->lcap * 4)) ||
            (rc = h->info.ensure(size_t(nroots) * sizeof(RootInfo))) || (rc = h->ws_cnt.ensure(size_t(nwaves) * cnt_bytes)) ||
            (rc = h->ws_b1c.ensure(size_t(nwaves) * d1cap * 4)) || (rc = h->ws_a1.ensure(size_t(nwaves) * d1cap * w1cap * 8)))
            return rc;
        HIPCHK(hipMemsetAsync(h->ws_cnt.p, 0, size_t(nwaves) * cnt_bytes, h->stream));
        DecompArgs da{};
        da.off = h->off.as<uint64_t>();
        da.nbr = h->nbr.as<uint32_t>();
        da.deg = h->deg.as<uint32_t>();
        da.dpos = h->dpos.as<int32_t>();
        da.dseq = h->dseq.as<int32_t>();
        da.pn = std::max<uint32_t>(h->pn, 1);
        da.k = P.k;
        da.q = P.q;
        da.pos_begin = pos_begin;
        da.nroots = nroots;
        da.blk = h->blk.as<uint32_t>();
        da.lft = h->lft.as<uint32_t>();
        da.lcap = h->lcap;
        da.info = h->info.as<RootInfo>();
        da.ws_cnt = h->ws_cnt.as<uint32_t>();
        da.ws_b1c = h->ws_b1c.as<uint32_t>();
        da.ws_a1 = h->ws_a1.as<uint64_t>();
        da.d1cap = d1cap;
        da.w1cap = w1cap;
        da.nwaves = nwaves;
        hipLaunchKernelGGL(decompose_kernel, dim3(nwg), dim3(256), 0, h->stream, da);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(info.data(), da.info, sizeof(RootInfo) * size_t(nroots), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));

        // ---- classes and arena layout ----
        std::vector<std::vector<uint32_t>> by_class(kNumClass);
        for (int64_t pos = 0; pos < nroots; pos++) {
            const RootInfo &ri = info[size_t(pos)];
            if (ri.status < 0) {
                if (ri.status == -1) {
                    kp::set_error("root at degeneracy position %lld: block exceeds %u vertices", (long long)(pos + pos_begin), NCAP);
                    return KP_ERANGE;
                }
                kp::set_error("root at degeneracy position %lld: left set exceeds %u vertices", (long long)(pos + pos_begin), h->lcap);
                return KP_EOVERFLOW;
            }
            if (ri.status != 1) continue;
            res->valid_roots++;
            res->max_n = std::max(res->max_n, ri.n);
            res->max_l = std::max(res->max_l, ri.l);
            by_class[class_of(ri.n)].push_back(uint32_t(pos));
        }
        uint64_t arena_words = 0;
        for (int c = 0; c < kNumClass; c++) {
            auto &v = by_class[c];
            // largest blocks first (longest-processing-time order)
            std::stable_sort(v.begin(), v.end(), [&](uint32_t x, uint32_t y) { return info[x].n > info[y].n; });
            class_begin[c] = uint32_t(descs.size());
            const uint32_t W = uint32_t(kClassW[c]);
            for (uint32_t pos : v) {
                const RootInfo &ri = info[pos];
                RootDesc d{arena_words, ri.n, ri.h, ri.l, pos};
                descs.push_back(d);
                dclass.push_back(c);
                class_maxn[c] = std::max(class_maxn[c], ri.n);
                arena_words += 3ull * ri.n * W + uint64_t(ri.l) * W + (ri.n + 1) / 2;
            }
        }
        class_begin[kNumClass] = uint32_t(descs.size());
        const uint32_t nd = uint32_t(descs.size());
        if (nd) {
            if ((rc = h->desc.ensure(sizeof(RootDesc) * nd)) || (rc = h->arena.ensure(8 * arena_words)) ||
                (rc = h->wr_extra.ensure(8 * nd)))
                return rc;
            HIPCHK(hipMemcpyAsync(h->desc.p, descs.data(), sizeof(RootDesc) * nd, hipMemcpyHostToDevice, h->stream));
            HIPCHK(hipMemsetAsync(h->arena.p, 0, 8 * arena_words, h->stream));
            for (int c = 0; c < kNumClass; c++) {
                const uint32_t b = class_begin[c], e = class_begin[c + 1];
                if (b == e) continue;
                BuildArgs ba{};
                ba.off = da.off;
                ba.nbr = da.nbr;
                ba.deg = da.deg;
                ba.orig = h->orig.as<uint32_t>();
                ba.blk = da.blk;
                ba.lft = da.lft;
                ba.lcap = h->lcap;
                ba.desc = h->desc.as<RootDesc>() + b;
                ba.ndesc = e - b;
                ba.arena = h->arena.as<uint64_t>();
                ba.ws_cnt = da.ws_cnt;
                ba.pn = da.pn;
                ba.nwaves = std::min<uint32_t>(nwaves, ((e - b) + WPB - 1) / WPB * WPB);
                ba.k = P.k;
                ba.q = P.q;
                ba.wr_extra = h->wr_extra.as<uint64_t>() + b;
                if ((rc = run_build_class(h, c, ba, (ba.nwaves + WPB - 1) / WPB))) return rc;
            }
            wr_extra.resize(nd);
            HIPCHK(hipMemcpyAsync(wr_extra.data(), h->wr_extra.p, 8 * nd, hipMemcpyDeviceToHost, h->stream));
        }
    }
    HIPCHK(hipEventRecord(e1, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < descs.size(); i++) {
        const RootInfo &ri = info[descs[i].pos];
        const uint64_t wr = ri.wr + wr_extra[i];
        res->adj_bytes += 4 * wr;
        res->alg_bytes += 4 * wr + (uint64_t(ri.n) * ri.n + 7) / 8 + uint64_t(ri.n) * ri.n;
    }

    // ---- persistent search: one launch per size class, all classes concurrently ----
    const bool want_list = (P.flags & KP_FLAG_LIST) != 0;
    const uint64_t list_cap = want_list ? (64ull << 20) : 0;
    if (want_list && (rc = h->list.ensure(list_cap * 4))) return rc;
    if ((rc = h->list_used.ensure(8))) return rc;
    HIPCHK(hipMemsetAsync(h->list_used.p, 0, 8, h->stream));
    uint64_t limit_ticks = 0;
    if (P.time_limit_s > 0) {
        int khz = 0;
        HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, h->device));
        limit_ticks = uint64_t(P.time_limit_s * double(khz) * 1e3);
    }
    // every class stream starts after the build; the main stream waits for all of them
    hipEvent_t ebuild;
    HIPCHK(hipEventCreateWithFlags(&ebuild, hipEventDisableTiming));
    HIPCHK(hipEventRecord(ebuild, h->stream));
    std::vector<int> order;
    for (int c = 0; c < kNumClass; c++)
        if (class_begin[c + 1] > class_begin[c]) order.push_back(c);
    // largest classes (by sum of n^2 over their roots) first
    std::vector<double> cwork(kNumClass, 0.0);
    for (size_t i = 0; i < descs.size(); i++) cwork[dclass[i]] += double(descs[i].n) * descs[i].n;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cwork[x] > cwork[y]; });
    ClassLaunch launches[kNumClass];
    std::vector<hipEvent_t> cdone;
    const bool serial_classes = std::getenv("KP_SERIAL_CLASSES") != nullptr;
    int launch_rc = KP_OK;
    for (int c : order) {
        const uint32_t b = class_begin[c], e = class_begin[c + 1];
        HIPCHK(hipStreamWaitEvent(h->cstream[c], ebuild, 0));
        SearchArgs sa{};
        sa.desc = h->desc.as<RootDesc>() + b;
        sa.ndesc = e - b;
        sa.arena = h->arena.as<uint64_t>();
        sa.k = P.k;
        sa.q = P.q;
        sa.list = want_list ? h->list.as<uint32_t>() : nullptr;
        sa.list_cap = list_cap;
        sa.list_used = h->list_used.as<uint64_t>();
        sa.limit_ticks = limit_ticks;
        if ((launch_rc = run_search_class(h, c, sa, class_maxn[c], &launches[c]))) break;
        hipEvent_t ev;
        HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        HIPCHK(hipEventRecord(ev, h->cstream[c]));
        HIPCHK(hipStreamWaitEvent(h->stream, ev, 0));
        cdone.push_back(ev);
        if (serial_classes) HIPCHK(hipStreamSynchronize(h->cstream[c]));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int c : order) HIPCHK(hipStreamSynchronize(h->cstream[c]));
    for (hipEvent_t ev : cdone) (void)hipEventDestroy(ev);
    (void)hipEventDestroy(ebuild);
    if (launch_rc) return launch_rc;
    for (int c : order) {
        Sched hs{};
        HIPCHK(hipMemcpy(&hs, h->csched[c].p, sizeof(Sched), hipMemcpyDeviceToHost));
        if (std::getenv("KP_STATS"))
            std::fprintf(stderr,
                         "[kp] search W=%d roots=%u waves=%u (%d blocks/CU) ring=%u back=%u frames=%llu tasks=%llu "
                         "donated=%llu pops=%llu spills=%llu busy=%.1f%% wave-s=%.1f\n",
                         kClassW[c], class_begin[c + 1] - class_begin[c], launches[c].nwaves,
                         launches[c].blocks_per_cu, launches[c].ring, launches[c].back, (unsigned long long)hs.frames,
                         (unsigned long long)hs.tasks, (unsigned long long)hs.donated, (unsigned long long)hs.pops,
                         (unsigned long long)hs.spills,
                         100.0 * double(hs.busy_ticks) / double(std::max<uint64_t>(1, hs.busy_ticks + hs.idle_ticks)),
                         double(hs.busy_ticks + hs.idle_ticks) / 1e8);
        if (hs.error) {
            kp::set_error("search kernel (class W=%d) reported error %d", kClassW[c], hs.error);
            return hs.error;
        }
        res->count += hs.count;
        res->hash_a += hs.ha;
        res->hash_b += hs.hb;
        res->tasks += hs.tasks;
        res->alg_bytes += hs.alg;
        res->search_alg_bytes += hs.alg;
    }
    std::vector<uint32_t> listed;
    if (want_list) {
        uint64_t used = 0;
        HIPCHK(hipMemcpy(&used, h->list_used.p, 8, hipMemcpyDeviceToHost));
        if (used > list_cap) {
            kp::set_error("plex list exceeds %llu words", (unsigned long long)list_cap);
            return KP_EOVERFLOW;
        }
        listed.resize(used);
        if (used) HIPCHK(hipMemcpy(listed.data(), h->list.p, 4 * used, hipMemcpyDeviceToHost));
    }
    HIPCHK(hipEventRecord(e2, h->stream));
    HIPCHK(hipEventSynchronize(e2));
    float ms1 = 0, ms2 = 0;
    HIPCHK(hipEventElapsedTime(&ms1, e0, e1));
    HIPCHK(hipEventElapsedTime(&ms2, e1, e2));
    res->decompose_ms = ms1;
    res->search_ms = ms2;
    res->wall_ms = res->preprocess_ms + ms1 + ms2;
    (void)G;
    if (want_list) {
        // records are (len, ids...); sort ids within a plex, then plexes lexicographically
        std::vector<std::pair<uint64_t, uint32_t>> recs;
        for (uint64_t i = 0; i < listed.size();) {
            const uint32_t len = listed[i];
            std::sort(listed.begin() + i + 1, listed.begin() + i + 1 + len);
            recs.emplace_back(i + 1, len);
            i += 1 + len;
        }
        std::sort(recs.begin(), recs.end(), [&](const auto &x, const auto &y) {
            return std::lexicographical_compare(listed.begin() + x.first, listed.begin() + x.first + x.second,
                                                listed.begin() + y.first, listed.begin() + y.first + y.second);
        });
        res->nlisted = recs.size();
        res->list_off = static_cast<uint64_t *>(std::malloc(8 * (recs.size() + 1)));
        uint64_t tot = 0;
        for (auto &r : recs) tot += r.second;
        res->list_ids = static_cast<uint32_t *>(std::malloc(4 * std::max<uint64_t>(tot, 1)));
        if (!res->list_off || !res->list_ids) return KP_ENOMEM;
        uint64_t o = 0;
        for (size_t i = 0; i < recs.size(); i++) {
            res->list_off[i] = o;
            std::memcpy(res->list_ids + o, listed.data() + recs[i].first, 4 * recs[i].second);
            o += recs[i].second;
        }
        res->list_off[recs.size()] = o;
    }
    return KP_OK;
}

} // namespace kp
