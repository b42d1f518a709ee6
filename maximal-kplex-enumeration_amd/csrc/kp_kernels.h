// kp_kernels.h -- device code of libkplex: the decompose, build, search and
// per-root kernels and the wave-level helpers they share.  Included by
// kp_device.hip (host driver, every kernel but the one-word search kernel)
// and kp_search1.hip (the one-word search kernel, built with its own
// compiler options; that file defines KP_ONE_WORD_SEARCH_ONLY, which leaves
// out the kernels that are not templates).
//
// Work unit: a job = (preprocessed graph in HBM, k, q, window of root
// positions).  One plain run is one job; a batch or a server holds many jobs
// whose roots share one pool of size-class launches.
//   decompose_kernel   one wave64 per (job, root position), grid-stride:
//                      later/earlier split of N(r), (q-2k)-core of the later
//                      1-hop on a 64-bit bitset adjacency, left filter and
//                      2-hop candidates counted with a per-wave counter array.
//                      [reference: decompose, src/device_funcs.cu:28-298]
//   build_kernel<WC>   one wave per valid root: bitset adjacency rows, the two
//                      pair-class bitsets (class >= UNLINK2EQUAL, class >=
//                      UNLINK2MORE), left rows, original ids.   [calculateDegrees,
//                      fillNeighbors, buildCommonMtx :300-460, :1727-1853]
//   search_kernel<W>   persistent: every wave64 is a worker with a DFS stack of
//                      (P, C, C2, X) bitset frames; k-search frames [kSearch
//                      :1855-2444] and Branch frames [BNB :1516-1724] share the
//                      stack; idle waves take resumed frames, then new roots from
//                      an atomic cursor, then frames donated by busy waves
//                      through a bounded MPMC queue in HBM.  On a pause request
//                      every wave writes its pending frames to a dump array and
//                      exits; the host resumes them in a later launch.
//
// Wave layout: lane L of a wave owns the block vertices L, 64+L, ... (word j of
// every bitset holds vertices 64j..64j+63, bit L = vertex 64j+L).  Sets are
// wave-uniform (SGPRs); per-vertex degrees are per-lane popcounts of the owned
// vertices' adjacency rows, which stay in VGPRs while a wave works on a root.
#pragma once
#include <hip/hip_runtime.h>

#include "kp_engine.h"

#include <algorithm>
#include <cstdint>
#include <utility>

namespace kpd {

constexpr uint32_t kSeedChunk = 16;   // resumed frames a wave claims at once, at most

constexpr uint32_t NCAP = KP_MAX_BLOCK;          // largest block
constexpr int WPB = 4;                           // waves per block (256 threads)
constexpr uint32_t EXCL = 0x80000000u;           // counter-array flags
constexpr uint32_t CLAIM = 0x40000000u;
constexpr uint32_t CMASK = 0x3FFFFFFFu;
constexpr uint32_t SPIN_LIMIT = 1u << 24;


// ---------------------------------------------------------------- helpers
__device__ __forceinline__ uint64_t lt_mask() { return (1ull << __lane_id()) - 1ull; }
__device__ __forceinline__ uint32_t rfl(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t rfl64(uint64_t x)
{
    return (uint64_t(rfl(uint32_t(x >> 32))) << 32) | rfl(uint32_t(x));
}
// Wave64 reductions: four DPP steps (quad_perm [1,0,3,2], [2,3,0,1],
// row_half_mirror, row_mirror) leave every lane of a 16-lane row holding the
// row's result; row_bcast:15 (rows 1 and 3 take the last lane of the row
// before) and row_bcast:31 (rows 2 and 3 take lane 31) then carry it across
// rows, so lane 63 holds the wave's result, read with one v_readlane.
#define KP_DPP(v, ctrl) uint32_t(__builtin_amdgcn_update_dpp(0, int(v), ctrl, 0xF, 0xF, false))
__device__ __forceinline__ uint32_t wmin32(uint32_t v)
{
    v = min(v, KP_DPP(v, 0xB1));
    v = min(v, KP_DPP(v, 0x4E));
    v = min(v, KP_DPP(v, 0x141));
    v = min(v, KP_DPP(v, 0x140));
    // the two cross-row steps as DPP-fused mins: rows outside row_mask keep
    // their value (s_nop 1 covers the VALU-write -> DPP-read hazard)
    asm volatile("s_nop 1\n\tv_min_u32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf" : "+v"(v));
    asm volatile("s_nop 1\n\tv_min_u32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf" : "+v"(v));
    return __builtin_amdgcn_readlane(v, 63);
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

// Root r = dseq[pos] (src/device_funcs.cu:28-298), restated on bitsets:
//  B1 = (q-2k)-core of G[later neighbours]   (fixpoint :139-171; the core is
//       unique, so any peeling order gives the reference's set)
//  L1 = earlier neighbours with >= q-2k+1 neighbours in B1          (:177-209)
//  B2 = later non-neighbours with >= q-2k+2 neighbours in B1        (:211-273)
//  L2 = earlier non-neighbours with >= q-2k+3 neighbours in B1
// Canonical layout: B = [r, B1 asc, B2 asc], L = [L1 asc, L2].
// Per-wave workspace of the decomposition: the counter array over the job's
// vertices (zero between roots), the later-neighbour list and its bitset
// adjacency in HBM, three LDS scratch rows.
struct DecompWs {
    uint32_t *cnt, *b1c;
    uint64_t *A1;
    uint64_t *row, *alive;
    uint32_t *sb2;
};

// Decomposes the root at degeneracy position pos of job J into blk/lft at
// `slot` and writes info[slot]; the whole wave calls it.
__device__ void decompose_root(const DecompArgs &a, const JobDev &J, int32_t pos, uint32_t slot, const DecompWs &ws);

#ifndef KP_ONE_WORD_SEARCH_ONLY
__global__ __launch_bounds__(256) void decompose_kernel(DecompArgs a)
{
    __shared__ uint64_t s_row[WPB][64];
    __shared__ uint64_t s_alive[WPB][64];
    __shared__ uint32_t s_b2[WPB][NCAP];
    const int wv = threadIdx.x >> 6;
    const uint32_t gw = blockIdx.x * WPB + wv;
    if (gw >= a.nwaves) return;
    const DecompWs ws{a.ws_cnt + size_t(gw) * a.cnt_stride, a.ws_b1c + size_t(gw) * a.d1cap,
                      a.ws_a1 + size_t(gw) * a.d1cap * a.w1cap, s_row[wv], s_alive[wv], s_b2[wv]};
    for (uint32_t slot = gw; slot < a.nslots; slot += a.nwaves) {
        const JobDev J = a.jobs[rfl(a.slot_job[slot])];
        decompose_root(a, J, int32_t(rfl(a.slot_pos[slot])), slot, ws);
    }
}
#endif


__device__ void decompose_root(const DecompArgs &a, const JobDev &J, int32_t pos, uint32_t slot, const DecompWs &ws)
{
    const int ln = __lane_id();
    uint32_t *cnt = ws.cnt, *b1c = ws.b1c, *sb2 = ws.sb2;
    uint64_t *A1 = ws.A1, *row = ws.row, *alive = ws.alive;
    {
        const int lb2k = J.q - 2 * J.k, bd = J.q - J.k;
        const uint32_t r = uint32_t(J.dseq[pos]);
        const uint64_t ro = J.off[r];
        const uint32_t dr = J.deg[r];
        uint32_t *B = a.blk + size_t(slot) * NCAP;
        uint32_t *L = a.lft + size_t(slot) * a.lcap;
        RootInfo info{0, 0, 0, 0, dr};

        // 1) split N(r) keeping list (ascending id) order
        uint32_t m1 = 0, ml = 0;
        uint64_t wsum = 0;
        for (uint32_t base = 0; base < dr; base += 64) {
            const uint32_t idx = base + ln;
            const bool v = idx < dr;
            const uint32_t w = v ? J.nbr[ro + idx] : 0u;
            const bool later = v && J.dpos[w] > pos;
            const bool earlier = v && !later;
            const uint64_t mL = __ballot(later), mE = __ballot(earlier);
            if (later) {
                const uint32_t t = m1 + __popcll(mL & lt_mask());
                if (t < a.d1cap) b1c[t] = w;
                wsum += J.deg[w];
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
            return;
        }
        if (m1 > a.d1cap || m1 > 64u * 64u) {
            info.status = -1;
            if (ln == 0) a.info[slot] = info;
            return;
        }
        const uint32_t w1 = (m1 + 63) >> 6;

        // 2) bitset adjacency among the later neighbours (index map in cnt[])
        for (uint32_t t = ln; t < m1; t += 64) ast(&cnt[b1c[t]], t + 1);
        fence_agent();
        for (uint32_t i = 0; i < m1; i++) {
            const uint32_t u = b1c[i];
            const uint32_t du = J.deg[u];
            if (int(du) < lb2k) continue;          // cannot be in the core: row never read
            const uint64_t uo = J.off[u];
            for (uint32_t t = ln; t < w1; t += 64) row[t] = 0;
            fence_wg();
            for (uint32_t base = 0; base < du; base += 64) {
                const uint32_t idx = base + ln;
                if (idx < du) {
                    const uint32_t c = ald(&cnt[J.nbr[uo + idx]]);
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
            const bool al = i < m1 && int(J.deg[b1c[i]]) >= lb2k;
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
            return;
        }
        const uint32_t hop = nb1 + 1;
        if (hop > NCAP) {
            info.status = -1;
            if (ln == 0) a.info[slot] = info;
            return;
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
        for (uint32_t t = ln; t < dr; t += 64) aor(&cnt[J.nbr[ro + t]], EXCL);
        for (uint32_t b = 1; b < hop; b++) {
            const uint32_t u = B[b];
            const uint64_t uo = J.off[u];
            const uint32_t du = J.deg[u];
            for (uint32_t t = ln; t < du; t += 64) aadd(&cnt[J.nbr[uo + t]], 1);
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
            const uint64_t uo = J.off[u];
            const uint32_t du = J.deg[u];
            for (uint32_t base = 0; base < du; base += 64) {
                const uint32_t idx = base + ln;
                bool own = false, later = false;
                uint32_t w = 0;
                if (idx < du) {
                    w = J.nbr[uo + idx];
                    const uint32_t c = ald(&cnt[w]);
                    if (!(c & (EXCL | CLAIM))) {
                        later = J.dpos[w] > pos;
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
        for (uint32_t t = ln; t < dr; t += 64) ast(&cnt[J.nbr[ro + t]], 0);
        for (uint32_t b = 1; b < hop; b++) {
            const uint32_t u = B[b];
            const uint64_t uo = J.off[u];
            const uint32_t du = J.deg[u];
            for (uint32_t t = ln; t < du; t += 64) ast(&cnt[J.nbr[uo + t]], 0);
        }
        fence_agent();
        const uint32_t n = hop + nb2, l = nl1 + nl2;
        info.n = n;
        info.h = hop;
        info.l = l;
        if (n > NCAP) info.status = -1;
        else if (ml > a.lcap || l > a.lcap) info.status = -2;
        else info.status = int(n) >= J.q ? 1 : 0;
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

// Per root: adjacency rows adj[i] (bit j = B[j] adjacent to B[i]), class rows
// cls2[i] (bit j: class(i,j) >= UNLINK2EQUAL) and cls4[i] (>= UNLINK2MORE)
// with the thresholds of buildCommonMtx (src/device_funcs.cu:1768-1773) over
// common = |N(i) ∩ N(j) ∩ B[0,h)|; left rows ladj[u] (bit i = L[u] adjacent
// to B[i]); original ids.  Row stride is the class width WC.
template <int WC>
__device__ void build_root(const BuildArgs &a, const RootDesc &d, const JobDev &J, uint32_t di, uint32_t *cnt,
                           uint64_t *row);

template <int WC>
__global__ __launch_bounds__(256) void build_kernel(BuildArgs a)
{
    __shared__ uint64_t s_row[WPB][WC];
    const int wv = threadIdx.x >> 6;
    const uint32_t gw = blockIdx.x * WPB + wv;
    if (gw >= a.nwaves) return;
    uint32_t *cnt = a.ws_cnt + size_t(gw) * a.cnt_stride;
    for (uint32_t di = gw; di < a.ndesc; di += a.nwaves) {
        const RootDesc d = a.desc[di];
        build_root<WC>(a, d, a.jobs[d.job], di, cnt, s_row[wv]);
    }
}

// Writes root d's rows into its arena (d.addr) and a.wr_extra[di]; the whole
// wave calls it, with a zeroed counter array over the job's vertices.
template <int WC>
__device__ void build_root(const BuildArgs &a, const RootDesc &d, const JobDev &J, uint32_t di, uint32_t *cnt,
                           uint64_t *row)
{
    const int ln = __lane_id();
    {
        const int k = d.k, q = d.q;
        const int km2 = max(k - 2, 0), km3 = max(k - 3, 0);
        const int thPP1 = q - k - 2 * km2, thPP2 = q - k - 2 * km3;
        const int thPC1 = q - 2 * k - km2, thPC2 = q - k - 1 - km2 - km3;
        const int thCC1 = q - 2 * k - (k - 1), thCC2 = q - 2 * k + 2 - (k - 1);
        const uint32_t n = d.n, h = d.h, l = d.l;
        const uint32_t *B = a.blk + size_t(d.slot) * NCAP;
        const uint32_t *L = a.lft + size_t(d.slot) * a.lcap;
        uint64_t *adj = reinterpret_cast<uint64_t *>(d.addr);
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
            const uint64_t uo = J.off[u];
            const uint32_t du = J.deg[u];
            if (ln < WC) row[ln] = 0;
            fence_wg();
            for (uint32_t base = 0; base < du; base += 64) {
                const uint32_t idx = base + ln;
                if (idx < du) {
                    const uint32_t c = ald(&cnt[J.nbr[uo + idx]]);
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
        for (uint32_t j = ln; j < l; j += 64) wsum += J.deg[L[j]];
        for (uint32_t j = ln; j < n; j += 64) origp[j] = J.orig[B[j]];
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


// writes the uniform 64-bit value v into lane L of the register pair (lo, hi)
template <int L>
__device__ __forceinline__ void wlane(uint32_t &lo, uint32_t &hi, uint64_t v)
{
    asm("v_writelane_b32 %0, %1, %2" : "+v"(lo) : "s"(uint32_t(v)), "i"(L));
    asm("v_writelane_b32 %0, %1, %2" : "+v"(hi) : "s"(uint32_t(v >> 32)), "i"(L));
}

template <int W>
__device__ __forceinline__ uint64_t rdl64(uint64_t x, uint32_t l)
{
    const uint32_t lo = __builtin_amdgcn_readlane(int(uint32_t(x)), int(l));
    const uint32_t hi = __builtin_amdgcn_readlane(int(uint32_t(x >> 32)), int(l));
    return (uint64_t(hi) << 32) | lo;
}

// Per-wave state that the frame loop touches rarely (root switches, checks,
// idle transitions).  It lives in LDS so that it takes no scalar registers
// from the frame loop; every lane writes the same (uniform) value.
struct Cold {
    uint64_t t_start, t_mark, busy_t, idle_t;
    uint64_t alg, task0, bpt, frames, pops;
    uint64_t plexes, ha, hb, tasks;   // accumulators of the current job (see Worker::tk, Worker::pacc)
    uint64_t la;                      // the current root's left rows
    uint32_t spills;
    uint64_t hungry_since;   // wall-clock tick at which the wave last became hungry
    uint32_t cjob, check_every;
};

// LW: left-vertex rows held in registers (left vertices 64t+lane, t < LW);
// LISTS: the kernel writes plex lists (kept out of the list-free kernel, whose
// frame loop then holds no single-lane branch)
template <int W, int LW, bool LISTS>
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
    uint32_t cur = 0xffffffffu, n = 0, l = 0;
    Cold *cold;
    uint32_t *sel;   // this wave's 64-entry LDS scratch (narrowing)
    // Rows of the vertices this lane owns (64j + lane) are kept in VGPRs.  The
    // search reads one class row per include step (the pivot's, a uniform
    // row), so beyond two words per row the class rows, the pivot's adjacency
    // row and the original ids stay in the root's arena and uniform rows come
    // in through scalar loads: the VGPRs then hold only the adjacency rows
    // that the per-vertex degrees need.
    static constexpr bool kRowRegs = W <= 2;
    static constexpr int RW = kRowRegs ? W : 1;
    uint64_t adjR[W][W], c2R[RW][RW], c4R[RW][RW];
    uint64_t lftR[LW > 0 ? LW : 1][W];
    uint32_t origR[RW];
    // the two per-vertex hash terms of the rows this lane owns, computed once
    // per root (list-free kernels with rows in registers)
    static constexpr bool kHashRegs = kRowRegs && !LISTS;
    uint64_t hAR[kHashRegs ? RW : 1], hBR[kHashRegs ? RW : 1];
    // One-word kernels keep the job's plex count, hash_a and hash_b since
    // the last flush in lanes 0, 1 and 2 of a register (other lanes hold
    // nothing of use); wider ones in the LDS record, which leaves their
    // registers to the frame.
    static constexpr bool kAccRegs = W == 1;
    uint64_t pacc = 0;
    const uint64_t *Ab = nullptr, *C2b = nullptr, *C4b = nullptr;   // the root's arena rows (uniform)
    const uint32_t *OG = nullptr;
    // Branch tasks since the last fold into Cold::tasks (a 32-bit uniform
    // counter: the job's accumulators live in LDS, which keeps the scalar
    // registers of the frame loop free of state it touches once per plex)
    uint32_t tk = 0;
    int err = 0;
    uint32_t nar_on = 0;     // narrowing: 0 not yet (epoch younger than nar.after_ticks), 1 on, 2 buffer full

    __device__ Worker(const SearchArgs &args, uint64_t *ring_, uint64_t *back_, Cold *cold_, uint32_t *sel_ = nullptr)
        : a(args), S(args.sched), ring(ring_), back(back_), cold(cold_), sel(sel_)
    {
        ln = __lane_id();
        k = 0;
        q = 0;
        const uint64_t now = wall_clock64();
        // The time limit runs from the first wave of the epoch, over every
        // class: a class whose grid waits for CUs held by another class does
        // not restart the clock when it starts.
        uint64_t start = now;
        if (a.t0) {
            if (lane0()) {
                const unsigned long long old = atomicCAS(a.t0, 0ull, (unsigned long long)now);
                start = old ? old : now;
            }
            start = rfl64(start);
        }
        cold->t_start = start;
        cold->t_mark = now;
        cold->busy_t = cold->idle_t = 0;
        cold->alg = cold->task0 = cold->bpt = cold->frames = cold->pops = 0;
        cold->plexes = cold->ha = cold->hb = cold->tasks = 0;
        cold->la = 0;
        cold->spills = 0;
        cold->hungry_since = 0;
        cold->cjob = 0xffffffffu;
        // frames between checks of the error flag, the deadline and the pause
        // sources; a frame budget shortens the interval so that small budgets
        // are met about evenly by all waves
        cold->check_every =
            a.pause_frames ? uint32_t(max(uint64_t(32), min(uint64_t(1024), a.pause_frames / max(1u, a.nwaves)))) : 1024u;
    }
    // Re-asserts that the wave-uniform state is uniform.  The optimizer merges
    // "if (uniform) { if (lane == 0) <memory op>; <uniform update> }" into one
    // branch on a lane-dependent condition, and the join then looks
    // lane-dependent to the divergence analysis; readfirstlane at the top of
    // the frame loop cuts that, so the loop body keeps uniform values in
    // scalar registers (a readfirstlane of a scalar register is a plain copy).
    __device__ __forceinline__ void launder()
    {
        rb = rfl(rb);
        rn = rfl(rn);
        bb = rfl(bb);
        bn = rfl(bn);
        cur = rfl(cur);
        n = rfl(n);
        l = rfl(l);
        k = int(rfl(uint32_t(k)));
        q = int(rfl(uint32_t(q)));
        if constexpr (!kRowRegs) {
            Ab = reinterpret_cast<const uint64_t *>(rfl64(reinterpret_cast<uint64_t>(Ab)));
            C2b = reinterpret_cast<const uint64_t *>(rfl64(reinterpret_cast<uint64_t>(C2b)));
            C4b = reinterpret_cast<const uint64_t *>(rfl64(reinterpret_cast<uint64_t>(C4b)));
            OG = reinterpret_cast<const uint32_t *>(rfl64(reinterpret_cast<uint64_t>(OG)));
        }
        tk = rfl(tk);
        err = int(rfl(uint32_t(err)));
        nar_on = rfl(nar_on);
    }
    static __device__ __forceinline__ void launder(Frame<W> &f)
    {
#pragma unroll
        for (int w = 0; w < W; w++) {
            f.P[w] = rfl64(f.P[w]);
            f.C[w] = rfl64(f.C[w]);
            f.D[w] = rfl64(f.D[w]);
            f.X[w] = rfl64(f.X[w]);
        }
        f.root = rfl(f.root);
        f.type = rfl(f.type);
        f.r = rfl(f.r);
    }
    static __device__ __forceinline__ bool uflag(bool b) { return rfl(uint32_t(b)) != 0; }

    // The branch to lane 0 starts with an empty volatile asm: it keeps the
    // optimizer from folding the lane test into an enclosing uniform branch
    // (which would make the join, and the wave-uniform state merged there,
    // look lane-dependent).
    __device__ __forceinline__ bool lane0() const
    {
        asm volatile("");
        return ln == 0;
    }

    __device__ __forceinline__ bool past_deadline() const
    {
        return a.limit_ticks && wall_clock64() - rfl64(cold->t_start) > a.limit_ticks;
    }
    // every lane makes the same compare-and-swap: the first one wins and the
    // rest find the same code, so no lane-dependent branch is needed
    __device__ __forceinline__ void set_error(int code)
    {
        err = code;
        atomicCAS(&S->error, 0, code);
    }

    // ---- frame slots ----
    // Slots of up to 64 words (W <= 4) are moved one word per lane with no
    // branch: lane j holds word j, and lanes past the last word repeat the
    // pad word (slot word FW-1, never read) with the same value, so every
    // lane's store is harmless.  Wider slots go through lane 0.  Loaded words
    // are broadcast with readlane / readfirstlane.
    static constexpr bool kLaneSlots = FW <= 64 && W <= 4;
    // Branch frames of the list-free kernels of these classes run in a
    // dedicated loop (fast_run, fast_run_w)
    static constexpr bool kFastRun = !LISTS && W <= 2;
    // Shared: a queue slot that another wave reads once its sequence number
    // is published.  Its words are written with agent-scope stores, which on
    // MI355X write through the XCD's L2, so the reader (another XCD) finds
    // them in memory once the stores are acknowledged; visibility then does
    // not rest on the release fence's L2 write-back completing before the
    // sequence-number store (the compiler emits no wait after the write-back).
    template <bool Shared = false>
    __device__ __forceinline__ void store_slot(uint64_t *dst, uint64_t meta, const uint64_t *P, const uint64_t *C,
                                               const uint64_t *X, const uint64_t *D, bool withD)
    {
        auto put = [](uint64_t *p, uint64_t v) {
            if constexpr (Shared) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else *p = v;
        };
        if constexpr (kLaneSlots) {
            uint64_t v = meta;
#pragma unroll
            for (int w = 0; w < W; w++) {
                v = inmask(1ull << (1 + w)) ? P[w] : v;
                v = inmask(1ull << (1 + W + w)) ? C[w] : v;
                v = inmask(1ull << (1 + 2 * W + w)) ? X[w] : v;
                if (withD) v = inmask(1ull << (1 + 3 * W + w)) ? D[w] : v;
            }
            put(dst + min(ln, uint32_t(FW - 1)), v);
        } else {
            if (lane0()) {
                put(dst, meta);
#pragma unroll
                for (int w = 0; w < W; w++) {
                    put(dst + 1 + w, P[w]);
                    put(dst + 1 + W + w, C[w]);
                    put(dst + 1 + 2 * W + w, X[w]);
                    if (withD) put(dst + 1 + 3 * W + w, D[w]);
                }
            }
            asm volatile("");
        }
    }
    __device__ __forceinline__ void load_slot(const uint64_t *src, Frame<W> &f)
    {
        if constexpr (kLaneSlots) {
            const uint64_t x = src[min(ln, uint32_t(FW - 1))];
            const uint64_t meta = rdl64<W>(x, 0);
            f.root = uint32_t(meta);
            f.type = uint32_t(meta >> 32) & 0xff;
            f.r = uint32_t(meta >> 40) & 0xffff;
#pragma unroll
            for (int w = 0; w < W; w++) {
                f.P[w] = rdl64<W>(x, 1 + w);
                f.C[w] = rdl64<W>(x, 1 + W + w);
                f.X[w] = rdl64<W>(x, 1 + 2 * W + w);
                f.D[w] = f.type == FT_KSEARCH ? rdl64<W>(x, 1 + 3 * W + w) : 0ull;
            }
        } else {
            uint64_t meta = 0;
            uint64_t t[4 * W];
            if (lane0()) {
                meta = src[0];
#pragma unroll
                for (int i = 0; i < 4 * W; i++) t[i] = src[1 + i];
            }
            asm volatile("");
            meta = rfl64(meta);
            f.root = uint32_t(meta);
            f.type = uint32_t(meta >> 32) & 0xff;
            f.r = uint32_t(meta >> 40) & 0xffff;
#pragma unroll
            for (int w = 0; w < W; w++) {
                f.P[w] = rfl64(t[w]);
                f.C[w] = rfl64(t[W + w]);
                f.X[w] = rfl64(t[2 * W + w]);
                f.D[w] = f.type == FT_KSEARCH ? rfl64(t[3 * W + w]) : 0ull;
            }
        }
    }
    static __device__ __forceinline__ uint64_t make_meta(uint32_t root, uint32_t type, uint32_t r)
    {
        return uint64_t(root) | (uint64_t(type) << 32) | (uint64_t(r) << 40);
    }
    __device__ __forceinline__ void copy_slot(uint64_t *dst, const uint64_t *src)
    {
        if constexpr (kLaneSlots) {
            const uint32_t i = min(ln, uint32_t(FW - 1));
            dst[i] = src[i];
        } else {
            if (lane0()) {
#pragma unroll
                for (int i = 0; i < FW; i++) dst[i] = src[i];
            }
            asm volatile("");
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
    // One-word frames: the LDS ring holds a wave's whole stack (launches size
    // it for the DFS depth bound), so the HBM backing stack is never used.
    static constexpr bool kRingOnly = W == 1;
    __device__ __forceinline__ uint32_t depth() const
    {
        if constexpr (kRingOnly) return rn;
        return rn + bn;
    }

    // a free ring slot on top: a full ring moves its oldest frame to the HBM
    // backing stack
    __device__ __forceinline__ bool make_room()
    {
        if constexpr (kRingOnly) {
            if (rn == a.ring_cap) {
                set_error(KP_EOVERFLOW);
                return false;
            }
            return true;
        }
        if (rn == a.ring_cap) {
            if (bn == a.back_cap) {
                set_error(KP_EOVERFLOW);
                return false;
            }
            copy_slot(back_slot(bn), ring_slot(0));
            bn++;
            rb = (rb + 1 == a.ring_cap) ? 0 : rb + 1;
            rn--;
            cold->spills = rfl(cold->spills) + 1;
        }
        return true;
    }
    __device__ bool push(uint64_t meta, const uint64_t *P, const uint64_t *C, const uint64_t *X, const uint64_t *D,
                         bool withD)
    {
        if (!make_room()) return false;
        store_slot(ring_slot(rn), meta, P, C, X, D, withD);
        rn++;
        return true;
    }
    // pushes a copy of a frame slot (a resumed frame)
    __device__ bool push_copy(const uint64_t *src)
    {
        if (!make_room()) return false;
        copy_slot(ring_slot(rn), src);
        rn++;
        return true;
    }
    __device__ __forceinline__ const uint64_t *seed_slot(uint32_t si) const
    {
        return si < a.nseed_dev ? a.seed + size_t(si) * FW : a.hseed + size_t(si - a.nseed_dev) * FW;
    }
    __device__ __forceinline__ bool push(const Frame<W> &f)
    {
        return push(make_meta(f.root, f.type, f.r), f.P, f.C, f.X, f.D, f.type == FT_KSEARCH);
    }
    __device__ void pop_top(Frame<W> &f)
    {
        if constexpr (kRingOnly) {
            rn--;
            load_slot(ring_slot(rn), f);
            return;
        }
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
        if (kRingOnly || bn == 0) {
            load_slot(ring_slot(0), f);
            rb = (rb + 1 == a.ring_cap) ? 0 : rb + 1;
            rn--;
            return;
        }
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
    __device__ __forceinline__ uint32_t home_shard() const { return rfl((blockIdx.x * 4u + (threadIdx.x >> 6)) % NSUBQ); }

    // Reserves a free slot in some shard (home shard first).  Returns the
    // global slot index, or ~0 when every shard is full.
    __device__ size_t q_reserve(uint32_t &shard_out, uint64_t &ticket_out)
    {
        uint32_t ok = 0, shard = 0;
        uint64_t t = 0;
        if (lane0()) {
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
        if (lane0()) {
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
        store_slot<true>(a.qframes + slot * FW, make_meta(f.root, f.type, f.r), f.P, f.C, f.X, f.D, f.type == FT_KSEARCH);
        // One agent-scope release (an L2 write-back on MI355X, whose L2 is
        // per XCD) orders the frame before the slot's sequence number; the
        // counters are relaxed.  A consumer validates the slot with an
        // acquire load of the sequence number.  The donor stays counted in
        // S->active until after this, and its seq_cst decrement of
        // S->active later orders the q_count increment for the termination
        // test.
        if (lane0()) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_store(&a.qseq[slot], t + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&a.subq[shard].count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&S->q_count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    // The caller has already counted itself in S->active.
    __device__ bool q_pop(Frame<W> &f)
    {
        uint32_t ok = 0, shard = 0;
        uint64_t hd = 0;
        if (lane0()) {
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
        if (lane0()) {
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
        if (lane0()) {
            __hip_atomic_store(&a.qseq[slot], hd + a.qcap, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            atomicSub(&a.subq[shard].size, 1u);
            __hip_atomic_fetch_sub(&S->q_count, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
        }
        cold->pops = cold->pops + 1;
        return true;
    }

    // adds the tasks counted since the last fold to Cold::tasks
    __device__ __forceinline__ void fold()
    {
        cold->tasks = rfl64(cold->tasks) + tk;
        tk = 0;
    }
    // task bytes of the current root so far go to Cold::alg
    __device__ void close_alg()
    {
        fold();
        const uint64_t tasks = rfl64(cold->tasks);
        cold->alg = rfl64(cold->alg) + (tasks - rfl64(cold->task0)) * rfl64(cold->bpt);
        cold->task0 = tasks;
    }
    // adds the accumulators to their job's totals (close_alg first)
    __device__ void flush()
    {
        // LDS values go through readfirstlane: the compiler then knows they
        // are wave-uniform and keeps what depends on them in scalar registers
        fold();
        const uint32_t job = rfl(cold->cjob);
        const uint64_t alg = rfl64(cold->alg);
        uint64_t cnt, ha, hb;
        if constexpr (kAccRegs) {
            cnt = rdl64<W>(pacc, 0);
            ha = rdl64<W>(pacc, 1);
            hb = rdl64<W>(pacc, 2);
        } else {
            cnt = rfl64(cold->plexes);
            ha = rfl64(cold->ha);
            hb = rfl64(cold->hb);
        }
        const uint64_t tasks = rfl64(cold->tasks);
        if (job != 0xffffffffu && (cnt | tasks) && lane0()) {
            JobAcc *J = a.acc + job;
            atomicAdd(&J->count, (unsigned long long)cnt);
            atomicAdd(&J->ha, (unsigned long long)ha);
            atomicAdd(&J->hb, (unsigned long long)hb);
            atomicAdd(&J->tasks, (unsigned long long)tasks);
            atomicAdd(&J->alg, (unsigned long long)alg);
        }
        pacc = 0;
        cold->plexes = cold->ha = cold->hb = cold->tasks = 0;
        cold->alg = 0;
        cold->task0 = 0;
    }

    // ---- root cache: this lane's rows of the current root ----
    // MayFlush: the root may belong to another job than the accumulators
    // (then they are flushed first); callers that pass false have checked
    __device__ __forceinline__ const RootDesc *desc_of(uint32_t ridx) const { return a.desc + ridx; }
    template <bool MayFlush>
    __device__ void load_root(uint32_t ridx)
    {
        const RootDesc *d = desc_of(ridx);
        n = rfl(d->n);
        // a derived root (narrowed from a wider block) keeps its block's size
        // in h for the per-task bytes; h is not needed after push_root
        const uint32_t dh = rfl(d->h);
        const uint32_t blk_n = (dh & 0x80000000u) ? (dh & 0x7fffffffu) : n;
        l = rfl(d->l);
        const uint32_t job = rfl(d->job);
        const uint32_t kq = rfl(*reinterpret_cast<const uint32_t *>(&d->k));
        k = int(kq & 0xffffu);
        q = int(kq >> 16);
        close_alg();
        if constexpr (MayFlush) {
            if (job != rfl(cold->cjob)) {
                flush();
                cold->cjob = job;
            }
        }
        cold->bpt = 2ull * (5ull * blk_n + 16ull);
        const uint64_t *A = reinterpret_cast<const uint64_t *>(rfl64(d->addr));
        const uint64_t *C2 = A + size_t(n) * W;
        const uint64_t *C4 = C2 + size_t(n) * W;
        const uint64_t *LA = C4 + size_t(n) * W;
        cold->la = reinterpret_cast<uint64_t>(LA);
        const uint32_t *og = reinterpret_cast<const uint32_t *>(LA + size_t(l) * W);
        Ab = A;
        C2b = C2;
        C4b = C4;
        OG = og;
        // Lanes past the end load the last row and zero it: no lane-dependent
        // branch, so the caller's join after a root switch stays uniform.
#pragma unroll
        for (int j = 0; j < W; j++) {
            const uint32_t v = 64 * j + ln;
            const bool in = v < n;
            const size_t vv = size_t(min(v, n - 1));
#pragma unroll
            for (int w = 0; w < W; w++) {
                const uint64_t x = A[vv * W + w];
                adjR[j][w] = in ? x : 0ull;
            }
            if constexpr (kRowRegs) {
#pragma unroll
                for (int w = 0; w < W; w++) {
                    const uint64_t y = C2[vv * W + w], z = C4[vv * W + w];
                    c2R[j][w] = in ? y : 0ull;
                    c4R[j][w] = in ? z : 0ull;
                }
                const uint32_t o = og[vv];
                origR[j] = in ? o : 0u;
                if constexpr (kHashRegs) {
                    hAR[j] = mix64(uint64_t(origR[j]) ^ kp::HASH_SA);
                    hBR[j] = mix64(uint64_t(origR[j]) ^ kp::HASH_SB);
                }
            }
        }
#pragma unroll
        for (int t = 0; t < LW; t++) {
            const uint32_t u = 64 * t + ln;
            const size_t uu = size_t(min(u, l - 1));
#pragma unroll
            for (int w = 0; w < W; w++) {
                uint64_t x = 0;
                if (l) x = LA[uu * W + w];
                lftR[t][w] = u < l ? x : 0ull;
            }
        }
        cur = ridx;
    }

    // row of vertex v (uniform) out of the owner lane's registers
    template <int RJ>
    __device__ __forceinline__ void row_of(const uint64_t (&R)[RJ][RJ], uint32_t v, uint64_t *out) const
    {
        const uint32_t pw = v >> 6, pl = v & 63;
#pragma unroll
        for (int j = 0; j < RJ; j++)
            if (uint32_t(j) == pw) {
#pragma unroll
                for (int w = 0; w < RJ; w++) out[w] = rdl64<W>(R[j][w], pl);
            }
    }
    // row of vertex v (uniform) of an arena matrix, through scalar loads (the
    // arena is read-only while the search kernel runs)
    typedef const __attribute__((address_space(4))) uint64_t *ConstRow;
    __device__ __forceinline__ void row_mem(const uint64_t *base, uint32_t v, uint64_t *out) const
    {
        ConstRow p = (ConstRow)(base + size_t(v) * W);
#pragma unroll
        for (int w = 0; w < W; w++) out[w] = p[w];
    }
    __device__ __forceinline__ void adj_row(uint32_t v, uint64_t *out) const
    {
        if constexpr (kRowRegs) row_of(adjR, v, out);
        else row_mem(Ab, v, out);
    }
    __device__ __forceinline__ void cls2_row(uint32_t v, uint64_t *out) const
    {
        if constexpr (kRowRegs) row_of(c2R, v, out);
        else row_mem(C2b, v, out);
    }
    __device__ __forceinline__ void cls4_row(uint32_t v, uint64_t *out) const
    {
        if constexpr (kRowRegs) row_of(c4R, v, out);
        else row_mem(C4b, v, out);
    }
    // original id of vertex 64j + lane
    __device__ __forceinline__ uint32_t orig_of(int j) const
    {
        if constexpr (kRowRegs) return origR[j];
        else return OG[min(64u * j + ln, n - 1)];
    }

    // ---- per-vertex helpers ----
    __device__ __forceinline__ uint32_t deg_in(int j, const uint64_t *Sx) const
    {
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < W; w++) c += __popcll(adjR[j][w] & Sx[w]);
        return c;
    }
    // lane predicate "vertex 64j+lane is in Sx": the uniform word is the lane
    // mask itself (no per-lane shift)
    static __device__ __forceinline__ bool has(const uint64_t *Sx, int j)
    {
        return __builtin_amdgcn_inverse_ballot_w64(Sx[j]);
    }
    static __device__ __forceinline__ bool inmask(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }
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
        uint64_t fail = 0;
#pragma unroll
        for (int j = 0; j < W; j++) fail |= P[j] & __ballot(int(deg_in(j, Sx)) + k < q);
        return fail == 0;
    }

    // kplex-compatibility mask w.r.t. P (isKplex2/isKplex3 :568-598)
    __device__ __forceinline__ void compat_mask(const uint64_t *P, uint32_t ps, uint64_t *OK) const
    {
        uint32_t dP[W];
        uint64_t Sat[W];
#pragma unroll
        for (int j = 0; j < W; j++) {
            dP[j] = deg_in(j, P);
            Sat[j] = P[j] & __ballot(int(dP[j]) + k == int(ps));
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
        for (int j = 0; j < W; j++) Sat[j] = Sx[j] & __ballot(int(deg_in(j, Sx)) + k < int(s) + 1);
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
        const uint64_t *LA = reinterpret_cast<const uint64_t *>(rfl64(cold->la));
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
                if constexpr (kHashRegs) {
                    sa += hAR[j];
                    sb += hBR[j];
                } else {
                    const uint64_t o = orig_of(j);
                    sa += mix64(o ^ kp::HASH_SA);
                    sb += mix64(o ^ kp::HASH_SB);
                }
            }
        sa = rfl64(wsum64(sa));
        sb = rfl64(wsum64(sb));
        if constexpr (kAccRegs) {
            uint32_t lo = 1u, hi = 0u;   // lane 0: one more plex
            wlane<1>(lo, hi, rfl64(mix64(sa + uint64_t(s) * kp::HASH_LA)));
            wlane<2>(lo, hi, rfl64(mix64(sb + uint64_t(s) * kp::HASH_LB)));
            pacc += (uint64_t(hi) << 32) | lo;
        } else {
            cold->plexes = rfl64(cold->plexes) + 1;
            cold->ha = rfl64(cold->ha) + mix64(sa + uint64_t(s) * kp::HASH_LA);
            cold->hb = rfl64(cold->hb) + mix64(sb + uint64_t(s) * kp::HASH_LB);
        }
        if constexpr (LISTS) {
            uint64_t base = 0;
            if (lane0()) base = atomicAdd((unsigned long long *)a.list_used, (unsigned long long)(s + 2));
            base = rfl64(base);
            if (base + s + 2 > a.list_cap) {
                set_error(KP_EOVERFLOW);
                return;
            }
            if (lane0()) {
                a.list[base] = s;
                a.list[base + 1] = cold->cjob;
            }
            uint64_t o = base + 2;
#pragma unroll
            for (int j = 0; j < W; j++) {
                if (has(Sx, j)) a.list[o + __popcll(Sx[j] & lt_mask())] = orig_of(j);
                o += __popcll(Sx[j]);
            }
            asm volatile("");   // the per-lane stores join here, before any other path merges
        }
    }

    __device__ __forceinline__ void count_task() { tk++; }

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
            adj_row(pp, adjPP);
            uint32_t k2 = ~0u;
#pragma unroll
            for (int j = 0; j < W; j++)
                if (inmask(C[j] & ~adjPP[j])) k2 = min(k2, (dG[j] << 20) | (dP[j] << 10) | (64u * j + ln));
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
                for (int j = 0; j < W; j++) Sat[j] = PC[j] & __ballot(int(dG[j]) + k == int(tot));
                uint64_t ext = 0;
#pragma unroll
                for (int j = 0; j < W; j++) ext |= X[j] & __ballot(int(dG[j]) + k >= int(tot) + 1 && subset_of_adj(j, Sat));
                if (ext) return false;
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
        cls2_row(pivot, c2v);
        uint64_t C2[W], PC2[W];
#pragma unroll
        for (int w = 0; w < W; w++) {
            C2[w] = C0[w] & c2v[w] & OK[w];
            PC2[w] = P2[w] | C2[w];
        }
        if (!ub_ok(P2, PC2)) return false;
        cls4_row(pivot, c4v);
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

    // One-word Branch frames of the current root, back to back.  The same
    // steps as branch() for W = 1, with the frame in scalar registers:
    // - the include child's degrees into P2 are dP plus the pivot's bit of
    //   the lane's row (the pivot is a candidate, not in P), not a new count;
    // - the include child is built before the exclude child is stored: when
    //   the include child is pruned, the exclude child is the next frame
    //   and never goes through the ring (branch() pushes it and pops it);
    // - a stored exclude child goes straight into the LDS ring slot above
    //   the top (tix), its words (meta, P, C0, X|pivot) written into lanes
    //   0-3 of one register pair whose lane 0 (the meta word) is set once
    //   (every lane starts with it);
    // - the top slot is read at the start of every frame, so that a frame
    //   that ends without a child finds the next one already loaded (the
    //   read is unused when the frame stores a child, or the stack is empty);
    // - a popped slot is taken only if it is a Branch frame of this root;
    //   any other frame (a k-search frame, another root's) stays on the
    //   stack for the general loop.
    // Runs until nfr reaches burst (the next frame, if any, is left in f and
    // the return value is true), the stack runs dry, the top frame is not
    // one this loop takes, or an error is set (false in those cases).
    __device__ bool fast_run(Frame<W> &f, uint32_t &nfr, uint32_t burst)
    {
        static_assert(W == 1, "one-word frames only");
        uint64_t P = f.P[0], C = f.C[0], X = f.X[0];
        const uint64_t meta = make_meta(cur, FT_BRANCH, 0);
        const uint64_t A = adjR[0][0];
        const uint32_t cap = a.ring_cap;
        const uint32_t lofs = min(ln, uint32_t(FW - 1));
        uint32_t slo = uint32_t(meta), shi = uint32_t(meta >> 32);
        // slots above the entry depth were written by this loop: Branch frames
        // of this root, whose meta word needs no check
        const uint32_t rn0 = rn;
        // ring index of the top slot (of the slot below the bottom when the
        // stack is empty)
        uint32_t tix = rb + rn + cap - 1;
        if (tix >= cap) tix -= cap;
        if (tix >= cap) tix -= cap;
        for (;;) {
            const uint64_t xs = ring[size_t(tix) * FW + lofs];
            bool next = false;   // (P, C, X) holds the next frame
            const uint32_t ps = __popcll(P), cs = __popcll(C);
            if (int(ps + cs) >= q) {
                if (cs == 0) {
                    if (X == 0 && int(ps) >= q && maximal(&P, ps)) record(&P, ps);
                } else {
                    const uint64_t PC = P | C;
                    const uint32_t dG = __popcll(A & PC);
                    uint32_t key = inmask(P) ? (dG << 10) | ln : ~0u;
                    key = rfl(wmin32(key));
                    const int minP = int(key >> 10);
                    if (minP + k >= max(q, int(ps))) {
                        const uint32_t dP = __popcll(A & P);
                        const uint32_t kk = (dG << 20) | (dP << 10) | ln;
                        uint32_t pivot = ~0u;
                        if (minP + k < int(ps + cs)) {
                            const uint64_t adjPP = rdl64<W>(A, key & 1023u);
                            const uint32_t k2 = rfl(wmin32(inmask(C & ~adjPP) ? kk : ~0u));
                            if (k2 == ~0u) set_error(KP_EINVAL);   // unreachable while the degrees are consistent
                            else pivot = k2 & 1023u;
                        } else {
                            const uint32_t k2 = rfl(wmin32(inmask(C) ? kk : ~0u));
                            const int mn = min(minP, int(k2 >> 20));
                            if (mn >= int(ps + cs) - k) {
                                // P ∪ C is a k-plex (isKplexPC2 :600-622 over X, then maximality)
                                const uint32_t tot = ps + cs;
                                const uint64_t Sat = PC & __ballot(int(dG) + k == int(tot));
                                const uint64_t ext = X & __ballot(int(dG) + k >= int(tot) + 1 && (Sat & ~A) == 0);
                                if (ext == 0 && maximal(&PC, tot)) record(&PC, tot);
                            } else {
                                pivot = k2 & 1023u;
                            }
                        }
                        if (pivot != ~0u) {
                            // branchInCand2 (:1467-1472): exclude child (P, C0, X|pivot),
                            // include child (:1273-1465)
                            const uint64_t pb = 1ull << pivot;
                            const uint64_t C0 = C & ~pb;
                            tk++;
                            const uint64_t P2 = P | pb;
                            const uint32_t ps2 = ps + 1;
                            const int dP2 = int(dP + uint32_t((A >> pivot) & 1ull));
                            const uint64_t Sat = P2 & __ballot(dP2 + k == int(ps2));
                            uint64_t OK = __ballot(dP2 + k >= int(ps2) + 1);
                            if (Sat) OK &= __ballot((Sat & ~A) == 0);   // no saturated vertex: nothing to be adjacent to
                            const uint64_t C2 = C0 & rdl64<W>(c2R[0][0], pivot) & OK;
                            const uint64_t PC2 = P2 | C2;
                            const uint64_t XE = X | pb;
                            next = true;
                            if ((P2 & __ballot(int(__popcll(A & PC2)) + k < q)) == 0) {
                                if (rn == cap) {
                                    set_error(KP_EOVERFLOW);
                                } else {
                                    tix = tix + 1 == cap ? 0 : tix + 1;
                                    wlane<1>(slo, shi, P);
                                    wlane<2>(slo, shi, C0);
                                    wlane<3>(slo, shi, XE);
                                    ring[size_t(tix) * FW + lofs] = (uint64_t(shi) << 32) | slo;
                                    rn++;
                                    X = X & rdl64<W>(c4R[0][0], pivot) & OK;
                                    P = P2;
                                    C = C2;
                                    tk++;
                                }
                            } else {
                                C = C0;
                                X = XE;
                            }
                        }
                    }
                }
            }
            nfr++;
            if (err) return false;
            if (nfr == burst) {
                if (next) {
                    f.P[0] = P;
                    f.C[0] = C;
                    f.X[0] = X;
                    f.D[0] = 0;
                    f.root = cur;
                    f.type = FT_BRANCH;
                    f.r = 0;
                }
                return next;
            }
            if (!next) {
                if (rn == 0) return false;
                if (rn <= rn0 && rdl64<W>(xs, 0) != meta) return false;
                rn--;
                tix = tix == 0 ? cap - 1 : tix - 1;
                P = rdl64<W>(xs, 1);
                C = rdl64<W>(xs, 2);
                X = rdl64<W>(xs, 3);
            }
        }
    }

    // fast_run for frames of two or more words: the same steps as branch(),
    // with the frame in scalar registers, the read-ahead of the top ring
    // slot and the exclude child run directly when the include child is
    // pruned (see fast_run).  The ring is backed by the HBM stack here: a full
    // ring moves its oldest frame there (make_room, which leaves the top
    // slot's index unchanged), and when the ring runs dry the general loop
    // pops from the HBM stack.
    template <int B, int... I>
    static __device__ __forceinline__ void wlanes(uint32_t &lo, uint32_t &hi, const uint64_t *v,
                                                  std::integer_sequence<int, I...>)
    {
        (wlane<B + I>(lo, hi, v[I]), ...);
    }
    __device__ bool fast_run_w(Frame<W> &f, uint32_t &nfr, uint32_t burst)
    {
        static_assert(W >= 2 && kLaneSlots, "lane-slot frames of two or more words");
        constexpr auto seq = std::make_integer_sequence<int, W>{};
        uint64_t P[W], C[W], X[W];
#pragma unroll
        for (int w = 0; w < W; w++) {
            P[w] = f.P[w];
            C[w] = f.C[w];
            X[w] = f.X[w];
        }
        const uint64_t meta = make_meta(cur, FT_BRANCH, 0);
        const uint32_t cap = a.ring_cap;
        const uint32_t lofs = min(ln, uint32_t(FW - 1));
        uint32_t slo = uint32_t(meta), shi = uint32_t(meta >> 32);
        uint32_t tix = rb + rn + cap - 1;
        if (tix >= cap) tix -= cap;
        if (tix >= cap) tix -= cap;
        for (;;) {
            const uint64_t xs = ring[size_t(tix) * FW + lofs];
            bool next = false;   // (P, C, X) holds the next frame
            const uint32_t ps = popc(P), cs = popc(C);
            if (int(ps + cs) >= q) {
                if (cs == 0) {
                    if (popc(X) == 0 && int(ps) >= q && maximal(P, ps)) record(P, ps);
                } else {
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
                    if (minP + k >= max(q, int(ps))) {
                        uint32_t pivot = ~0u;
                        if (minP + k < int(ps + cs)) {
                            uint64_t adjPP[W];
                            adj_row(key & 1023u, adjPP);
                            uint32_t k2 = ~0u;
#pragma unroll
                            for (int j = 0; j < W; j++)
                                if (inmask(C[j] & ~adjPP[j])) k2 = min(k2, (dG[j] << 20) | (dP[j] << 10) | (64u * j + ln));
                            k2 = rfl(wmin32(k2));
                            if (k2 == ~0u) set_error(KP_EINVAL);   // unreachable while the degrees are consistent
                            else pivot = k2 & 1023u;
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
                                for (int j = 0; j < W; j++) Sat[j] = PC[j] & __ballot(int(dG[j]) + k == int(tot));
                                uint64_t ext = 0;
#pragma unroll
                                for (int j = 0; j < W; j++)
                                    ext |= X[j] & __ballot(int(dG[j]) + k >= int(tot) + 1 && subset_of_adj(j, Sat));
                                if (ext == 0 && maximal(PC, tot)) record(PC, tot);
                            } else {
                                pivot = k2 & 1023u;
                            }
                        }
                        if (pivot != ~0u) {
                            // branchInCand2 (:1467-1472): exclude child (P, C0, X|pivot),
                            // include child (:1273-1465)
                            const uint32_t pw = pivot >> 6, pl = pivot & 63u;
                            const uint64_t pb = 1ull << pl;
                            uint64_t C0[W], P2[W], XE[W];
#pragma unroll
                            for (int w = 0; w < W; w++) {
                                const uint64_t m = w == int(pw) ? pb : 0ull;
                                C0[w] = C[w] & ~m;
                                P2[w] = P[w] | m;
                                XE[w] = X[w] | m;
                            }
                            tk++;
                            const uint32_t ps2 = ps + 1;
                            // degrees into P2: the pivot (a candidate) adds its bit of each row
                            int dP2[W];
#pragma unroll
                            for (int j = 0; j < W; j++) {
                                uint64_t col = 0;
#pragma unroll
                                for (int w = 0; w < W; w++)
                                    if (w == int(pw)) col = adjR[j][w];
                                dP2[j] = int(dP[j] + uint32_t((col >> pl) & 1ull));
                            }
                            uint64_t Sat[W], OK[W];
#pragma unroll
                            for (int j = 0; j < W; j++) Sat[j] = P2[j] & __ballot(dP2[j] + k == int(ps2));
#pragma unroll
                            for (int j = 0; j < W; j++)
                                OK[j] = __ballot(dP2[j] + k >= int(ps2) + 1 && subset_of_adj(j, Sat));
                            uint64_t c2v[W], C2[W], PC2[W];
                            cls2_row(pivot, c2v);
#pragma unroll
                            for (int w = 0; w < W; w++) {
                                C2[w] = C0[w] & c2v[w] & OK[w];
                                PC2[w] = P2[w] | C2[w];
                            }
                            next = true;
                            if (ub_ok(P2, PC2)) {
                                if (make_room()) {
                                    tix = tix + 1 == cap ? 0 : tix + 1;
                                    wlanes<1>(slo, shi, P, seq);
                                    wlanes<1 + W>(slo, shi, C0, seq);
                                    wlanes<1 + 2 * W>(slo, shi, XE, seq);
                                    ring[size_t(tix) * FW + lofs] = (uint64_t(shi) << 32) | slo;
                                    rn++;
                                    uint64_t c4v[W];
                                    cls4_row(pivot, c4v);
#pragma unroll
                                    for (int w = 0; w < W; w++) {
                                        X[w] = X[w] & c4v[w] & OK[w];
                                        P[w] = P2[w];
                                        C[w] = C2[w];
                                    }
                                    tk++;
                                }
                            } else {
#pragma unroll
                                for (int w = 0; w < W; w++) {
                                    C[w] = C0[w];
                                    X[w] = XE[w];
                                }
                            }
                        }
                    }
                }
            }
            nfr++;
            if (err) return false;
            if (nfr == burst) {
                if (next) {
#pragma unroll
                    for (int w = 0; w < W; w++) {
                        f.P[w] = P[w];
                        f.C[w] = C[w];
                        f.X[w] = X[w];
                        f.D[w] = 0;
                    }
                    f.root = cur;
                    f.type = FT_BRANCH;
                    f.r = 0;
                }
                return next;
            }
            if (!next) {
                if (rn == 0) return false;
                if (rdl64<W>(xs, 0) != meta) return false;
                rn--;
                tix = tix == 0 ? cap - 1 : tix - 1;
#pragma unroll
                for (int w = 0; w < W; w++) {
                    P[w] = rdl64<W>(xs, 1 + w);
                    C[w] = rdl64<W>(xs, 1 + W + w);
                    X[w] = rdl64<W>(xs, 1 + 2 * W + w);
                }
            }
        }
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
        cls2_row(u, c2u);
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

    // ---- narrowing (wide classes): re-root a frame whose live set fits one word ----
    // Bits of a W-word row r at the positions of the uniform set V, packed in
    // order (V holds at most 64 vertices).
    static __device__ __forceinline__ uint64_t gather_bits(const uint64_t *r, const uint64_t *V)
    {
        uint64_t out = 0;
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < W; w++) {
            uint64_t mw = V[w];
            const uint64_t x = r[w];   // a per-lane row in memory, read word by word
            while (mw) {
                const uint32_t b = uint32_t(__builtin_ctzll(mw));
                mw &= mw - 1;
                out |= ((x >> b) & 1ull) << t;
                t++;
            }
        }
        return out;
    }
    // lane predicate: vertex s (per lane) is in the uniform set Sx
    static __device__ __forceinline__ bool has_v(const uint64_t *Sx, uint32_t s)
    {
        uint64_t word = 0;
#pragma unroll
        for (int w = 0; w < W; w++) word = (s >> 6) == uint32_t(w) ? Sx[w] : word;
        return (word >> (s & 63)) & 1ull;
    }
    // Left vertex u can still extend a plex S of this subtree (P ⊆ S ⊆ U =
    // P ∪ C ∪ D, |S| >= q) only if it misses at most k-1 vertices of P, is
    // adjacent to every saturated vertex of P (they stay saturated in S), and
    // has at least q+1-k neighbours in U: maximal() tests exactly these
    // conditions on S, so a left vertex that fails one never extends.
    __device__ __forceinline__ bool left_keeps(const uint64_t *LA, uint32_t u, const uint64_t *P, const uint64_t *U,
                                               const uint64_t *SatP, uint32_t ps) const
    {
        const size_t uu = size_t(min(u, l - 1));
        int c = 0, cp = 0;
        uint64_t bad = 0;
#pragma unroll
        for (int w = 0; w < W; w++) {
            const uint64_t x = LA[uu * W + w];
            c += __popcll(x & U[w]);
            cp += __popcll(x & P[w]);
            bad |= SatP[w] & ~x;
        }
        return u < l && bad == 0 && cp + k >= int(ps) + 1 && c + k >= q + 1;
    }
    // Called before a frame is processed.  The excluded set X is tested only
    // for extension (an x in X blocks a record when S + x is a k-plex, and is
    // kept at an include step only when compatible with the new P), so an x
    // that is incompatible with P (then with every P' ⊇ P) or has fewer than
    // q+1-k neighbours in U (then never extends a plex of size >= q) changes no
    // test; such vertices are dropped for good.  If P ∪ C ∪ D ∪ X then fits
    // one word, the frame becomes a derived root of the one-word class with
    // those vertices in block order (pivot and k-search choices depend only on
    // that order, so the subtree is the same) and returns true.
    __device__ bool try_narrow(Frame<W> &f)
    {
        const NarrowArgs &N = a.nar;
        uint64_t U[W];
        uint32_t cu = 0, cd = 0;
#pragma unroll
        for (int w = 0; w < W; w++) {
            U[w] = f.P[w] | f.C[w] | f.D[w];
            cu += __popcll(U[w]);
            cd += __popcll(f.C[w] | f.D[w]);
        }
        const uint32_t ps = popc(f.P);
        if (cu > 64 || ps + cd < uint32_t(q) + N.min_slack) return false;
        uint64_t OK[W], V[W], SatP[W];
        compat_mask(f.P, ps, OK);
        uint32_t m = 0;
#pragma unroll
        for (int j = 0; j < W; j++) {
            f.X[j] &= OK[j] & __ballot(int(deg_in(j, U)) + k >= q + 1);
            V[j] = U[j] | f.X[j];
            m += __popcll(V[j]);
            SatP[j] = f.P[j] & __ballot(int(deg_in(j, f.P)) + k == int(ps));
        }
        if (m > 64) return false;
        const uint64_t *LA = reinterpret_cast<const uint64_t *>(rfl64(cold->la));
        uint32_t nl = 0;
        for (uint32_t t = 0; t < l; t += 64) nl += uint32_t(__popcll(__ballot(left_keeps(LA, t + ln, f.P, U, SatP, ps))));
        // subgraph words: adj, cls2, cls4 rows (one word each), left rows, original ids
        const uint64_t words = 3ull * m + nl + (m + 1) / 2;
        uint64_t base = 0, idx = 0;
        uint32_t ok = 0;
        if (lane0()) {
            base = atomicAdd(&N.ctr[0], (unsigned long long)words);
            if (base + words <= N.arena_cap) {
                idx = atomicAdd(&N.ctr[1], 1ull);
                ok = idx < N.cap;
            }
        }
        if (!rfl(ok)) {
            // no more narrowing in this launch (the next reservations would
            // fail too); the class pauses if asked to, so that the next epoch
            // resumes it with an empty buffer
            nar_on = 2;
            if (N.pause_when_full && lane0()) __hip_atomic_store(&S->pause, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        base = rfl64(base);
        idx = rfl64(idx);
        uint64_t *A = N.arena + base;
        // lane t takes the vertex of compact index t
        uint32_t off = 0;
#pragma unroll
        for (int j = 0; j < W; j++) {
            if (V[j]) {
                if (inmask(V[j])) sel[off + __popcll(V[j] & lt_mask())] = 64u * j + ln;
                off += __popcll(V[j]);
            }
        }
        const uint32_t s = sel[min(ln, m - 1)];
        // the three rows of vertex s, gathered word by word
        uint64_t ga = 0, g2 = 0, g4 = 0;
        {
            uint32_t t = 0;
#pragma unroll
            for (int w = 0; w < W; w++) {
                const uint64_t xa = Ab[size_t(s) * W + w], x2 = C2b[size_t(s) * W + w], x4 = C4b[size_t(s) * W + w];
                uint64_t mw = V[w];
                while (mw) {
                    const uint32_t b = uint32_t(__builtin_ctzll(mw));
                    mw &= mw - 1;
                    ga |= ((xa >> b) & 1ull) << t;
                    g2 |= ((x2 >> b) & 1ull) << t;
                    g4 |= ((x4 >> b) & 1ull) << t;
                    t++;
                }
            }
        }
        const uint32_t og = OG[s];
        uint32_t *ogd = reinterpret_cast<uint32_t *>(A + 3ull * m + nl);
        if (ln < m) {
            A[ln] = ga;
            A[m + ln] = g2;
            A[2 * m + ln] = g4;
            ogd[ln] = og;
        }
        uint32_t pos = 0;
        for (uint32_t t = 0; t < l; t += 64) {
            const uint32_t u = t + ln;
            const bool kp = left_keeps(LA, u, f.P, U, SatP, ps);
            const uint64_t bm = __ballot(kp);
            if (bm) {
                const uint64_t gl = gather_bits(LA + size_t(min(u, l - 1)) * W, V);
                if (kp) A[3ull * m + pos + __popcll(bm & lt_mask())] = gl;
                pos += uint32_t(__popcll(bm));
            }
        }
        const uint64_t live = m == 64 ? ~0ull : ((1ull << m) - 1ull);
        const bool in = ln < m;
        const uint64_t P1 = __ballot(in && has_v(f.P, s)), C1 = __ballot(in && has_v(f.C, s));
        const uint64_t X1 = __ballot(in && has_v(f.X, s)), D1 = __ballot(in && has_v(f.D, s));
        if (lane0()) {
            RootDesc d{};
            d.addr = reinterpret_cast<uint64_t>(A);
            d.n = m;
            d.h = 0x80000000u | n;
            d.l = nl;
            d.slot = 0;
            d.job = cold->cjob;
            d.k = uint16_t(k);
            d.q = uint16_t(q);
            N.desc[idx] = d;
            uint64_t *F = N.frames + idx * uint64_t(FrameLayout<1>::FW);
            F[0] = make_meta(N.tag + uint32_t(idx), f.type, f.r);
            F[1] = P1 & live;
            F[2] = C1 & live;
            F[3] = X1 & live;
            F[4] = D1 & live;
            F[5] = 0;
        }
        asm volatile("");
        return true;
    }

    __device__ void push_root(uint32_t ridx)
    {
        const RootDesc *d = a.desc + ridx;
        const uint32_t rn_ = rfl(d->n), rh = rfl(d->h);
        const uint32_t rk = rfl(*reinterpret_cast<const uint32_t *>(&d->k)) & 0xffffu;
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
        push(make_meta(ridx, FT_KSEARCH, rk - 1u), P, C, X, D, true);
    }

    // Pause sources: the scheduler flag, the host word, the frame budget.
    // Returns true once a pause is in force.
    __device__ bool pause_requested(uint32_t new_frames)
    {
        uint32_t pz = 0;
        if (lane0()) {
            pz = __hip_atomic_load(&S->pause, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (!pz && a.host_pause &&
                __hip_atomic_load(a.host_pause, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0)
                pz = 1;
            if (!pz && a.class_pause &&
                __hip_atomic_load(a.class_pause, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0)
                pz = 1;
            if (!pz && a.pause_frames && new_frames) {
                const uint64_t old = atomicAdd((unsigned long long *)&S->epoch_frames, (unsigned long long)new_frames);
                if (old + new_frames >= a.pause_frames) pz = 1;
            }
            if (pz) __hip_atomic_store(&S->pause, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        return rfl(pz) != 0;
    }

    // Writes this wave's pending frames (the next frame, then the stack) to
    // the dump array.
    __device__ void dump_frames(bool have, const Frame<W> &f)
    {
        const uint32_t m = depth() + (have ? 1u : 0u);
        if (m == 0) return;
        uint32_t base = 0;
        if (lane0()) base = atomicAdd(&S->dump_used, m);
        base = rfl(base);
        if (uint64_t(base) + m > a.dump_cap) {
            set_error(KP_EOVERFLOW);
            return;
        }
        uint64_t *d = a.dump + size_t(base) * FW;
        if (have) {
            store_slot(d, make_meta(f.root, f.type, f.r), f.P, f.C, f.X, f.D, f.type == FT_KSEARCH);
            d += FW;
        }
        for (uint32_t i = 0; i < bn; i++, d += FW) copy_slot(d, back_slot(i));
        for (uint32_t i = 0; i < rn; i++, d += FW) copy_slot(d, ring_slot(i));
        rn = bn = 0;
    }

    // Drops the accumulators of the root in progress (its subtree goes to
    // another path, which counts it from the start).
    __device__ void discard()
    {
        tk = 0;
        pacc = 0;
        cold->plexes = cold->ha = cold->hb = cold->tasks = 0;
        cold->alg = 0;
        cold->task0 = 0;
        rn = bn = 0;
    }

    // Monolithic per-root mode: this wave walks the whole subtree of root
    // ridx on its own stack, with no queue, donation or pause.  Returns true
    // when the subtree is done (accumulators flushed to the root's job), false
    // once more than `budget` frames were processed (0 = no budget) or on an
    // error; then the root's accumulators are dropped.
    __device__ bool run_single(uint32_t ridx, uint64_t budget)
    {
        cur = 0xffffffffu;   // the descriptor slot is reused root after root
        if (past_deadline()) set_error(KP_ETIMEDOUT);
        if (err) return false;
        push_root(ridx);
        uint64_t frames = 0;
        Frame<W> f;
        while (depth() > 0 && !err) {
            launder();
            pop_top(f);
            for (;;) {
                if (f.root != cur) load_root<true>(f.root);
                Frame<W> nf;
                const bool more = (f.type == FT_KSEARCH) ? ksearch(f, nf) : branch(f, nf);
                frames++;
                if ((frames & 1023u) == 0 && past_deadline()) set_error(KP_ETIMEDOUT);
                if (err || !more) break;
                f = nf;
                if (budget && frames > budget) break;
            }
            if (budget && frames > budget) break;
        }
        if (err || (budget && frames > budget)) {
            discard();
            return false;
        }
        close_alg();
        flush();
        return true;
    }

    // Termination: a wave counts itself in S->active only while it holds
    // frames, and it increments S->active before taking a frame, a seed or a
    // root (seq_cst), so "no seeds or roots left, queue empty, nobody
    // active", read in that order, means no frame exists anywhere.  A wave
    // that reads it exits; a wave that still holds work never exits before
    // its stack is empty, unless a pause is in force: then it dumps its
    // frames and exits, and the last wave to exit moves the queued frames to
    // the dump as well.  Waves without work count themselves in S->hungry;
    // busy waves donate their oldest frames (the largest pending subtrees)
    // while hungry waves outnumber the queued frames.
    // Narrowing starts once the epoch has run for nar.after_ticks: a short
    // epoch gains nothing from it (its derived roots would wait for the next
    // epoch), a long one (blocks of heavy hub roots) gains much.
    __device__ __forceinline__ uint32_t narrow_state() const
    {
        if (nar_on != 0 || a.nar.cap == 0) return nar_on;
        return wall_clock64() - rfl64(cold->t_start) >= a.nar.after_ticks ? 1u : 0u;
    }
    __device__ void run()
    {
        if (lane0()) atomicAdd(&S->started, 1u);
        if constexpr (W >= 2) nar_on = narrow_state();
        bool active_self = false, hungry_self = false, have = false, paused = false;
        uint32_t to_donate = a.donate_period, backoff = 1;   // frames left until the next donation check
        uint32_t to_check = rfl(cold->check_every);   // frames left until the next check
        const uint32_t gw = rfl(blockIdx.x * 4u + (threadIdx.x >> 6));
        Frame<W> f;
        for (;;) {
            launder();
            launder(f);
            have = uflag(have);
            active_self = uflag(active_self);
            hungry_self = uflag(hungry_self);
            to_donate = rfl(to_donate);
            to_check = rfl(to_check);
            backoff = rfl(backoff);
            if (!have) {
                if (depth() == 0) {
                    if (active_self) {
                        const uint64_t now = wall_clock64();
                        cold->busy_t = cold->busy_t + (now - cold->t_mark);
                        cold->t_mark = now;
                        if (lane0()) __hip_atomic_fetch_sub(&S->active, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                        active_self = false;
                    }
                    if (err) break;
                    // hints only (relaxed): claims below are atomic read-modify-writes
                    bool got = false;
                    uint32_t qc = 0, nr = 0, ns = 0;
                    if (lane0()) {
                        qc = __hip_atomic_load(&S->q_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ns = __hip_atomic_load(&S->next_seed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        nr = __hip_atomic_load(&S->next_root, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    qc = rfl(qc);
                    ns = rfl(ns);
                    nr = rfl(nr);
                    if (qc > 0) {
                        if (lane0()) __hip_atomic_fetch_add(&S->active, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                        got = q_pop(f);
                        if (got) {
                            active_self = true;
                            have = true;
                        } else if (lane0()) {
                            __hip_atomic_fetch_sub(&S->active, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                        }
                        if (err) break;
                    }
                    if (!got && ns < a.nseed) {
                        // Resumed frames are claimed a few at a time (every
                        // wave shares the one cursor): the first is processed
                        // now, the others go on this wave's stack, from which
                        // donations can still move them.
                        const uint32_t want = max(1u, min(a.seed_chunk, (a.nseed - ns) / (2u * a.nwaves)));
                        uint32_t si = 0xffffffffu;
                        if (lane0()) {
                            __hip_atomic_fetch_add(&S->active, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                            si = __hip_atomic_fetch_add(&S->next_seed, want, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                            if (si >= a.nseed)
                                __hip_atomic_fetch_sub(&S->active, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                        }
                        si = rfl(si);
                        if (si < a.nseed) {
                            active_self = true;
                            const uint32_t cnt = min(want, a.nseed - si);
                            for (uint32_t j = cnt; j-- > 1;)
                                if (!push_copy(seed_slot(si + j))) break;
                            load_slot(seed_slot(si), f);
                            have = true;
                            got = true;
                        }
                        if (err) break;
                    }
                    if (!got && nr < a.ntodo) {
                        uint32_t rr = 0xffffffffu;
                        if (lane0()) {
                            __hip_atomic_fetch_add(&S->active, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                            rr = __hip_atomic_fetch_add(&S->next_root, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                            if (rr >= a.ntodo)
                                __hip_atomic_fetch_sub(&S->active, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                        }
                        rr = rfl(rr);
                        if (rr < a.ntodo) {
                            active_self = true;
                            push_root(a.todo[rr]);
                            got = true;
                        }
                    }
                    if (!got) {
                        if (!hungry_self) {
                            if (lane0()) atomicAdd(&S->hungry, 1u);
                            hungry_self = true;
                            cold->hungry_since = wall_clock64();
                        }
                        uint32_t done = 0, stop = 0;
                        if (lane0()) {
                            // a relaxed look first; the seq_cst reads (in this
                            // order) confirm termination
                            const bool maybe =
                                __hip_atomic_load(&S->active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0 &&
                                __hip_atomic_load(&S->q_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
                            if (maybe) {
                                const uint32_t ns2 = __hip_atomic_load(&S->next_seed, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                                const uint32_t nr2 = __hip_atomic_load(&S->next_root, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                                const uint32_t qc2 = __hip_atomic_load(&S->q_count, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                                const uint32_t ac = __hip_atomic_load(&S->active, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                                done = ns2 >= a.nseed && nr2 >= a.ntodo && qc2 == 0 && ac == 0;
                            }
                            stop = __hip_atomic_load(&S->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
                            if (past_deadline()) stop = 2;
                        }
                        done = rfl(done);
                        stop = rfl(stop);
                        if (stop == 2) set_error(KP_ETIMEDOUT);
                        if (done || stop) break;
                        if (pause_requested(0)) {
                            paused = true;
                            break;
                        }
                        // Retirement: once the class has nothing left to start,
                        // a wave that has stayed hungry for retire_ticks while
                        // hungry waves outnumber the busy ones by more than a
                        // margin exits, so that the CU slots it holds go to
                        // other size classes whose blocks wait for them.  While
                        // busy waves hold large subtrees their donations feed
                        // hungry waves within a millisecond or so, so only a
                        // class whose remaining work is thin sheds waves; the
                        // margin keeps enough hungry waves for donations to
                        // double the busy set each round.
                        // A class whose own grid is not yet all resident, or
                        // that runs alone (retire_ticks 0), keeps its waves.
                        if (a.retire_ticks && backoff >= 16 && ns >= a.nseed && nr >= a.ntodo &&
                            wall_clock64() - rfl64(cold->hungry_since) > a.retire_ticks) {
                            uint32_t retire = 0;
                            if (lane0() && __hip_atomic_load(&S->started, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= a.nwaves) {
                                const uint32_t ac = __hip_atomic_load(&S->active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                const uint32_t qc2 = __hip_atomic_load(&S->q_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                const uint32_t old = atomicSub(&S->hungry, 1u);
                                if (qc2 == 0 && old > ac + a.retire_margin + 1u) {
                                    retire = 1;
                                    atomicAdd(&S->retired, 1u);
                                } else {
                                    atomicAdd(&S->hungry, 1u);
                                }
                            }
                            if (rfl(retire)) {
                                hungry_self = false;
                                break;
                            }
                        }
                        // jittered exponential backoff: waves do not poll in lockstep
                        const uint32_t jit = (gw * 2654435761u + backoff * 97u) >> 28;
                        for (uint32_t z = 0; z < backoff + jit; z++) __builtin_amdgcn_s_sleep(16);
                        if (backoff < 16) backoff <<= 1;
                        continue;
                    }
                    if (hungry_self) {
                        if (lane0()) atomicSub(&S->hungry, 1u);
                        hungry_self = false;
                    }
                    backoff = 1;
                    const uint64_t now = wall_clock64();
                    cold->idle_t = cold->idle_t + (now - cold->t_mark);
                    cold->t_mark = now;
                }
                if (!have) pop_top(f);
            }
            // A burst of frames from the wave's own stack, up to the next
            // donation or check point.  The burst loop does no cross-wave work,
            // so its control flow stays uniform and its state scalar.
            const uint32_t burst = min(to_donate, to_check);
            if (f.root != cur) load_root<true>(f.root);
            have = true;   // f holds the next frame
            uint32_t nfr = 0;
            for (;;) {
                if (f.root != cur) {
                    // a root of another job: its accumulators are flushed
                    // outside the burst (the frame stays pending in f)
                    if (rfl(desc_of(f.root)->job) != rfl(cold->cjob)) break;
                    load_root<false>(f.root);
                }
                Frame<W> nf;
                bool narrowed = false;
                if constexpr (W >= 2) narrowed = nar_on == 1 && try_narrow(f);
                if (narrowed) {
                    have = false;
                } else if (kFastRun && f.type == FT_BRANCH && nar_on != 1) {
                    if constexpr (kFastRun) {
                        // counts its own frames (nfr) and pops its own stack
                        if constexpr (W == 1) have = fast_run(f, nfr, burst);
                        else have = fast_run_w(f, nfr, burst);
                        if (err || nfr == burst) break;
                        if (!have) {
                            if (depth() == 0) break;
                            pop_top(f);
                            have = true;
                        }
                    }
                    continue;
                } else if constexpr (W >= 2) {
                    // wide frames: the child is built in place (soc-epinions,
                    // mostly two-word frames: 188 -> 160 ms); one-word frames
                    // keep the copy, which the register allocator handles
                    // better (75k window: 303 vs 309 ms in place)
                    have = (f.type == FT_KSEARCH) ? ksearch(f, f) : branch(f, f);
                } else {
                    have = (f.type == FT_KSEARCH) ? ksearch(f, nf) : branch(f, nf);
                    if (have) f = nf;
                }
                nfr++;
                if (err) break;
                if (nfr == burst) break;
                if (!have) {
                    if (depth() == 0) break;
                    pop_top(f);
                    have = true;
                }
            }
            if (err) break;
            to_donate -= nfr;
            to_check -= nfr;
            if (to_donate == 0) {
                to_donate = a.donate_period;
                const uint32_t avail = depth() + (have ? 1u : 0u);
                if (avail >= 2 && depth() >= 1) {
                    uint32_t want = 0;
                    if (lane0()) {
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
                        if (lane0()) atomicAdd((unsigned long long *)&S->donated, 1ull);
                    }
                    if (err) break;
                }
            }
            if (to_check == 0) {
                const uint32_t ce = rfl(cold->check_every);
                to_check = ce;
                fold();
                cold->frames = cold->frames + ce;
                if constexpr (W >= 2) nar_on = narrow_state();
                uint32_t stop = 0;
                if (lane0()) {
                    stop = __hip_atomic_load(&S->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
                    if (past_deadline()) stop = 2;
                }
                stop = rfl(stop);
                if (stop == 2) set_error(KP_ETIMEDOUT);
                if (stop) break;
                if (pause_requested(ce)) {
                    paused = true;
                    break;
                }
            }
        }
        if (paused && !err) dump_frames(have, f);
        if (active_self && lane0()) __hip_atomic_fetch_sub(&S->active, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
        if (hungry_self && lane0()) atomicSub(&S->hungry, 1u);
        {
            const uint64_t now = wall_clock64();
            if (active_self) cold->busy_t = cold->busy_t + (now - cold->t_mark);
            else cold->idle_t = cold->idle_t + (now - cold->t_mark);
        }
        close_alg();
        flush();
        const uint64_t busy_t = rfl64(cold->busy_t), idle_t = rfl64(cold->idle_t), pops = rfl64(cold->pops);
        const uint32_t spills = rfl(cold->spills);
        const uint64_t frames = rfl64(cold->frames) + (rfl(cold->check_every) - to_check);
        uint32_t last = 0;
        if (lane0()) {
            // zero terms are skipped: most waves of a launch over little work
            // never ran a frame, and every add here hits the same line
            if (busy_t) atomicAdd((unsigned long long *)&S->busy_ticks, (unsigned long long)busy_t);
            if (idle_t) atomicAdd((unsigned long long *)&S->idle_ticks, (unsigned long long)idle_t);
            if (frames) atomicAdd((unsigned long long *)&S->frames, (unsigned long long)frames);
            if (spills) atomicAdd((unsigned long long *)&S->spills, (unsigned long long)spills);
            if (pops) atomicAdd((unsigned long long *)&S->pops, (unsigned long long)pops);
            last = __hip_atomic_fetch_add(&S->exited, 1u, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT) + 1 == a.nwaves;
        }
        // The last wave out moves frames still queued to the dump: every other
        // wave has exited, so nothing publishes or pops concurrently.
        if (rfl(last)) {
            uint32_t pz = 0, qc = 0;
            if (lane0()) {
                pz = __hip_atomic_load(&S->pause, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
                qc = __hip_atomic_load(&S->q_count, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (rfl(pz) && !err) {
                for (uint32_t i = rfl(qc); i > 0 && !err; i--) {
                    if (!q_pop(f)) break;
                    dump_frames(true, f);
                }
            }
        }
    }
};

// Waves per SIMD (list-free kernels, compiler report): one word 6-7 (74
// VGPRs, 71 in the kp_search1.hip build; asked for at least 4), two to four
// words 4-5 (89-109 VGPRs), six words 2 (170), eight words 2 when asked for
// (228, no scratch); sixteen words 1 with 4,800 bytes of scratch per lane
// (628 without the narrowing code, whose gathers need the registers that the
// sixteen-word frame already fills).
// V tags a second build of the same kernel (kp_search1.hip: V = 1).
template <int W, int LW, bool LISTS, int V = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W == 1 ? 4 : (W == 8 ? 2 : 1)))) void search_kernel(SearchArgs a)
{
    extern __shared__ uint64_t lds[];
    __shared__ Cold s_cold[4];
    __shared__ uint32_t s_sel[4][W >= 2 ? 64 : 1];
    constexpr int FW = FrameLayout<W>::FW;
    // the wave index is uniform within a wave; readfirstlane tells the
    // compiler, so that everything derived from it (ring, stack, LDS state,
    // the frame loop's control flow) stays scalar
    const int wv = int(rfl(threadIdx.x >> 6));
    const uint32_t gw = blockIdx.x * (blockDim.x >> 6) + wv;
    if (gw >= a.nwaves) return;
    uint64_t *ring = lds + size_t(wv) * a.ring_cap * FW;
    uint64_t *back = a.gstack ? a.gstack + size_t(gw) * a.back_cap * FW : nullptr;
    Worker<W, LW, LISTS> wk(a, ring, back, &s_cold[wv], s_sel[wv]);
    wk.run();
}


// ============================================================ per-root
// Monolithic per-root kernel: one wave per root decomposes it, builds its
// subgraph into the wave's own arena slot and walks the whole subtree on its
// own stack (Worker::run_single): one launch, no queue, no donation, no host
// step between the phases.  [reference: the per-root kSearch3/BNB3 path,
// src/device_funcs.cu:3027-4403, its "naive root-based" form]  Roots whose
// block is wider than the kernel's W words, or whose subtree passes the frame
// budget, are listed in `deferred` for a wider kernel or the pooled path.
template <int W, int LW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W == 1 ? 4 : 1))) void root_kernel(RootKernArgs a)
{
    __shared__ uint64_t s_row[WPB][64];
    __shared__ uint64_t s_alive[WPB][64];
    __shared__ uint32_t s_b2[WPB][NCAP];
    __shared__ Cold s_cold[WPB];
    extern __shared__ uint64_t lds[];
    constexpr int FW = FrameLayout<W>::FW;
    const int wv = int(rfl(threadIdx.x >> 6));
    const uint32_t gw = blockIdx.x * WPB + wv;
    if (gw >= a.d.nwaves) return;
    const int ln = __lane_id();
    const DecompWs ws{a.d.ws_cnt + size_t(gw) * a.d.cnt_stride, a.d.ws_b1c + size_t(gw) * a.d.d1cap,
                      a.d.ws_a1 + size_t(gw) * a.d.d1cap * a.d.w1cap, s_row[wv], s_alive[wv], s_b2[wv]};
    uint64_t *ring = lds + size_t(wv) * a.s.ring_cap * FW;
    uint64_t *back = a.s.gstack ? a.s.gstack + size_t(gw) * a.s.back_cap * FW : nullptr;
    Worker<W, LW, false> wk(a.s, ring, back, &s_cold[wv]);
    uint64_t *arena = a.arena + size_t(gw) * a.arena_words;
    for (uint32_t slot = gw; slot < a.d.nslots; slot += a.d.nwaves) {
        const uint32_t job = rfl(a.d.slot_job[slot]);
        const JobDev J = a.d.jobs[job];
        decompose_root(a.d, J, int32_t(rfl(a.d.slot_pos[slot])), gw, ws);
        fence_agent();
        RootInfo ri = a.d.info[gw];
        ri.status = int32_t(rfl(uint32_t(ri.status)));
        ri.n = rfl(ri.n);
        ri.h = rfl(ri.h);
        ri.l = rfl(ri.l);
        ri.wr = rfl64(ri.wr);
        if (ri.status < 0) {
            atomicCAS(&a.s.sched->error, 0, ri.status == -1 ? KP_ERANGE : KP_EOVERFLOW);
            break;
        }
        if (ri.status != 1) continue;
        if (ri.n > 64u * W) {
            if (ln == 0) a.deferred[atomicAdd(a.ndeferred, 1u)] = slot;
            continue;
        }
        RootDesc d{};
        d.addr = reinterpret_cast<uint64_t>(arena);
        d.n = ri.n;
        d.h = ri.h;
        d.l = ri.l;
        d.slot = gw;
        d.job = job;
        d.k = uint16_t(J.k);
        d.q = uint16_t(J.q);
        const uint64_t words = 3ull * ri.n * W + uint64_t(ri.l) * W + (ri.n + 1) / 2;
        for (uint64_t t = ln; t < words; t += 64) arena[t] = 0;
        fence_agent();
        build_root<W>(a.b, d, J, gw, ws.cnt, s_row[wv]);
        if (ln == 0) a.desc[gw] = d;
        fence_agent();
        if (wk.run_single(gw, a.budget)) {
            const uint64_t w = ri.wr + rfl64(a.b.wr_extra[gw]);
            const uint64_t n = ri.n;
            if (ln == 0) {
                JobAcc *A = a.s.acc + job;
                atomicAdd(&A->roots, 1ull);
                atomicAdd(&A->adj, (unsigned long long)(4 * w));
                atomicAdd(&A->dalg, (unsigned long long)(4 * w + (n * n + 7) / 8 + n * n));
                atomicMax(&A->max_n, ri.n);
                atomicMax(&A->max_l, ri.l);
            }
        } else {
            if (wk.err) break;
            if (ln == 0) a.deferred[atomicAdd(a.ndeferred, 1u)] = slot;
        }
    }
}

// fills n words with v (a kernel rather than hipMemsetAsync, so that the
// launch sequence stays capturable as a graph and needs no runtime call that
// may allocate)
#ifndef KP_ONE_WORD_SEARCH_ONLY
__global__ void fill_u64_kernel(uint64_t *p, uint32_t n, uint64_t v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
#endif


// slot j of every shard starts with sequence number j (free for ticket j)
#ifndef KP_ONE_WORD_SEARCH_ONLY
__global__ void iota_kernel(uint64_t *seq, uint32_t n, uint32_t per_shard)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) seq[i] = i % per_shard;
}
#endif


} // namespace kpd
