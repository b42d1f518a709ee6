// kp_device.hip -- MI355X (gfx950) kernels and device driver of libkplex.
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
    uint64_t plexes, ha, hb, tasks;   // accumulators of the current job (see Worker::tk)
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
        const uint64_t cnt = rfl64(cold->plexes), ha = rfl64(cold->ha), hb = rfl64(cold->hb);
        const uint64_t tasks = rfl64(cold->tasks);
        if (job != 0xffffffffu && (cnt | tasks) && lane0()) {
            JobAcc *J = a.acc + job;
            atomicAdd(&J->count, (unsigned long long)cnt);
            atomicAdd(&J->ha, (unsigned long long)ha);
            atomicAdd(&J->hb, (unsigned long long)hb);
            atomicAdd(&J->tasks, (unsigned long long)tasks);
            atomicAdd(&J->alg, (unsigned long long)alg);
        }
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
                const uint64_t o = orig_of(j);
                sa += mix64(o ^ kp::HASH_SA);
                sb += mix64(o ^ kp::HASH_SB);
            }
        sa = rfl64(wsum64(sa));
        sb = rfl64(wsum64(sb));
        cold->plexes = rfl64(cold->plexes) + 1;
        cold->ha = rfl64(cold->ha) + mix64(sa + uint64_t(s) * kp::HASH_LA);
        cold->hb = rfl64(cold->hb) + mix64(sb + uint64_t(s) * kp::HASH_LB);
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
                            const uint64_t OK = __ballot(dP2 + k >= int(ps2) + 1 && (Sat & ~A) == 0);
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
                if (rdl64<W>(xs, 0) != meta) return false;
                rn--;
                tix = tix == 0 ? cap - 1 : tix - 1;
                P = rdl64<W>(xs, 1);
                C = rdl64<W>(xs, 2);
                X = rdl64<W>(xs, 3);
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
                } else if constexpr (W >= 2) {
                    // wide frames: the child is built in place (soc-epinions,
                    // mostly two-word frames: 188 -> 160 ms); one-word frames
                    // keep the copy, which the register allocator handles
                    // better (75k window: 303 vs 309 ms in place)
                    have = (f.type == FT_KSEARCH) ? ksearch(f, f) : branch(f, f);
                } else if (!LISTS && f.type == FT_BRANCH) {
                    // counts its own frames (nfr) and pops its own stack
                    have = fast_run(f, nfr, burst);
                    if (err || nfr == burst) break;
                    if (!have) {
                        if (depth() == 0) break;
                        pop_top(f);
                        have = true;
                    }
                    continue;
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

// Waves per SIMD (list-free kernels, compiler report): one word 7 (68
// VGPRs; asked for at least 4), two to four words 4 (104-128 VGPRs), six
// words 2 (189), eight words 2 when asked for (235, no scratch; 1 otherwise,
// with AGPR spills); sixteen words 1 with 724 bytes of scratch per lane.
template <int W, int LW, bool LISTS>
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
__global__ void fill_u64_kernel(uint64_t *p, uint32_t n, uint64_t v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// slot j of every shard starts with sequence number j (free for ticket j)
__global__ void iota_kernel(uint64_t *seq, uint32_t n, uint32_t per_shard)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) seq[i] = i % per_shard;
}

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
        if (pending < uint64_t(nb) * uint64_t(E.cus) * 4u) nb = std::min(nb, W >= 2 ? 2 : 4);
    }
    const uint32_t nwg = uint32_t(nb) * uint32_t(E.cus);
    SearchArgs sa{};
    sa.nwaves = nwg * 4;
    sa.seed_chunk = kSeedChunk;
    if (const char *e = std::getenv("KP_SEED_CHUNK")) sa.seed_chunk = std::min<uint32_t>(kSeedChunk, std::max(1, std::atoi(e)));
    sa.qcap = T.queue_slots ? std::max<uint32_t>(16, T.queue_slots) : std::max<uint32_t>(256, sa.nwaves * 4 / NSUBQ);
    sa.ring_cap = ring;
    sa.back_cap = back;
    sa.donate_period = T.donate_period ? T.donate_period : 256u;
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
        float t_best = 0, t = 0;
        if ((rc = E.time_class(c, o, &t_best))) return rc;   // warm-up
        if ((rc = E.time_class(c, o, &t_best))) return rc;
        auto trial = [&](ClassTune cand) -> int {
            g_tune[c] = cand;
            int r = E.time_class(c, o, &t);
            g_tune[c] = best;
            if (r) return r;
            if (t < 0.97f * t_best) {
                best = cand;
                t_best = t;
                g_tune[c] = best;
            }
            return KP_OK;
        };
        const int nbmax = std::max(1, E.cls[c].blocks_per_cu);
        for (int nb = nbmax - 1; nb >= 1; nb--) {
            ClassTune x = best;
            x.blocks_per_cu = nb;
            if ((rc = trial(x))) return rc;
        }
        for (uint32_t dp : {32u, 64u, 128u, 512u, 1024u}) {
            ClassTune x = best;
            x.donate_period = dp;
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
