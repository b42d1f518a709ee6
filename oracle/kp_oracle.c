/*
 * kp_oracle.c -- CPU oracle (TEST INFRASTRUCTURE, never the product).
 *
 * Plain C11 + OpenMP restatement of the reference's maximal k-plex path.
 * Each function names the reference file:line it follows.  The per-root
 * search follows the reference's list/array formulation (labels, neiInG,
 * neiInP, local neighbour lists, pair-class matrix) so that it is an
 * independent implementation from the product's bitset formulation.
 */
#define _GNU_SOURCE
#include "kp_oracle.h"
#include <limits.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------ */
/* hashing: SplitMix64 finalizer (Steele, Lea & Flood 2014)            */
/* ------------------------------------------------------------------ */
uint64_t kpo_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
/* Per-plex hash of the vertex set S (original ids):
 *   a(S) = mix64( sum_{v in S} mix64(id_v ^ SA) + |S| * LA )
 *   b(S) = mix64( sum_{v in S} mix64(id_v ^ SB) + |S| * LB )
 * and the result hash is (sum_S a(S), sum_S b(S)) mod 2^64.  Both levels are
 * sums, so the value does not depend on enumeration order. */
#define KP_SA 0x243F6A8885A308D3ULL
#define KP_LA 0x13198A2E03707344ULL
#define KP_SB 0xA4093822299F31D0ULL
#define KP_LB 0x082EFA98EC4E6C89ULL

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* ------------------------------------------------------------------ */
/* graph I/O  (inc/graph.h:37-65: u32 n, u32 m, u32 deg[n], u32 nbrs[m]) */
/* ------------------------------------------------------------------ */
void kpo_graph_free(kpo_graph *g)
{
    if (!g) return;
    free(g->off); free(g->nbr); free(g->deg);
    memset(g, 0, sizeof(*g));
}

static int graph_alloc(kpo_graph *g, uint32_t n, uint64_t m)
{
    g->n = n; g->m = m;
    g->off = (uint64_t *)malloc((size_t)(n + 1) * sizeof(uint64_t));
    g->deg = (uint32_t *)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
    g->nbr = (uint32_t *)malloc((size_t)(m ? m : 1) * sizeof(uint32_t));
    if (!g->off || !g->deg || !g->nbr) { kpo_graph_free(g); return -12; }
    return 0;
}

int kpo_load_bin(const char *path, kpo_graph *g)
{
    memset(g, 0, sizeof(*g));
    FILE *f = fopen(path, "rb");
    if (!f) return -2;
    uint32_t hdr[2];
    if (fread(hdr, 4, 2, f) != 2) { fclose(f); return -5; }
    if (graph_alloc(g, hdr[0], hdr[1])) { fclose(f); return -12; }
    if (fread(g->deg, 4, g->n, f) != g->n) { fclose(f); kpo_graph_free(g); return -5; }
    g->off[0] = 0;
    for (uint32_t i = 0; i < g->n; i++) g->off[i + 1] = g->off[i] + g->deg[i];
    if (g->off[g->n] != g->m) { fclose(f); kpo_graph_free(g); return -22; }
    if (fread(g->nbr, 4, g->m, f) != g->m) { fclose(f); kpo_graph_free(g); return -5; }
    fclose(f);
    return 0;
}

int kpo_graph_from_csr(uint32_t n, const uint32_t *deg, const uint32_t *nbr, kpo_graph *g)
{
    memset(g, 0, sizeof(*g));
    uint64_t m = 0;
    for (uint32_t i = 0; i < n; i++) m += deg[i];
    if (graph_alloc(g, n, m)) return -12;
    g->off[0] = 0;
    for (uint32_t i = 0; i < n; i++) { g->deg[i] = deg[i]; g->off[i + 1] = g->off[i] + deg[i]; }
    if (m) memcpy(g->nbr, nbr, m * 4);
    return 0;
}

/* ------------------------------------------------------------------ */
/* (q-k)-core peel with stable relabel  (inc/host_funcs.h:16-122)      */
/* Vertices with degree < bd are removed, cascading; survivors are     */
/* renumbered by rank in the old numbering (sequence::pack is stable,  */
/* inc/sequence.h:475-490) and neighbour order is preserved.           */
/* ------------------------------------------------------------------ */
static int peel(const kpo_graph *g, int bd, kpo_graph *out, uint32_t *old_of_new /* >= n */,
                uint32_t *pn_out)
{
    uint32_t n = g->n;
    int32_t *res = (int32_t *)malloc((size_t)(n ? n : 1) * 4);
    uint8_t *alive = (uint8_t *)malloc(n ? n : 1);
    uint32_t *Q = (uint32_t *)malloc((size_t)(n ? n : 1) * 4);
    uint32_t *map = (uint32_t *)malloc((size_t)(n ? n : 1) * 4);
    if (!res || !alive || !Q || !map) { free(res); free(alive); free(Q); free(map); return -12; }
    size_t qh = 0, qt = 0;
    for (uint32_t i = 0; i < n; i++) {
        res[i] = (int32_t)g->deg[i];
        if ((long)g->deg[i] < bd) { alive[i] = 0; Q[qt++] = i; }
        else alive[i] = 1;
    }
    while (qh < qt) {
        uint32_t e = Q[qh++];
        for (uint64_t p = g->off[e]; p < g->off[e + 1]; p++) {
            uint32_t v = g->nbr[p];
            if (alive[v]) {
                int32_t old = res[v]--;
                if (old == bd) { alive[v] = 0; Q[qt++] = v; }
            }
        }
    }
    uint32_t pn = 0;
    for (uint32_t i = 0; i < n; i++) if (alive[i]) { map[i] = pn; old_of_new[pn] = i; pn++; }
    uint64_t pm = 0;
    for (uint32_t i = 0; i < pn; i++) {
        uint32_t o = old_of_new[i];
        for (uint64_t p = g->off[o]; p < g->off[o + 1]; p++) pm += alive[g->nbr[p]];
    }
    int rc = graph_alloc(out, pn, pm);
    if (!rc) {
        uint64_t w = 0;
        out->off[0] = 0;
        for (uint32_t i = 0; i < pn; i++) {
            uint32_t o = old_of_new[i];
            for (uint64_t p = g->off[o]; p < g->off[o + 1]; p++) {
                uint32_t v = g->nbr[p];
                if (alive[v]) out->nbr[w++] = map[v];
            }
            out->off[i + 1] = w;
            out->deg[i] = (uint32_t)(w - out->off[i]);
        }
    }
    *pn_out = pn;
    free(res); free(alive); free(Q); free(map);
    return rc;
}

/* ------------------------------------------------------------------ */
/* (q-2k)-truss prefilter (inc/host_funcs.h:761-825 driver,            */
/* :463-559 fast_truss_peeling, :405-415 find_pos_sorted).             */
/* Transcribed with its exact behaviour: supports are per directed     */
/* entry, only the u->w and v->w copies are decremented, the stack is  */
/* LIFO, and an edge may be popped more than once.                     */
/* ------------------------------------------------------------------ */
static long find_pos(const kpo_graph *g, uint32_t u, uint32_t v)
{
    uint64_t lo = g->off[u], hi = g->off[u + 1];
    while (lo < hi) {
        uint64_t mid = lo + ((hi - lo) >> 1);
        if (g->nbr[mid] < v) lo = mid + 1; else hi = mid;
    }
    if (lo == g->off[u + 1] || g->nbr[lo] != v) return -1;
    return (long)lo;
}

static int truss_prune(kpo_graph *g, int q, int k)
{
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const int thr = q - 2 * k;
    int32_t *tri = (int32_t *)malloc((size_t)(m ? m : 1) * 4);
    uint8_t *dead = (uint8_t *)calloc(m ? m : 1, 1);
    if (!tri || !dead) { free(tri); free(dead); return -12; }
    #pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < (long)n; i++) {
        for (uint64_t p = g->off[i]; p < g->off[i + 1]; p++) {
            uint32_t v = g->nbr[p];
            uint64_t iu = g->off[i], eu = g->off[i + 1], iv = g->off[v], ev = g->off[v + 1];
            int common = 0;
            while (iu < eu && iv < ev) {
                uint32_t a = g->nbr[iu], b = g->nbr[iv];
                if (a == b) { common++; iu++; iv++; }
                else if (a < b) iu++;
                else iv++;
            }
            tri[p] = common;
        }
    }
    /* seeds in (i ascending, neighbour order) -- inc/host_funcs.h:800-819 */
    size_t cap = 1024, sz = 0;
    uint32_t (*Qe)[2] = malloc(cap * sizeof(*Qe));
    if (!Qe) { free(tri); free(dead); return -12; }
#define QPUSH(a, b) do { if (sz == cap) { cap *= 2; void *nq = realloc(Qe, cap * sizeof(*Qe)); \
        if (!nq) { free(Qe); free(tri); free(dead); return -12; } Qe = nq; } Qe[sz][0] = (a); Qe[sz][1] = (b); sz++; } while (0)
    for (uint32_t i = 0; i < n; i++)
        for (uint64_t p = g->off[i]; p < g->off[i + 1]; p++) {
            uint32_t v = g->nbr[p];
            if (i < v && tri[p] < thr) QPUSH(i, v);
        }
    while (sz) {
        sz--;
        uint32_t u = Qe[sz][0], v = Qe[sz][1];
        long puv = find_pos(g, u, v), pvu = find_pos(g, v, u);
        if (puv >= 0 && !dead[puv]) { dead[puv] = 1; g->deg[u]--; }
        if (pvu >= 0 && !dead[pvu]) { dead[pvu] = 1; g->deg[v]--; }
        uint64_t iu = g->off[u], eu = g->off[u + 1], iv = g->off[v], ev = g->off[v + 1];
        while (iu < eu && iv < ev) {
            while (iu < eu && dead[iu]) ++iu;
            while (iv < ev && dead[iv]) ++iv;
            if (iu >= eu || iv >= ev) break;
            uint32_t a = g->nbr[iu], b = g->nbr[iv];
            if (a == b) {
                uint32_t w = a;
                if (!dead[iu]) {
                    int prev = tri[iu];
                    if (prev > 0) { tri[iu] = prev - 1; if (prev == thr) QPUSH(u, w); }
                }
                if (!dead[iv]) {
                    int prev = tri[iv];
                    if (prev > 0) { tri[iv] = prev - 1; if (prev == thr) QPUSH(v, w); }
                }
                iu++; iv++;
            } else if (a < b) iu++;
            else iv++;
        }
    }
#undef QPUSH
    uint64_t w = 0;
    for (uint32_t u = 0; u < n; u++) {
        uint64_t b = g->off[u], e = g->off[u + 1];
        g->off[u] = w;
        for (uint64_t i = b; i < e; i++) if (!dead[i]) g->nbr[w++] = g->nbr[i];
    }
    g->off[n] = w;
    g->m = w;
    free(Qe); free(tri); free(dead);
    return 0;
}

/* ------------------------------------------------------------------ */
/* degeneracy order: ListLinearHeap (inc/degeneracy.h:23-203) driven   */
/* as in inc/host_funcs.h:829-851.  Buckets are LIFO.                  */
/* ------------------------------------------------------------------ */
static int degeneracy(const kpo_graph *g, int32_t *dseq, int32_t *dpos)
{
    const int n = (int)g->n;
    if (n == 0) return 0;
    const int key_cap = n - 1;
    int *keys = malloc(n * sizeof(int)), *pres = malloc(n * sizeof(int)), *nexts = malloc(n * sizeof(int));
    int *heads = malloc((key_cap + 1) * sizeof(int));
    if (!keys || !pres || !nexts || !heads) { free(keys); free(pres); free(nexts); free(heads); return -12; }
    int min_key = key_cap, max_key = 0;
    for (int i = 0; i <= key_cap; i++) heads[i] = n;
    for (int i = 0; i < n; i++) dpos[i] = INT_MAX;
#define H_INSERT(id, key) do { int _id = (id), _k = (key); keys[_id] = _k; pres[_id] = n; nexts[_id] = heads[_k]; \
        if (heads[_k] != n) pres[heads[_k]] = _id; heads[_k] = _id; if (_k < min_key) min_key = _k; if (_k > max_key) max_key = _k; } while (0)
#define H_REMOVE(id) do { int _id = (id); if (pres[_id] == n) { heads[keys[_id]] = nexts[_id]; if (nexts[_id] != n) pres[nexts[_id]] = n; } \
        else { int _p = pres[_id]; nexts[_p] = nexts[_id]; if (nexts[_id] != n) pres[nexts[_id]] = _p; } } while (0)
    for (int i = 0; i < n; i++) H_INSERT(i, (int)g->deg[i]);
    for (int i = 0; i < n; i++) {
        while (heads[min_key] == n) ++min_key;         /* tighten() */
        int u = heads[min_key];
        heads[min_key] = nexts[u];
        if (heads[min_key] != n) pres[heads[min_key]] = n;
        dpos[u] = i;
        dseq[i] = u;
        for (uint64_t p = g->off[u]; p < g->off[u] + g->deg[u]; p++) {
            int nei = (int)g->nbr[p];
            if (dpos[nei] == INT_MAX) {
                int nk = keys[nei] - 1;
                H_REMOVE(nei);
                H_INSERT(nei, nk);
            }
        }
    }
#undef H_INSERT
#undef H_REMOVE
    (void)max_key;
    free(keys); free(pres); free(nexts); free(heads);
    return 0;
}

void kpo_prep_free(kpo_prep *p)
{
    if (!p) return;
    kpo_graph_free(&p->g);
    free(p->orig); free(p->dseq); free(p->dpos);
    memset(p, 0, sizeof(*p));
}

/* decomposableSearch's host half (inc/host_funcs.h:722-851) */
int kpo_preprocess(const kpo_graph *g, int k, int q, int truss, kpo_prep *p)
{
    memset(p, 0, sizeof(*p));
    const int bd = q - k;
    uint32_t *o1 = malloc((size_t)(g->n ? g->n : 1) * 4);
    if (!o1) return -12;
    uint32_t pn;
    int rc = peel(g, bd, &p->g, o1, &pn);
    if (rc) { free(o1); return rc; }
    if (truss) {
        rc = truss_prune(&p->g, q, k);
        if (rc) { free(o1); kpo_prep_free(p); return rc; }
        kpo_graph tg = p->g;
        /* after compaction the CSR is consistent with deg[] */
        uint32_t *o2 = malloc((size_t)(tg.n ? tg.n : 1) * 4);
        kpo_graph g2;
        uint32_t pn2;
        rc = o2 ? peel(&tg, bd, &g2, o2, &pn2) : -12;
        if (rc) { free(o1); free(o2); kpo_prep_free(p); return rc; }
        for (uint32_t i = 0; i < pn2; i++) o2[i] = o1[o2[i]];
        free(o1);
        o1 = o2;
        kpo_graph_free(&p->g);
        p->g = g2;
        pn = pn2;
    }
    p->orig = o1;
    p->dseq = malloc((size_t)(pn ? pn : 1) * 4);
    p->dpos = malloc((size_t)(pn ? pn : 1) * 4);
    if (!p->dseq || !p->dpos) { kpo_prep_free(p); return -12; }
    return degeneracy(&p->g, p->dseq, p->dpos);
}

/* ------------------------------------------------------------------ */
/* per-thread workspace                                                */
/* ------------------------------------------------------------------ */
typedef struct {
    /* stamps over the preprocessed vertex set */
    uint32_t *vis, *inB1, *idxB, *idxL;
    uint32_t *bpos, *lpos;   /* peeled id -> block / left index (valid where idxB/idxL match) */
    uint32_t stamp;
    uint32_t *B, *L, *B1c, *Lc, *B2, *L2, *cnt;
    size_t pn;
} dws_t;

static int dws_init(dws_t *w, size_t pn)
{
    memset(w, 0, sizeof(*w));
    size_t s = pn ? pn : 1;
    w->pn = pn;
    w->vis = calloc(s, 4); w->inB1 = calloc(s, 4); w->idxB = calloc(s, 4); w->idxL = calloc(s, 4);
    w->B = malloc(s * 4); w->L = malloc(s * 4); w->B1c = malloc(s * 4); w->Lc = malloc(s * 4);
    w->B2 = malloc(s * 4); w->L2 = malloc(s * 4); w->cnt = malloc(s * 4);
    w->bpos = malloc(s * 4); w->lpos = malloc(s * 4);
    return (w->bpos && w->lpos && w->vis && w->inB1 && w->idxB && w->idxL && w->B && w->L && w->B1c && w->Lc && w->B2 && w->L2 && w->cnt) ? 0 : -12;
}
static void dws_free(dws_t *w)
{
    free(w->vis); free(w->inB1); free(w->idxB); free(w->idxL); free(w->B); free(w->L);
    free(w->B1c); free(w->Lc); free(w->B2); free(w->L2); free(w->cnt);
    free(w->bpos); free(w->lpos);
}
static uint32_t next_stamp(dws_t *w)
{
    if (++w->stamp == 0) {
        memset(w->vis, 0, w->pn * 4); memset(w->inB1, 0, w->pn * 4);
        memset(w->idxB, 0, w->pn * 4); memset(w->idxL, 0, w->pn * 4);
        w->stamp = 1;
    }
    return w->stamp;
}

static int cmp_u32(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

static int count_in_B1(const kpo_graph *g, const dws_t *w, uint32_t st, uint32_t u)
{
    int c = 0;
    for (uint64_t p = g->off[u]; p < g->off[u + 1]; p++) c += (w->inB1[g->nbr[p]] == st);
    return c;
}

/* decompose (src/device_funcs.cu:28-298).  Returns 1 valid, 0 invalid.  */
static int decompose_root(const kpo_graph *g, const int32_t *dpos, const int32_t *dseq, long pos,
                          int k, int q, dws_t *w, uint32_t *n_out, uint32_t *hop_out, uint32_t *l_out)
{
    const int lb2k = q - 2 * k, bd = q - k;
    const uint32_t r = (uint32_t)dseq[pos];
    const uint32_t st = next_stamp(w);
    uint32_t n1 = 0, nl = 0;
    *n_out = 0; *hop_out = 0; *l_out = 0;
    w->vis[r] = st;
    for (uint64_t p = g->off[r]; p < g->off[r + 1]; p++) {
        uint32_t v = g->nbr[p];
        w->vis[v] = st;
        if (dpos[v] > pos) w->B1c[n1++] = v; else w->Lc[nl++] = v;
    }
    /* fixpoint filter of the later 1-hop (:139-171) */
    uint32_t nb1 = n1;
    memcpy(w->B, w->B1c, n1 * 4);          /* B holds B1 (without the root) while filtering */
    for (uint32_t i = 0; i < nb1; i++) w->inB1[w->B[i]] = st;
    for (;;) {
        if ((long)nb1 < (long)bd) return 0;   /* |B| - 1 < q - k */
        for (uint32_t i = 0; i < nb1; i++) w->cnt[i] = (uint32_t)count_in_B1(g, w, st, w->B[i]);
        uint32_t wr = 0;
        for (uint32_t i = 0; i < nb1; i++) {
            if ((long)w->cnt[i] >= lb2k) w->B[wr++] = w->B[i];
            else { w->inB1[w->B[i]] = 0; w->vis[w->B[i]] = 0; }
        }
        if (wr == nb1) break;
        nb1 = wr;
    }
    const uint32_t hop = nb1 + 1;
    /* earlier 1-hop filter (:177-209) */
    uint32_t nlk = 0;
    for (uint32_t i = 0; i < nl; i++) {
        uint32_t u = w->Lc[i];
        if (count_in_B1(g, w, st, u) >= lb2k + 1) w->L[nlk++] = u;
        else w->vis[u] = 0;
    }
    /* 2-hop (:211-273): scan all later neighbours of the root */
    uint32_t nb2 = 0, nl2 = 0;
    for (uint32_t i = 0; i < n1; i++) {
        uint32_t nei = w->B1c[i];
        for (uint64_t p = g->off[nei]; p < g->off[nei + 1]; p++) {
            uint32_t th = g->nbr[p];
            if (w->vis[th] == st) continue;
            const int later = dpos[th] > pos;
            const int thr = lb2k + (later ? 2 : 3);
            long mn = (long)g->deg[th] < (long)(hop - 1) ? (long)g->deg[th] : (long)(hop - 1);
            if (mn < thr) continue;
            int c = count_in_B1(g, w, st, th);
            w->vis[th] = st;
            if (later) { if (c >= lb2k + 2) w->B2[nb2++] = th; }
            else { if (c >= lb2k + 3) w->L2[nl2++] = th; }
        }
    }
    qsort(w->B2, nb2, 4, cmp_u32);
    qsort(w->L2, nl2, 4, cmp_u32);
    /* canonical layout: B = [r, B1 (ascending), B2 (ascending)] */
    memmove(w->B + 1, w->B, nb1 * 4);
    w->B[0] = r;
    memcpy(w->B + hop, w->B2, nb2 * 4);
    memcpy(w->L + nlk, w->L2, nl2 * 4);
    *n_out = hop + nb2;
    *hop_out = hop;
    *l_out = nlk + nl2;
    return (long)(*n_out) >= q ? 1 : 0;
}

/* ------------------------------------------------------------------ */
/* local subgraph (calculateDegrees :300-351, fillNeighbors :353-460,  */
/* commonEle/buildCommonMtx :1727-1853)                                */
/* ------------------------------------------------------------------ */
enum { LAB_P = 0, LAB_C = 1, LAB_X = 2, LAB_U = 3, LAB_H = 5 };
enum { UNLINK2LESS = 0, LINK2LESS = 1, UNLINK2EQUAL = 2, LINK2EQUAL = 3, UNLINK2MORE = 4, LINK2MORE = 5 };

typedef struct {
    int n, h, l, k, q;
    uint8_t *adj;          /* n*n */
    uint8_t *cls;          /* n*n */
    int *noff, *nbr;       /* local CSR, sorted */
    int *dhop;
    int *loff, *lnb;       /* left CSR, sorted left indices */
    const uint32_t *B;     /* peeled ids */
    const uint32_t *orig;  /* peeled -> original */
} sub_t;

static int max_i(int a, int b) { return a > b ? a : b; }

static int build_sub(const kpo_graph *g, dws_t *w, uint32_t n, uint32_t hop, uint32_t l,
                     int k, int q, sub_t *s)
{
    memset(s, 0, sizeof(*s));
    s->n = (int)n; s->h = (int)hop; s->l = (int)l; s->k = k; s->q = q;
    s->adj = calloc((size_t)n * n, 1);
    s->cls = calloc((size_t)n * n, 1);
    s->noff = malloc((n + 1) * sizeof(int));
    s->dhop = malloc(n * sizeof(int));
    s->loff = malloc((n + 1) * sizeof(int));
    if (!s->adj || !s->cls || !s->noff || !s->dhop || !s->loff) return -12;
    const uint32_t st = next_stamp(w);
    for (uint32_t i = 0; i < n; i++) w->idxB[w->B[i]] = st, w->bpos[w->B[i]] = i;
    for (uint32_t j = 0; j < l; j++) w->idxL[w->L[j]] = st, w->lpos[w->L[j]] = j;
    size_t tot = 0, ltot = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t u = w->B[i];
        for (uint64_t p = g->off[u]; p < g->off[u + 1]; p++) {
            uint32_t v = g->nbr[p];
            if (w->idxB[v] == st) tot++;
            else if (w->idxL[v] == st) ltot++;
        }
    }
    s->nbr = malloc((tot ? tot : 1) * sizeof(int));
    s->lnb = malloc((ltot ? ltot : 1) * sizeof(int));
    if (!s->nbr || !s->lnb) return -12;
    int o = 0, lo = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t u = w->B[i];
        s->noff[i] = o; s->loff[i] = lo;
        for (uint64_t p = g->off[u]; p < g->off[u + 1]; p++) {
            uint32_t v = g->nbr[p];
            if (w->idxB[v] == st) { int j = (int)w->bpos[v]; s->nbr[o++] = j; s->adj[(size_t)i * n + j] = 1; }
            else if (w->idxL[v] == st) s->lnb[lo++] = (int)w->lpos[v];
        }
        qsort(s->nbr + s->noff[i], o - s->noff[i], sizeof(int), (int (*)(const void *, const void *))cmp_u32);
        qsort(s->lnb + s->loff[i], lo - s->loff[i], sizeof(int), (int (*)(const void *, const void *))cmp_u32);
        int dh = 0;
        for (int t = s->noff[i]; t < o; t++) dh += (s->nbr[t] < (int)hop);
        s->dhop[i] = dh;
    }
    s->noff[n] = o; s->loff[n] = lo;
    /* pair classes (buildCommonMtx :1768-1830) */
    const int thPP1 = q - k - 2 * max_i(k - 2, 0), thPP2 = q - k - 2 * max_i(k - 3, 0);
    const int thPC1 = q - 2 * k - max_i(k - 2, 0), thPC2 = q - k - 1 - max_i(k - 2, 0) - max_i(k - 3, 0);
    const int thCC1 = q - 2 * k - (k - 1), thCC2 = q - 2 * k + 2 - (k - 1);
    for (int i = 0; i < (int)n; i++) {
        for (int j = 0; j < i; j++) {
            int t1, t2;
            if (i < (int)hop) { t1 = thCC1; t2 = thCC2; }
            else if (j < (int)hop) { t1 = thPC1; t2 = thPC2; }
            else {
                if (k == 2) { uint8_t c = s->adj[(size_t)i * n + j]; s->cls[(size_t)i * n + j] = c; s->cls[(size_t)j * n + i] = c; continue; }
                t1 = thPP1; t2 = thPP2;
            }
            /* commonEle: |N(i) ∩ N(j) ∩ B[0,hop)| over the first degreeHop entries */
            int a = s->noff[i], ea = s->noff[i] + s->dhop[i], b = s->noff[j], eb = s->noff[j] + s->dhop[j], c = 0;
            while (a < ea && b < eb) {
                if (s->nbr[a] < s->nbr[b]) a++;
                else if (s->nbr[a] > s->nbr[b]) b++;
                else { c++; a++; b++; }
            }
            uint8_t cl;
            if (s->adj[(size_t)i * n + j]) cl = c > t1 ? LINK2MORE : (c == t1 ? LINK2EQUAL : LINK2LESS);
            else cl = c > t2 ? UNLINK2MORE : (c == t2 ? UNLINK2EQUAL : UNLINK2LESS);
            s->cls[(size_t)i * n + j] = cl;
            s->cls[(size_t)j * n + i] = cl;
        }
    }
    return 0;
}

static void free_sub(sub_t *s)
{
    free(s->adj); free(s->cls); free(s->noff); free(s->nbr); free(s->dhop); free(s->loff); free(s->lnb);
}

/* ------------------------------------------------------------------ */
/* search context                                                      */
/* ------------------------------------------------------------------ */
typedef struct {
    uint64_t *ids;   /* (len, id...) records */
    size_t sz, cap;
    uint64_t n;
} plist_t;

typedef struct {
    const sub_t *s;
    int n, k, q;
    /* k-search state */
    int *P; int PS;
    uint8_t *c1, *c2;        /* membership flags */
    int C1S, C2S;
    int *X; int XS;
    int *nG, *nP;            /* neiInG / neiInP (int for the k-search) */
    int *drops; size_t dsz, dcap;
    /* Branch recursion buffers, one level per depth */
    uint8_t *lab; uint16_t *bG, *bP; int levels;
    int *PL, *CL, *XL;       /* per-depth P/C/X lists: levels * n */
    /* maximality scratch */
    uint8_t *uni; int *commons, *commons2, *sat;
    /* results */
    uint64_t count, tasks, ha, hb;
    uint32_t flags;
    plist_t *list;
    uint32_t *tmp_ids;
} ctx_t;

static inline int adjb(const ctx_t *c, int a, int b) { return c->s->adj[(size_t)a * c->n + b]; }
static inline int clsb(const ctx_t *c, int a, int b) { return c->s->cls[(size_t)a * c->n + b]; }
static inline void addG(ctx_t *c, int v, int *arr) { const sub_t *s = c->s; for (int t = s->noff[v]; t < s->noff[v + 1]; t++) arr[s->nbr[t]]++; }
static inline void subG(ctx_t *c, int v, int *arr) { const sub_t *s = c->s; for (int t = s->noff[v]; t < s->noff[v + 1]; t++) arr[s->nbr[t]]--; }

static int drops_push(ctx_t *c, int v)
{
    if (c->dsz == c->dcap) {
        size_t nc = c->dcap ? c->dcap * 2 : 1024;
        int *nd = realloc(c->drops, nc * sizeof(int));
        if (!nd) return -12;
        c->drops = nd; c->dcap = nc;
    }
    c->drops[c->dsz++] = v;
    return 0;
}

static void record_plex(ctx_t *c, const int *S1, int n1, const int *S2, int n2)
{
    const sub_t *s = c->s;
    int len = n1 + n2;
    uint64_t sa = 0, sb = 0;
    for (int i = 0; i < len; i++) {
        int v = i < n1 ? S1[i] : S2[i - n1];
        uint32_t id = s->orig[s->B[v]];
        c->tmp_ids[i] = id;
        sa += kpo_mix64((uint64_t)id ^ KP_SA);
        sb += kpo_mix64((uint64_t)id ^ KP_SB);
    }
    c->count++;
    c->ha += kpo_mix64(sa + (uint64_t)len * KP_LA);
    c->hb += kpo_mix64(sb + (uint64_t)len * KP_LB);
    if (c->flags & KPO_LIST) {
        plist_t *L = c->list;
        qsort(c->tmp_ids, len, 4, cmp_u32);
        if (L->sz + len + 1 > L->cap) {
            size_t nc = L->cap ? L->cap * 2 : 4096;
            while (nc < L->sz + len + 1) nc *= 2;
            uint64_t *ni = realloc(L->ids, nc * 8);
            if (!ni) return;   /* listing is best effort; count/hash are unaffected */
            L->ids = ni; L->cap = nc;
        }
        L->ids[L->sz++] = (uint64_t)len;
        for (int i = 0; i < len; i++) L->ids[L->sz++] = c->tmp_ids[i];
        L->n++;
    }
}

/* binary search of u in the sorted left list of block vertex v */
static int in_left(const sub_t *s, int v, int u)
{
    int lo = s->loff[v], hi = s->loff[v + 1];
    while (lo < hi) { int mid = (lo + hi) >> 1; if (s->lnb[mid] < u) lo = mid + 1; else hi = mid; }
    return lo < s->loff[v + 1] && s->lnb[lo] == u;
}

/* canGlobalAdd / canGlobalAdd2 (:643-679) */
static int can_global_add(const ctx_t *c, int u, const int *S, int len)
{
    int cnt = 0;
    for (int i = 0; i < len; i++) cnt += in_left(c->s, S[i], u);
    return cnt + c->k >= len + 1;
}

/* isMaximal_opt / isMaximalPC_opt (:703-1000) on S = plex (+ cand), with
 * per-vertex in-set degree deg[] (neiInP for P, neiInG for P+C). */
static int is_maximal(ctx_t *c, const int *S, int len, const uint16_t *deg)
{
    const sub_t *s = c->s;
    const int k = c->k;
    int nsat = 0;
    for (int i = 0; i < len; i++) if (deg[S[i]] + k < len + 1) c->sat[nsat++] = S[i];
    int szcom = 0;
    if (nsat) {
        int v0 = c->sat[0];
        for (int t = s->loff[v0]; t < s->loff[v0 + 1]; t++) c->commons[szcom++] = s->lnb[t];
        for (int i = 1; i < nsat; i++) {   /* intersection (:680-702) */
            int v = c->sat[i], a = 0, b = s->loff[v], eb = s->loff[v + 1], d = 0;
            while (a < szcom && b < eb) {
                if (c->commons[a] == s->lnb[b]) { c->commons[d++] = c->commons[a]; a++; b++; }
                else if (c->commons[a] < s->lnb[b]) a++;
                else b++;
            }
            szcom = d;
        }
        if (nsat == 1) {
            for (int t = 0; t < k && t < len; t++) {
                int u = S[t];
                for (int j = s->loff[u]; j < s->loff[u + 1]; j++) c->uni[s->lnb[j]] = 1;
            }
            int wr = 0;
            for (int i = 0; i < szcom; i++) if (c->uni[c->commons[i]]) c->commons[wr++] = c->commons[i];
            szcom = wr;
            for (int t = 0; t < k && t < len; t++) {
                int u = S[t];
                for (int j = s->loff[u]; j < s->loff[u + 1]; j++) c->uni[s->lnb[j]] = 0;
            }
        }
    } else {
        for (int t = 0; t < k && t < len; t++) {
            int u = S[t];
            for (int j = s->loff[u]; j < s->loff[u + 1]; j++) {
                int v = s->lnb[j];
                if (c->uni[v]) continue;
                c->commons[szcom++] = v;
                c->uni[v] = 1;
            }
        }
        for (int i = 0; i < szcom; i++) c->uni[c->commons[i]] = 0;
    }
    for (int i = 0; i < szcom; i++)
        if (can_global_add(c, c->commons[i], S, len)) return 0;
    return 1;
}

/* ---------------- Branch (BNB :1516-1724) as a DFS ------------------ */
static void bnb(ctx_t *c, int depth, int PS, int CS, int XS);

/* build child at level d+1 from level d's arrays */
static inline uint8_t *LAB(ctx_t *c, int d) { return c->lab + (size_t)d * c->n; }
static inline uint16_t *BG(ctx_t *c, int d) { return c->bG + (size_t)d * c->n; }
static inline uint16_t *BP(ctx_t *c, int d) { return c->bP + (size_t)d * c->n; }

static int grow_levels(ctx_t *c, int need)
{
    if (need <= c->levels) return 0;
    int nl = c->levels ? c->levels : 16;
    while (nl < need) nl *= 2;
    size_t n = (size_t)c->n;
    uint8_t *l = realloc(c->lab, nl * n);
    if (!l) return -12;
    c->lab = l;
    uint16_t *g = realloc(c->bG, nl * n * 2);
    if (!g) return -12;
    c->bG = g;
    uint16_t *p = realloc(c->bP, nl * n * 2);
    if (!p) return -12;
    c->bP = p;
    int *a = realloc(c->PL, nl * n * sizeof(int)); if (!a) return -12; c->PL = a;
    a = realloc(c->CL, nl * n * sizeof(int)); if (!a) return -12; c->CL = a;
    a = realloc(c->XL, nl * n * sizeof(int)); if (!a) return -12; c->XL = a;
    c->levels = nl;
    return 0;
}

static void bnb(ctx_t *c, int d, int PS, int CS, int XS)
{
    const int n = c->n, k = c->k, q = c->q;
    const sub_t *s = c->s;
    uint8_t *lab = LAB(c, d);
    uint16_t *nG = BG(c, d), *nP = BP(c, d);
    int *P = c->PL + (size_t)d * n, *C = c->CL + (size_t)d * n, *X = c->XL + (size_t)d * n;
    /* initializePCX (:1474-1513): lists in ascending local index */
    int ps = 0, cs = 0, xs = 0;
    for (int i = 0; i < n; i++) {
        if (lab[i] == LAB_P) P[ps++] = i;
        else if (lab[i] == LAB_C) C[cs++] = i;
        else if (lab[i] == LAB_X) X[xs++] = i;
    }
    (void)PS; (void)CS; (void)XS;
    if (ps + cs < q) return;
    if (cs == 0) {
        if (xs == 0 && ps >= q && is_maximal(c, P, ps, nP)) record_plex(c, P, ps, NULL, 0);
        return;
    }
    int minP = INT_MAX, pp = -1;
    for (int i = 0; i < ps; i++) if (nG[P[i]] < minP) { minP = nG[P[i]]; pp = P[i]; }
    if (minP + k < max_i(q, ps)) return;
    int pivot = -1;
    if (minP + k < ps + cs) {
        int bg = INT_MAX, bp = INT_MAX;
        for (int i = 0; i < cs; i++) {
            int v = C[i];
            if (adjb(c, v, pp)) continue;
            if (nG[v] < bg || (nG[v] == bg && nP[v] < bp)) { bg = nG[v]; bp = nP[v]; pivot = v; }
        }
        if (pivot < 0) return;   /* unreachable while neiInG is consistent */
    } else {
        int bg = INT_MAX, bp = INT_MAX;
        for (int i = 0; i < cs; i++) {
            int v = C[i];
            if (nG[v] < bg || (nG[v] == bg && nP[v] < bp)) { bg = nG[v]; bp = nP[v]; pivot = v; }
        }
        int mn = minP < bg ? minP : bg;
        if (mn >= ps + cs - k) {
            /* P+C is a k-plex: reject if some x in X extends it (isKplexPC2 :600-622) */
            const int tot = ps + cs;
            for (int i = 0; i < xs; i++) {
                int x = X[i];
                if (nG[x] + k < tot + 1) continue;
                int ok = 1;
                for (int j = 0; j < ps && ok; j++) { int u = P[j]; if (nG[u] + k == tot && !adjb(c, u, x)) ok = 0; }
                for (int j = 0; j < cs && ok; j++) { int u = C[j]; if (nG[u] + k == tot && !adjb(c, u, x)) ok = 0; }
                if (ok) return;
            }
            /* S = P then C (isMaximalPC_opt reads plex[t] then cand[t-PlexSz]) */
            int *S = c->commons2;
            memcpy(S, P, ps * sizeof(int));
            memcpy(S + ps, C, cs * sizeof(int));
            if (is_maximal(c, S, tot, nG)) record_plex(c, P, ps, C, cs);
            return;
        }
    }
    /* branchInCand2 (:1467-1472): exclude child then include child */
    const int v = pivot;
    uint8_t *cl = LAB(c, d + 1);
    uint16_t *cG = BG(c, d + 1), *cP = BP(c, d + 1);
    /* exclude child (:1197-1270) */
    memcpy(cl, lab, n);
    memcpy(cG, nG, n * 2);
    memcpy(cP, nP, n * 2);
    for (int t = s->noff[v]; t < s->noff[v + 1]; t++) cG[s->nbr[t]]--;
    cl[v] = LAB_X;
    c->tasks++;
    bnb(c, d + 1, ps, cs - 1, xs + 1);
    /* include child (:1273-1465), built from level d's (unmodified) arrays */
    memcpy(cl, lab, n);
    memcpy(cG, nG, n * 2);
    memcpy(cP, nP, n * 2);
    cl[v] = LAB_P;
    int nps = ps + 1;
    for (int t = s->noff[v]; t < s->noff[v + 1]; t++) cP[s->nbr[t]]++;
    /* P list with v appended (order irrelevant for the checks below) */
    int *Pn = c->sat;   /* scratch: size n */
    memcpy(Pn, P, ps * sizeof(int));
    Pn[ps] = v;
    int ncs = 0;
    for (int i = 0; i < cs; i++) {
        int u = C[i];
        if (u == v) continue;
        int keep = clsb(c, v, u) >= UNLINK2EQUAL;
        if (keep) {   /* isKplex3 (:586-598) */
            if (cP[u] + k < nps + 1) keep = 0;
            for (int j = 0; j < nps && keep; j++) { int w = Pn[j]; if (cP[w] + k == nps && !adjb(c, w, u)) keep = 0; }
        }
        if (keep) ncs++;
        else { cl[u] = LAB_U; for (int t = s->noff[u]; t < s->noff[u + 1]; t++) cG[s->nbr[t]]--; }
    }
    /* upperBound2 (:1022-1040) */
    for (int j = 0; j < nps; j++) if (cG[Pn[j]] + k < q) return;
    int nxs = 0;
    for (int i = 0; i < xs; i++) {
        int x = X[i];
        int keep = clsb(c, v, x) >= UNLINK2MORE;
        if (keep) {
            if (cP[x] + k < nps + 1) keep = 0;
            for (int j = 0; j < nps && keep; j++) { int w = Pn[j]; if (cP[w] + k == nps && !adjb(c, w, x)) keep = 0; }
        }
        if (keep) nxs++;
        else cl[x] = LAB_U;
    }
    c->tasks++;
    bnb(c, d + 1, nps, ncs, nxs);
}

/* emit a Branch task from the k-search state (kSearch :2005-2082, :2317-2380) */
static void ks_emit(ctx_t *c)
{
    const int n = c->n;
    uint8_t *lab = LAB(c, 0);
    uint16_t *nG = BG(c, 0), *nP = BP(c, 0);
    for (int i = 0; i < n; i++) { lab[i] = LAB_U; nG[i] = (uint16_t)c->nG[i]; nP[i] = (uint16_t)c->nP[i]; }
    for (int i = 0; i < c->PS; i++) lab[c->P[i]] = LAB_P;
    for (int i = 0; i < n; i++) { if (c->c1[i]) lab[i] = LAB_C; else if (c->c2[i]) lab[i] = LAB_H; }
    for (int i = 0; i < c->XS; i++) lab[c->X[i]] = LAB_X;
    c->tasks++;
    bnb(c, 0, c->PS, c->C1S, c->XS);
}

/* upperBoundK2 (:1002-1020) */
static int ks_ub(const ctx_t *c)
{
    for (int i = 0; i < c->PS; i++) if (c->nG[c->P[i]] + c->k < c->q) return 0;
    return 1;
}

/* isKplex2 (:568-584) */
static int ks_iskplex(const ctx_t *c, int v)
{
    if (c->nP[v] + c->k < c->PS + 1) return 0;
    for (int i = 0; i < c->PS; i++) {
        int u = c->P[i];
        if (c->nP[u] + c->k == c->PS && !adjb(c, u, v)) return 0;
    }
    return 1;
}

static void ks_emit0(ctx_t *c)
{
    if (c->PS + c->C1S < c->q) return;
    if (c->PS > 1 && !ks_ub(c)) return;
    ks_emit(c);
}

/* The kSearch state machine (:1936-2443) restated recursively.  A frame with
 * budget r over the ordered C2 either excludes the top vertex u (child with
 * the same budget) or includes it; an include at budget 1 is the state-4
 * emit (C1 filtered by isKplex2, rest of C2 labelled H), otherwise the
 * include continues with budget r-1.  Drops made by updateCand13/23 are
 * undone when the frame returns (recoverCand12/23). */
static int ks_top(const ctx_t *c)
{
    for (int i = c->n - 1; i >= c->s->h; i--) if (c->c2[i]) return i;
    return -1;
}

static void ks_G(ctx_t *c, int r)
{
    const int n = c->n, q = c->q;
    int u = ks_top(c);
    if (u < 0) { ks_emit0(c); return; }
    /* exclude u */
    c->c2[u] = 0; c->C2S--; c->X[c->XS++] = u;
    ks_G(c, r);
    c->XS--; c->c2[u] = 1; c->C2S++;
    if (r <= 0) return;   /* budget 0: exclude-only path (k == 1) */
    /* include u */
    c->c2[u] = 0; c->C2S--;
    c->P[c->PS++] = u;
    addG(c, u, c->nP);
    addG(c, u, c->nG);
    size_t mark = c->dsz;
    for (int v = 0; v < n; v++) {   /* updateCand13 */
        if (c->c1[v] && clsb(c, u, v) < UNLINK2EQUAL) {
            c->c1[v] = 0; c->C1S--; subG(c, v, c->nG);
            if (drops_push(c, v)) return;
        }
    }
    size_t mark2 = c->dsz;
    for (int v = 0; v < n; v++) {   /* updateCand23 */
        if (c->c2[v] && clsb(c, u, v) < UNLINK2EQUAL) {
            c->c2[v] = 0; c->C2S--;
            if (drops_push(c, v)) return;
        }
    }
    if (r == 1) {
        if (c->PS + c->C1S >= q && (c->PS <= 1 || ks_ub(c))) {
            size_t mark3 = c->dsz;
            for (int v = 0; v < n; v++) {
                if (c->c1[v] && !ks_iskplex(c, v)) {
                    c->c1[v] = 0; c->C1S--; subG(c, v, c->nG);
                    if (drops_push(c, v)) return;
                }
            }
            ks_emit(c);
            for (size_t t = mark3; t < c->dsz; t++) { int v = c->drops[t]; c->c1[v] = 1; c->C1S++; addG(c, v, c->nG); }
            c->dsz = mark3;
        }
    } else {
        ks_G(c, r - 1);
    }
    for (size_t t = mark2; t < c->dsz; t++) { c->c2[c->drops[t]] = 1; c->C2S++; }
    for (size_t t = mark; t < mark2; t++) { int v = c->drops[t]; c->c1[v] = 1; c->C1S++; addG(c, v, c->nG); }
    c->dsz = mark;
    c->PS--;
    subG(c, u, c->nP);
    subG(c, u, c->nG);
    c->c2[u] = 1; c->C2S++;
}

static int search_root(ctx_t *c, const sub_t *s)
{
    const int n = s->n;
    c->s = s; c->n = n;
    c->P = malloc(n * sizeof(int)); c->X = malloc(n * sizeof(int));
    c->c1 = calloc(n, 1); c->c2 = calloc(n, 1);
    c->nG = malloc(n * sizeof(int)); c->nP = malloc(n * sizeof(int));
    c->uni = calloc(s->l ? s->l : 1, 1);
    c->commons = malloc((s->l + 1) * sizeof(int));
    c->commons2 = malloc((n + 1) * sizeof(int));
    c->sat = malloc((n + 1) * sizeof(int));
    c->tmp_ids = malloc((n + 1) * sizeof(uint32_t));
    c->levels = 0; c->lab = NULL; c->bG = NULL; c->bP = NULL; c->PL = c->CL = c->XL = NULL;
    c->dsz = 0;
    int rc = grow_levels(c, n + 3);   /* Branch depth <= |C| + 1 <= n: buffers never move during the DFS */
    if (!c->P || !c->X || !c->c1 || !c->c2 || !c->nG || !c->nP || !c->uni || !c->commons ||
        !c->commons2 || !c->sat || !c->tmp_ids) rc = -12;
    if (!rc) {
        /* kSearch init (:1946-1976): neiInG = degreeHop, neiInP = [adj to root] */
        for (int i = 0; i < n; i++) { c->nG[i] = s->dhop[i]; c->nP[i] = 0; }
        for (int t = s->noff[0]; t < s->noff[1]; t++) c->nP[s->nbr[t]]++;
        c->PS = 0; c->P[c->PS++] = 0;
        c->C1S = 0; c->C2S = 0; c->XS = 0;
        for (int i = 1; i < s->h; i++) { c->c1[i] = 1; c->C1S++; }
        for (int i = s->h; i < n; i++) { c->c2[i] = 1; c->C2S++; }
        ks_G(c, c->k - 1);
    }
    free(c->P); free(c->X); free(c->c1); free(c->c2); free(c->nG); free(c->nP);
    free(c->uni); free(c->commons); free(c->commons2); free(c->sat); free(c->tmp_ids);
    free(c->lab); free(c->bG); free(c->bP); free(c->PL); free(c->CL); free(c->XL);
    return rc;
}

/* ------------------------------------------------------------------ */
/* driver (decomposableSearch :722-1162)                               */
/* ------------------------------------------------------------------ */
typedef struct { uint32_t *ids; uint64_t len; } plex_ref_t;

static const uint32_t *g_sort_base;
static int cmp_plex(const void *a, const void *b)
{
    const uint64_t *x = (const uint64_t *)a, *y = (const uint64_t *)b;
    /* x[0] = offset, x[1] = len */
    const uint32_t *px = g_sort_base + x[0], *py = g_sort_base + y[0];
    uint64_t lx = x[1], ly = y[1];
    for (uint64_t i = 0; i < lx && i < ly; i++) if (px[i] != py[i]) return px[i] < py[i] ? -1 : 1;
    return (lx > ly) - (lx < ly);
}

void kpo_result_free(kpo_result *r)
{
    if (!r) return;
    free(r->list_off); free(r->list_ids);
    r->list_off = NULL; r->list_ids = NULL; r->nlisted = 0;
}

int kpo_enumerate(const kpo_graph *g, int k, int q, int truss, int nthreads,
                  uint32_t flags, kpo_result *res)
{
    return kpo_enumerate_range(g, k, q, truss, nthreads, flags, 0, -1, res);
}

int kpo_enumerate_range(const kpo_graph *g, int k, int q, int truss, int nthreads,
                        uint32_t flags, long pos_begin, long pos_end, kpo_result *res)
{
    memset(res, 0, sizeof(*res));
    if (k < 1 || q < 2 * k) return -22;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    double t0 = now_ms();
    kpo_prep pr;
    int rc = kpo_preprocess(g, k, q, truss, &pr);
    if (rc) return rc;
    double t1 = now_ms();
    res->preprocess_ms = t1 - t0;
    const kpo_graph *G = &pr.g;
    res->pn = G->n; res->pm = G->m;
    long nroots = (long)G->n - (long)q + 2;
    if (nroots < 0) nroots = 0;
    if (pos_end >= 0 && pos_end < nroots) nroots = pos_end;
    long rbeg = pos_begin < 0 ? 0 : pos_begin;
    uint64_t count = 0, tasks = 0, valid = 0, ha = 0, hb = 0, algb = 0, adjb_ = 0;
    uint32_t maxn = 0, maxl = 0;
    int err = 0;
    int nth = omp_get_max_threads();
    plist_t *lists = calloc(nth, sizeof(plist_t));
    #pragma omp parallel reduction(+:count, tasks, valid, ha, hb, algb, adjb_) reduction(max:maxn, maxl)
    {
        dws_t w;
        int tid = omp_get_thread_num();
        int lrc = dws_init(&w, G->n);
        if (lrc) {
            #pragma omp atomic write
            err = lrc;
        }
        ctx_t c;
        memset(&c, 0, sizeof(c));
        c.k = k; c.q = q; c.flags = flags; c.list = &lists[tid];
        #pragma omp for schedule(dynamic, 1)
        for (long pos = rbeg; pos < nroots; pos++) {
            if (lrc) continue;
            uint32_t n, hop, l;
            int v = decompose_root(G, pr.dpos, pr.dseq, pos, k, q, &w, &n, &hop, &l);
            if (v <= 0) continue;
            valid++;
            if (n > maxn) maxn = n;
            if (l > maxl) maxl = l;
            /* W_r = d(r) + sum_{v in N+(r)} d(v) + sum_{u in B u L} d(u) */
            uint32_t r = (uint32_t)pr.dseq[pos];
            uint64_t Wr = G->deg[r];
            for (uint64_t p = G->off[r]; p < G->off[r + 1]; p++)
                if (pr.dpos[G->nbr[p]] > pos) Wr += G->deg[G->nbr[p]];
            for (uint32_t i = 0; i < n; i++) Wr += G->deg[w.B[i]];
            for (uint32_t i = 0; i < l; i++) Wr += G->deg[w.L[i]];
            adjb_ += 4 * Wr;
            algb += 4 * Wr + ((uint64_t)n * n + 7) / 8 + (uint64_t)n * n;
            sub_t s;
            if (build_sub(G, &w, n, hop, l, k, q, &s)) { free_sub(&s); err = -12; continue; }
            s.B = w.B; s.orig = pr.orig;
            uint64_t t_before = c.tasks;
            if (search_root(&c, &s)) err = -12;
            algb += (c.tasks - t_before) * 2ull * (5ull * n + 16ull);
            free_sub(&s);
        }
        count += c.count; tasks += c.tasks; ha += c.ha; hb += c.hb;
        dws_free(&w);
    }
    res->count = count; res->tasks = tasks; res->valid_roots = valid;
    res->hash_a = ha; res->hash_b = hb; res->alg_bytes = algb; res->adj_bytes = adjb_;
    res->max_n = maxn; res->max_l = maxl;
    if (flags & KPO_LIST) {
        uint64_t tot = 0, np = 0;
        for (int t = 0; t < nth; t++) { np += lists[t].n; tot += lists[t].sz - lists[t].n; }
        res->nlisted = np;
        res->list_ids = malloc((tot ? tot : 1) * 4);
        uint64_t (*idx)[2] = malloc((np ? np : 1) * sizeof(*idx));
        uint32_t *tmp = malloc((tot ? tot : 1) * 4);
        res->list_off = malloc((np + 1) * 8);
        if (!res->list_ids || !idx || !tmp || !res->list_off) { err = -12; }
        else {
            uint64_t o = 0, pi = 0;
            for (int t = 0; t < nth; t++) {
                uint64_t i = 0;
                while (i < lists[t].sz) {
                    uint64_t len = lists[t].ids[i++];
                    idx[pi][0] = o; idx[pi][1] = len; pi++;
                    for (uint64_t j = 0; j < len; j++) tmp[o++] = (uint32_t)lists[t].ids[i++];
                }
            }
            g_sort_base = tmp;
            qsort(idx, np, sizeof(*idx), cmp_plex);
            o = 0;
            for (uint64_t p = 0; p < np; p++) {
                res->list_off[p] = o;
                memcpy(res->list_ids + o, tmp + idx[p][0], idx[p][1] * 4);
                o += idx[p][1];
            }
            res->list_off[np] = o;
        }
        free(idx); free(tmp);
    }
    for (int t = 0; t < nth; t++) free(lists[t].ids);
    free(lists);
    kpo_prep_free(&pr);
    res->wall_ms = now_ms() - t0;
    return err;
}

int kpo_root_block(const kpo_prep *p, int k, int q, long pos,
                   uint32_t *B, uint32_t *n, uint32_t *hop,
                   uint32_t *L, uint32_t *l, uint8_t *cls, size_t cls_cap)
{
    const kpo_graph *G = &p->g;
    if (pos < 0 || pos >= (long)G->n) return -22;
    dws_t w;
    if (dws_init(&w, G->n)) { dws_free(&w); return -12; }
    int v = decompose_root(G, p->dpos, p->dseq, pos, k, q, &w, n, hop, l);
    if (v > 0 || *n) {
        memcpy(B, w.B, (size_t)(*n) * 4);
        memcpy(L, w.L, (size_t)(*l) * 4);
    }
    if (v > 0 && cls) {
        sub_t s;
        if ((size_t)(*n) * (*n) > cls_cap) { dws_free(&w); return -34; }
        if (build_sub(G, &w, *n, *hop, *l, k, q, &s)) { free_sub(&s); dws_free(&w); return -12; }
        memcpy(cls, s.cls, (size_t)(*n) * (*n));
        free_sub(&s);
    }
    dws_free(&w);
    return v;
}
