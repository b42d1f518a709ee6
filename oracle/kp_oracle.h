/*
 * kp_oracle.h -- CPU oracle for maximal k-plex enumeration (TEST INFRASTRUCTURE).
 *
 * This is the checker, not the product.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  It is a plain C/OpenMP
 * restatement of the reference's hot path, written from the reference's
 * behaviour:
 *   loader            inc/graph.h:37-65
 *   (q-k)-core peel   inc/host_funcs.h:16-122
 *   (q-2k)-truss      inc/host_funcs.h:761-825, :463-559, :405-415
 *   degeneracy order  inc/degeneracy.h:23-203, inc/host_funcs.h:829-851
 *   decompose         src/device_funcs.cu:28-298
 *   local subgraph    src/device_funcs.cu:300-460, :1727-1853
 *   k-search          src/device_funcs.cu:1855-2444 (+ helpers :568-641, :1002-1195)
 *   Branch (BNB)      src/device_funcs.cu:1516-1724 (+ :586-622, :643-1040, :1197-1513)
 *
 * Deliberate, result-invariant deviations (the reference is racy here):
 *   - canonical in-block order: root, kept later 1-hop ascending by peeled id,
 *     later 2-hop ascending; left set: kept earlier 1-hop ascending, earlier
 *     2-hop ascending;
 *   - k-search branches on the highest-index vertex left in C2;
 *   - Branch pivot ties broken by (neiInG, neiInP, local index);
 *   - 64-bit counters, no capacity truncation, no queue-capacity abort.
 *
 * Parity pins: README known-answer counts (soc-epinions k4 q30, email-euall
 * k3 q12), closed forms (complete multipartite graphs, disjoint cliques),
 * brute force on tiny graphs, and invariants (relabelling, truss on/off).
 */
#ifndef KP_ORACLE_H
#define KP_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t n;
    uint64_t m;        /* directed adjacency entries (= 2 x undirected edges) */
    uint64_t *off;     /* n+1 */
    uint32_t *nbr;     /* m, each list sorted ascending */
    uint32_t *deg;     /* n */
} kpo_graph;

typedef struct {
    uint64_t count;        /* maximal k-plexes of size >= q */
    uint64_t tasks;        /* Branch tasks produced (k-search emits + BNB children) */
    uint64_t valid_roots;  /* roots whose block has >= q vertices */
    uint64_t hash_a, hash_b;
    uint64_t alg_bytes;    /* sum_r [4 W_r + ceil(n_r^2/8) + n_r^2] + sum_t 2 (5 n + 16) */
    uint64_t adj_bytes;    /* sum_r 4 W_r */
    uint32_t pn;           /* vertices after preprocessing */
    uint64_t pm;           /* directed entries after preprocessing */
    uint32_t max_n, max_l; /* largest block / left set */
    double wall_ms, preprocess_ms;
    /* optional listing (flags & KPO_LIST): plex i = ids[off[i] .. off[i+1]),
     * original vertex ids sorted ascending; plexes sorted lexicographically. */
    uint64_t nlisted;
    uint64_t *list_off;
    uint32_t *list_ids;
} kpo_result;

#define KPO_LIST 1u

int  kpo_load_bin(const char *path, kpo_graph *g);
int  kpo_graph_from_csr(uint32_t n, const uint32_t *deg, const uint32_t *nbr, kpo_graph *g);
void kpo_graph_free(kpo_graph *g);

int  kpo_enumerate(const kpo_graph *g, int k, int q, int truss, int nthreads,
                   uint32_t flags, kpo_result *res);
/* Same, restricted to root positions [pos_begin, pos_end) of the degeneracy
 * order (a bounded sample for timing); pos_end < 0 means all roots. */
int  kpo_enumerate_range(const kpo_graph *g, int k, int q, int truss, int nthreads,
                         uint32_t flags, long pos_begin, long pos_end, kpo_result *res);
void kpo_result_free(kpo_result *r);

/* Preprocessing only: peeled (and optionally trussed) graph, new->original id
 * map, degeneracy order.  Arrays are malloc'd; free with kpo_prep_free. */
typedef struct {
    kpo_graph g;
    uint32_t *orig;   /* pn */
    int32_t  *dseq;   /* pn */
    int32_t  *dpos;   /* pn */
} kpo_prep;
int  kpo_preprocess(const kpo_graph *g, int k, int q, int truss, kpo_prep *p);
void kpo_prep_free(kpo_prep *p);

/* One root's decomposition at degeneracy position pos.  Returns 1 if the
 * root is valid (block >= q), 0 if not, <0 on error.  B/L get peeled ids in
 * canonical order (caller buffers of size pn); cls (optional, n*n bytes) gets
 * the pair-class matrix. */
int  kpo_root_block(const kpo_prep *p, int k, int q, long pos,
                    uint32_t *B, uint32_t *n, uint32_t *hop,
                    uint32_t *L, uint32_t *l, uint8_t *cls, size_t cls_cap);

uint64_t kpo_mix64(uint64_t z);

#ifdef __cplusplus
}
#endif
#endif
