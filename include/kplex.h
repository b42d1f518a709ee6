/*
 * kplex.h -- C ABI of the MI355X-native maximal k-plex enumeration library
 * (libkplex.so).  Plain C types only; no HIP or torch types cross it.
 *
 * The reference (CudaPlex) has no library boundary: it is one executable
 * whose de-facto interface is
 *   - the binary CSR file          inc/graph.h:37-65
 *   - the CLI  main <g> -k -q -t -tr   main.cu:11-62
 *   - the parameter globals k, lb(=q), thres, truss, bd   inc/kPlexEnum.h:20-24
 *   - the entry void decomposableSearch(const graph<int>&)  inc/host_funcs.h:722
 *   - its stdout lines              inc/host_funcs.h:1149-1157, inc/graph.h:47
 * Each entry point below names the piece of that interface it replaces.
 *
 * Conventions: every call returns 0 (KP_OK) or a negative KP_E* code.  The
 * caller owns its inputs; the library owns what it returns through out
 * structs and frees it with the matching *_free call.  Handles are not
 * shared between threads without external locking.
 */
#ifndef KPLEX_H
#define KPLEX_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------- */
#define KP_OK          0
#define KP_ENOENT     (-2)    /* file not found (reference: utils::open_file exits, inc/utils.h:24-33) */
#define KP_EIO        (-5)    /* short read / write */
#define KP_ENOMEM     (-12)   /* host or device allocation failed */
#define KP_ENODEV     (-19)   /* no HIP device / library built without one */
#define KP_EINVAL     (-22)   /* bad argument (k < 1, q < 2k, malformed graph) */
#define KP_ERANGE     (-34)   /* a root block exceeds KP_MAX_BLOCK vertices */
#define KP_EOVERFLOW  (-75)   /* an internal capacity (queue, stack, list buffer) was exceeded */
#define KP_ETIMEDOUT  (-110)  /* the time limit in kp_params was reached; results are partial */
#define KP_EBUSY      (-16)   /* server: admission refused (memory / slots) */
#define KP_EHIP       (-1000) /* HIP runtime error; kp_last_error() has the text */

#define KP_MAX_BLOCK  1024    /* largest root block (reference MAX_BLK_SIZE, inc/common.h:8) */

/* ---- graph (inc/graph.h) -------------------------------------------- */
typedef struct kp_graph {
    uint32_t  n;     /* vertices */
    uint64_t  m;     /* directed adjacency entries = 2 x undirected edges */
    uint64_t *off;   /* n+1 offsets */
    uint32_t *nbr;   /* m neighbour ids, each list sorted ascending */
    uint32_t *deg;   /* n degrees */
} kp_graph;

/* Replaces graph<intT>::graph(const char*) (inc/graph.h:37-65): reads
 * u32 n, u32 m, u32 deg[n], u32 nbrs[m].  Unlike the reference it fails
 * with KP_ENOENT/KP_EIO/KP_EINVAL instead of exiting or reading past the end. */
int  kp_graph_load_bin(const char *path, kp_graph *out);
/* Writes the same format (the write order of inc/graphIO.h:19-25). */
int  kp_graph_save_bin(const kp_graph *g, const char *path);
/* Copies a CSR given as degrees + concatenated neighbour lists. */
int  kp_graph_from_csr(uint32_t n, const uint32_t *deg, const uint32_t *nbr, kp_graph *out);
/* Checks sortedness, symmetry and absence of self loops / duplicates
 * (the reference assumes all of them: its lookups are binary searches). */
int  kp_graph_validate(const kp_graph *g);
void kp_graph_free(kp_graph *g);

/* ---- parameters (inc/kPlexEnum.h:20-24; main.cu:31-53) -------------- */
typedef struct kp_params {
    int      k;            /* -k : at most k-1 non-neighbours per vertex inside the plex */
    int      q;            /* -q : minimum plex size (reference global lb) */
    float    thres;        /* -t : fraction of queue_capacity kept on the device as resumable
                              frames after a pause; the rest is spilled to pinned host memory */
    int      truss;        /* -tr: 1 = (q-2k)-truss prefilter before the core */
    uint32_t flags;        /* KP_FLAG_* */
    double   time_limit_s; /* 0 = none; else KP_ETIMEDOUT after this many seconds of search */
    uint64_t queue_capacity; /* device task capacity in frames per size class; 0 = 4,194,304
                              (the reference's MAX_CAP) */
    int64_t  root_begin;   /* enumerate only plexes whose first vertex in the degeneracy */
    int64_t  root_end;     /* order sits at a position in [root_begin, root_end); 0,0 = all */
    uint64_t pause_every;  /* 0 = never; else each search launch pauses after about this many
                              frames, its pending frames are carried (device / pinned host)
                              and a new launch resumes them.  Results do not change. */
} kp_params;

#define KP_FLAG_LIST      1u  /* return every maximal k-plex (original ids, sorted) */
#define KP_FLAG_VALIDATE  2u  /* run kp_graph_validate first */
#define KP_FLAG_VERBOSE   4u  /* print the reference's progress lines to stdout */
#define KP_FLAG_GPU_PREP  8u  /* (q-k)-core peel, relabel and truss supports on the GPU; the truss
                                 peel and the degeneracy order stay on the host (same result) */
#define KP_FLAG_NO_NARROW 16u /* keep every frame in its root's size class: no re-rooting of
                                 frames of wide blocks into the one-word class (same result) */

/* ---- result (inc/host_funcs.h:1144-1157) ---------------------------- */
typedef struct kp_result {
    uint64_t count;        /* maximal k-plexes of size >= q ("Maximal k-Plexes", 64-bit) */
    uint64_t tasks;        /* Branch tasks produced ("Total tasks generated") */
    uint64_t valid_roots;  /* roots whose block has >= q vertices ("Total Valid Blocks") */
    uint64_t hash_a, hash_b; /* order-independent 128-bit multiset hash, see below */
    uint32_t pn;           /* vertices after preprocessing ("pn = ") */
    uint64_t pm;           /* directed entries after preprocessing */
    uint32_t max_n, max_l; /* largest root block / left set */
    uint64_t alg_bytes;    /* sum_r [4 W_r + ceil(n_r^2/8) + n_r^2] + sum_tasks 2 (5 n + 16) */
    uint64_t adj_bytes;    /* sum_r 4 W_r (CSR bytes read to build the root subgraphs) */
    uint64_t search_alg_bytes; /* sum_tasks 2 (5 n + 16): the search kernels' share of alg_bytes */
    double   preprocess_ms;/* host core/truss/degeneracy */
    double   upload_ms;    /* host -> HBM copies */
    double   decompose_ms; /* GPU root decomposition + subgraph build */
    double   search_ms;    /* GPU persistent k-search + Branch */
    double   wall_ms;      /* preprocess + decompose + search (reference "Total Time Elapsed") */
    uint64_t spilled;      /* tasks spilled to pinned host memory */
    uint64_t pauses;       /* pause/resume cycles */
    uint64_t nlisted;      /* KP_FLAG_LIST: plex i is list_ids[list_off[i] .. list_off[i+1]) */
    uint64_t *list_off;
    uint32_t *list_ids;
    uint32_t pn_first;     /* vertices after the first (q-k)-core; with truss on the reference
                              prints it and then pn after the second core (inc/host_funcs.h:115,
                              called at :749 and :823) */
    int32_t  status;       /* kp_server_poll: KP_OK, or the KP_E* code that failed this query
                              (its counts are then zero) */
    uint64_t narrowed;     /* frames of blocks wider than 64 vertices whose live set fit one
                              word and that were re-rooted into the one-word class */
} kp_result;
/*
 * Hash: for each maximal k-plex S of original vertex ids
 *   a(S) = mix64( sum_{v in S} mix64(v ^ 0x243F6A8885A308D3) + |S| * 0x13198A2E03707344 )
 *   b(S) = mix64( sum_{v in S} mix64(v ^ 0xA4093822299F31D0) + |S| * 0x082EFA98EC4E6C89 )
 * hash_a = sum_S a(S), hash_b = sum_S b(S) (mod 2^64), mix64 = SplitMix64
 * finalizer (Steele, Lea & Flood 2014).  Independent of enumeration order.
 */

/* Replaces decomposableSearch(const graph<int>&) (inc/host_funcs.h:722-1162)
 * with the globals k, lb, thres, truss passed in *p: preprocessing on the
 * host, enumeration on the current HIP device. */
int  kp_enumerate(const kp_graph *g, const kp_params *p, kp_result *r);
void kp_result_free(kp_result *r);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same enumeration as kp_enumerate, same arguments and
 * results, with no device or pinned-host allocation.  Every device buffer is
 * carved out of one static device array of KP_NOALLOC_SCRATCH_BYTES that each
 * GPU holds from the library's first use on it until the process ends; the
 * enumeration runs as the monolithic per-root kernels (one wave per root) on
 * the null stream.  Limits: the preprocessed graph and the per-wave buffers
 * must fit the array (else KP_ENOMEM, with the bytes needed in
 * kp_last_error), root blocks of at most 256 vertices (else KP_ERANGE); no
 * plex lists and no pauses (KP_EINVAL).  Calls that overlap on one GPU race
 * on the shared array: serialize them. */
#define KP_NOALLOC_SCRATCH_BYTES (64u << 20)
int  kp_enumerate_noalloc(const kp_graph *g, const kp_params *p, kp_result *r);

/* Host preprocessing only (no device needed): the (q-k)-core after the
 * optional truss (inc/host_funcs.h:749-823), its new->original id map and the
 * degeneracy order (inc/host_funcs.h:829-851).  out receives the preprocessed
 * CSR; orig and dseq must have room for g->n entries. */
int  kp_preprocess(const kp_graph *g, int k, int q, int truss, kp_graph *out,
                   uint32_t *orig, int32_t *dseq);

/* The same as kp_preprocess, with the (q-k)-core peel, the stable relabel
 * and the truss supports (inc/host_funcs.h:16-122, :761-825) computed on the
 * current HIP device; the one-sided truss peel (:463-559) and the degeneracy
 * order (inc/degeneracy.h) stay on the host, whose orders they follow.  The
 * output equals kp_preprocess's. */
int  kp_preprocess_gpu(const kp_graph *g, int k, int q, int truss, kp_graph *out,
                       uint32_t *orig, int32_t *dseq);

/* ---- prepared (HBM-resident) graphs --------------------------------- */
/* Split of kp_enumerate: kp_prepare runs the host preprocessing for (k, q,
 * truss) and uploads the preprocessed CSR + degeneracy order to HBM;
 * kp_run then enumerates with every input already resident.  No reference
 * counterpart: the reference re-runs the whole pipeline per invocation. */
typedef struct kp_prepared kp_prepared;
int  kp_prepare(const kp_graph *g, const kp_params *p, kp_prepared **out);
int  kp_run(kp_prepared *h, kp_result *r);
uint64_t kp_prepared_bytes(const kp_prepared *h);
/* Sets the root window of later kp_run calls (the meaning of kp_params.root_begin
 * / root_end), so that a resident graph is enumerated window by window without
 * preprocessing it again. */
int  kp_set_window(kp_prepared *h, int64_t root_begin, int64_t root_end);
/* Inspection hook for root-by-root parity: the decomposition and subgraph of
 * the root at degeneracy position pos of a prepared graph, as the GPU builds
 * them -- what the reference's decompose, calculateDegrees / fillNeighbors and
 * buildCommonMtx compute for one root (src/device_funcs.cu:28-460,
 * :1727-1853).  block (room for KP_MAX_BLOCK) and left (room for left_cap)
 * receive peeled vertex ids in the canonical order (the left set's 2-hop part
 * in the order the GPU claimed it: the left set is used as a set); pairs, unless NULL,
 * receives n*n bytes, pairs[i*n+j] bit 0 = block[i] adjacent to block[j],
 * bit 1 = pair class >= UNLINK2EQUAL, bit 2 = pair class >= UNLINK2MORE.
 * Returns 1 for a valid root (n >= q), 0 for an invalid one, < 0 on error. */
int  kp_root_subgraph(kp_prepared *h, int64_t pos, uint32_t *block, uint32_t *n, uint32_t *hop,
                      uint32_t *left, uint32_t *l, uint32_t left_cap, uint8_t *pairs, size_t pairs_cap);
void kp_prepared_free(kp_prepared *h);

/* ---- launch settings per size class (autotuner) --------------------- */
/* No reference counterpart (the reference has fixed WARPS and queue sizes,
 * inc/common.h).  Index c is the size class of blocks with at most 64 * W_c
 * vertices, W = {1, 2, 3, 4, 6, 8, 16}; a 0 entry keeps the built-in default. */
typedef struct kp_tune {
    int32_t  blocks_per_cu[7];  /* resident 256-thread blocks per CU (<= occupancy limit) */
    uint32_t donate_period[7];  /* frames between donation checks (default 256) */
    uint32_t ring_frames[7];    /* frames per wave in the LDS stack ring (default 8 KiB worth) */
    uint32_t queue_slots[7];    /* frame-queue slots per shard (default max(256, waves / 8)) */
    double   class_ms[7];       /* kp_autotune: best search time of the class alone */
} kp_tune;
/* Times each size class of the prepared workload alone under candidate
 * settings (coordinate descent) and makes the fastest the process-wide
 * settings, which every later run, batch and server launch uses. */
int  kp_autotune(kp_prepared *h, kp_tune *out);
/* Sets the process-wide settings (NULL = defaults). */
void kp_set_tune(const kp_tune *t);

/* ---- batched: many graphs in one persistent launch ------------------ */
/* No reference counterpart (the reference handles one graph per process).
 * p[i] holds graph i's k, q, truss and root window; flags, time limit and
 * pause settings are taken from p[0].  All roots of all graphs share one set
 * of size-class launches; r[i] is graph i's result (decompose/search times are
 * the batch's). */
int  kp_enumerate_batch(const kp_graph *const *g, const kp_params *p, size_t n, kp_result *r);

/* ---- served: continuous batching of (graph, k, q) queries ----------- */
/* No reference counterpart.  Registered graphs stay in host memory; each
 * (graph, q-k, truss) is preprocessed once and its CSR + degeneracy order stay
 * resident in HBM.  Queries run in epochs of the persistent search launches:
 * while queries wait, the running epoch is paused after slice_ms (every wave
 * writes its pending frames out), the waiting queries are admitted (root
 * decomposition + subgraph build) as job slots and arena memory allow, and
 * the next epoch resumes the paused frames together with the new roots.  A
 * query's result is the same alone or in any mix. */
typedef struct kp_server kp_server;
typedef struct kp_server_config {
    size_t   max_inflight;      /* queries admitted at once (0 = 64) */
    uint64_t mem_budget_bytes;  /* admission budget: subgraph arena bytes of admitted queries
                                   (0 = 4 GiB); the search stacks and queues are a fixed
                                   per-size-class cost, not per query */
    double   slice_ms;          /* minimum epoch length before a pause for waiting queries
                                   (0 = 30 ms) */
} kp_server_config;
int  kp_server_create(const kp_server_config *cfg, kp_server **out);
/* Registers a graph (copied). */
int  kp_server_add_graph(kp_server *s, const kp_graph *g, uint32_t *graph_id);
/* Preprocesses (graph_ids[i], p[i].k, p[i].q, p[i].truss) for i < n, in
 * parallel on the host, and keeps each result resident in HBM, so that later
 * queries on them start with decomposition. */
int  kp_server_preload(kp_server *s, const uint32_t *graph_ids, const kp_params *p, size_t n);
/* Queues a query (k, q, truss of *p; lists and root windows are not served);
 * returns its ticket.  Admission happens in kp_server_poll. */
int  kp_server_submit(kp_server *s, uint32_t graph_id, const kp_params *p, uint64_t *ticket);
/* One scheduling step, never blocking on the device: finishes an epoch that
 * has ended, admits waiting queries, starts the next epoch, requests a pause
 * when queries wait.  Returns up to cap completed queries; in each result
 * wall_ms is submit -> completion, search_ms admission -> completion and
 * decompose_ms the admission (decompose + build) time of its group.  A query
 * that cannot run (preprocessing failed, a block larger than KP_MAX_BLOCK, a
 * subgraph arena larger than the whole budget, an error reported by a search
 * kernel) comes back with status set and zero counts; the call itself fails
 * only when the HIP runtime does. */
int  kp_server_poll(kp_server *s, uint64_t *tickets, kp_result *results, size_t cap, size_t *ndone);
/* Number of queries submitted and not yet returned by kp_server_poll. */
size_t kp_server_outstanding(const kp_server *s);
int  kp_server_destroy(kp_server *s);

/* ---- misc ------------------------------------------------------------ */
const char *kp_strerror(int code);
const char *kp_last_error(void);
/* Fills buf with "<device name>, <CUs> CUs, <HBM GiB> GiB" of the current device. */
int  kp_device_info(char *buf, size_t len);

/* ---- synthetic inputs (data layer; no reference counterpart) ---------- */
/* G(n, p) plus `groups` disjoint planted groups of `gsize` vertices, each a
 * clique from which up to `max_removed` edges per vertex are removed. */
int  kp_gen_er_planted(uint32_t n, double p, uint32_t groups, uint32_t gsize,
                       uint32_t max_removed, uint64_t seed, kp_graph *out);
/* Chung-Lu graph with power-law expected degrees (exponent gamma, largest
 * expected degree max_deg, about m_target undirected edges) plus `ncomm`
 * disjoint communities of size U[cmin, cmax] at intra-density `dens`. */
int  kp_gen_chung_lu(uint32_t n, double gamma, uint64_t m_target, double max_deg,
                     uint32_t ncomm, uint32_t cmin, uint32_t cmax, double dens,
                     uint64_t seed, kp_graph *out);
/* Applies a seeded random permutation to the vertex ids. */
int  kp_graph_relabel(const kp_graph *g, uint64_t seed, kp_graph *out, uint32_t *perm_out);

#ifdef __cplusplus
}
#endif
#endif
