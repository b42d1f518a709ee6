// kp_prep.cpp -- host preprocessing (stays on the host, C++/OpenMP):
//   (q-k)-core peel with stable relabel     inc/host_funcs.h:16-122
//   optional (q-2k)-truss + second core     inc/host_funcs.h:761-825, :463-559
//   degeneracy order (LIFO bucket heap)     inc/degeneracy.h:23-203, inc/host_funcs.h:829-851
#include "kp_internal.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <ctime>

namespace kp {

namespace {
thread_local char g_err[512];
}

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}
const char *last_error() { return g_err; }

double now_ms()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

namespace {

// adjacency entries below which preprocessing stays on the calling thread
constexpr uint64_t kParallelEntries = 1u << 18;

} // namespace

namespace {

// The same core and relabel on the calling thread, for small graphs: one
// stack of peeled vertices, one pass each for the new degrees and lists.
int core_peel_small(const kp_graph *g, int bd, kp_graph *out, std::vector<uint32_t> &old_of_new)
{
    const uint32_t n = g->n;
    std::vector<int32_t> res(n);
    std::vector<uint32_t> map(n), stack;
    stack.reserve(n);
    for (uint32_t i = 0; i < n; i++) {
        res[i] = int32_t(g->deg[i]);
        map[i] = int64_t(g->deg[i]) >= bd ? 0u : UINT32_MAX;
        if (map[i]) stack.push_back(i);
    }
    while (!stack.empty()) {
        const uint32_t u = stack.back();
        stack.pop_back();
        for (uint64_t p = g->off[u]; p < g->off[u + 1]; p++) {
            const uint32_t v = g->nbr[p];
            if (map[v] == 0 && res[v]-- == bd) {
                map[v] = UINT32_MAX;
                stack.push_back(v);
            }
        }
    }
    old_of_new.clear();
    uint64_t pm = 0;
    for (uint32_t i = 0; i < n; i++)
        if (map[i] == 0) {
            map[i] = uint32_t(old_of_new.size());
            old_of_new.push_back(i);
            pm += uint64_t(res[i]);   // survivors keep exactly their surviving neighbours
        }
    const uint32_t pn = uint32_t(old_of_new.size());
    int rc = graph_alloc(out, pn, pm);
    if (rc) return rc;
    uint64_t w = 0;
    out->off[0] = 0;
    for (uint32_t i = 0; i < pn; i++) {
        const uint32_t o = old_of_new[i];
        for (uint64_t p = g->off[o]; p < g->off[o + 1]; p++) {
            const uint32_t v = map[g->nbr[p]];
            if (v != UINT32_MAX) out->nbr[w++] = v;
        }
        out->deg[i] = uint32_t(w - out->off[i]);
        out->off[i + 1] = w;
    }
    return KP_OK;
}

} // namespace

// The bd-core is unique, so the parallel cascade below gives the same vertex
// set as the reference's per-thread queues with CAS decrements.
int core_peel(const kp_graph *g, int bd, kp_graph *out, std::vector<uint32_t> &old_of_new)
{
    if (g->m <= kParallelEntries) return core_peel_small(g, bd, out, old_of_new);
    const uint32_t n = g->n;
    std::vector<int32_t> res(n);
    std::vector<uint8_t> alive(n);
    std::vector<uint32_t> frontier, next;
    for (uint32_t i = 0; i < n; i++) {
        res[i] = int32_t(g->deg[i]);
        alive[i] = int64_t(g->deg[i]) >= bd;
        if (!alive[i]) frontier.push_back(i);
    }
    // Small inputs run on the calling thread: waking an OpenMP team costs
    // more than peeling a few thousand vertices.
    const bool par = g->m > kParallelEntries;
    while (!frontier.empty()) {
        next.clear();
        #pragma omp parallel if (par && frontier.size() > 1024)
        {
            std::vector<uint32_t> mine;
            #pragma omp for schedule(dynamic, 256) nowait
            for (size_t t = 0; t < frontier.size(); t++) {
                uint32_t e = frontier[t];
                for (uint64_t p = g->off[e]; p < g->off[e + 1]; p++) {
                    uint32_t v = g->nbr[p];
                    if (__atomic_load_n(&alive[v], __ATOMIC_RELAXED)) {
                        int32_t old = __atomic_fetch_sub(&res[v], 1, __ATOMIC_RELAXED);
                        if (old == bd) {
                            __atomic_store_n(&alive[v], uint8_t(0), __ATOMIC_RELAXED);
                            mine.push_back(v);
                        }
                    }
                }
            }
            #pragma omp critical
            next.insert(next.end(), mine.begin(), mine.end());
        }
        frontier.swap(next);
    }
    // stable relabel: new id = rank among survivors (sequence::pack, inc/sequence.h:475-490)
    std::vector<uint32_t> map(n, UINT32_MAX);
    old_of_new.clear();
    for (uint32_t i = 0; i < n; i++)
        if (alive[i]) {
            map[i] = uint32_t(old_of_new.size());
            old_of_new.push_back(i);
        }
    const uint32_t pn = uint32_t(old_of_new.size());
    std::vector<uint32_t> ndeg(pn);
    #pragma omp parallel for schedule(dynamic, 256) if (par)
    for (long i = 0; i < long(pn); i++) {
        uint32_t o = old_of_new[i], c = 0;
        for (uint64_t p = g->off[o]; p < g->off[o + 1]; p++) c += alive[g->nbr[p]];
        ndeg[i] = c;
    }
    uint64_t pm = 0;
    for (uint32_t i = 0; i < pn; i++) pm += ndeg[i];
    int rc = graph_alloc(out, pn, pm);
    if (rc) return rc;
    out->off[0] = 0;
    for (uint32_t i = 0; i < pn; i++) {
        out->deg[i] = ndeg[i];
        out->off[i + 1] = out->off[i] + ndeg[i];
    }
    #pragma omp parallel for schedule(dynamic, 256) if (par)
    for (long i = 0; i < long(pn); i++) {
        uint32_t o = old_of_new[i];
        uint64_t w = out->off[i];
        for (uint64_t p = g->off[o]; p < g->off[o + 1]; p++) {
            uint32_t v = g->nbr[p];
            if (alive[v]) out->nbr[w++] = map[v];
        }
    }
    return KP_OK;
}

namespace {

long find_entry(const kp_graph *g, uint32_t u, uint32_t v)
{
    const uint32_t *b = g->nbr + g->off[u], *e = g->nbr + g->off[u + 1];
    const uint32_t *it = std::lower_bound(b, e, v);
    return (it == e || *it != v) ? -1 : long(it - g->nbr);
}

// (q-2k)-truss prefilter.  Behaviour kept exactly: per-directed-entry
// supports, seeds (i < v, support < q-2k) in (i, list) order, LIFO stack,
// only the u->w and v->w entries decremented for a common neighbour w, and
// an edge may be popped again after it is dead.
} // namespace

// Support of every directed entry (u, v): |N(u) ∩ N(v)| by a sorted merge.
void truss_supports(const kp_graph *g, std::vector<int32_t> &tri)
{
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    tri.assign(m, 0);
    #pragma omp parallel for schedule(dynamic, 64) if (m > kParallelEntries)
    for (long i = 0; i < long(n); i++) {
        for (uint64_t p = g->off[i]; p < g->off[i + 1]; p++) {
            const uint32_t v = g->nbr[p];
            const uint32_t *a = g->nbr + g->off[i], *ae = g->nbr + g->off[i + 1];
            const uint32_t *b = g->nbr + g->off[v], *be = g->nbr + g->off[v + 1];
            int32_t c = 0;
            while (a < ae && b < be) {
                if (*a == *b) { c++; a++; b++; }
                else if (*a < *b) a++;
                else b++;
            }
            tri[p] = c;
        }
    }
}

// The truss peel with the supports already counted (tri is consumed).
int truss_peel(kp_graph *g, int q, int k, std::vector<int32_t> &tri)
{
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const int thr = q - 2 * k;
    std::vector<std::pair<uint32_t, uint32_t>> stack;
    for (uint32_t i = 0; i < n; i++)
        for (uint64_t p = g->off[i]; p < g->off[i + 1]; p++)
            if (i < g->nbr[p] && tri[p] < thr) stack.emplace_back(i, g->nbr[p]);
    std::vector<uint8_t> dead(m, 0);
    while (!stack.empty()) {
        auto [u, v] = stack.back();
        stack.pop_back();
        long puv = find_entry(g, u, v), pvu = find_entry(g, v, u);
        if (puv >= 0 && !dead[puv]) { dead[puv] = 1; g->deg[u]--; }
        if (pvu >= 0 && !dead[pvu]) { dead[pvu] = 1; g->deg[v]--; }
        uint64_t iu = g->off[u], eu = g->off[u + 1], iv = g->off[v], ev = g->off[v + 1];
        while (iu < eu && iv < ev) {
            while (iu < eu && dead[iu]) ++iu;
            while (iv < ev && dead[iv]) ++iv;
            if (iu >= eu || iv >= ev) break;
            const uint32_t a = g->nbr[iu], b = g->nbr[iv];
            if (a == b) {
                if (tri[iu] > 0) { int32_t prev = tri[iu]--; if (prev == thr) stack.emplace_back(u, a); }
                if (tri[iv] > 0) { int32_t prev = tri[iv]--; if (prev == thr) stack.emplace_back(v, a); }
                iu++;
                iv++;
            } else if (a < b) {
                iu++;
            } else {
                iv++;
            }
        }
    }
    uint64_t w = 0;
    for (uint32_t u = 0; u < n; u++) {
        uint64_t b = g->off[u], e = g->off[u + 1];
        g->off[u] = w;
        for (uint64_t i = b; i < e; i++)
            if (!dead[i]) g->nbr[w++] = g->nbr[i];
    }
    g->off[n] = w;
    g->m = w;
    return KP_OK;
}

// Min-degree peeling order with the reference's bucket heap: inserts go to
// the head of a bucket (LIFO), pop takes the head of the lowest non-empty
// bucket, a decrement moves the vertex to the head of its new bucket.
void degeneracy_order(const kp_graph *g, std::vector<int32_t> &dseq, std::vector<int32_t> &dpos, uint32_t *max_later)
{
    const int n = int(g->n);
    dseq.assign(n, 0);
    dpos.assign(n, INT_MAX);
    if (!n) return;
    std::vector<int> key(n), prv(n), nxt(n), head(n, n);
    int min_key = n - 1;
    auto insert = [&](int id, int k) {
        key[id] = k;
        prv[id] = n;
        nxt[id] = head[k];
        if (head[k] != n) prv[head[k]] = id;
        head[k] = id;
        if (k < min_key) min_key = k;
    };
    auto remove = [&](int id) {
        if (prv[id] == n) {
            head[key[id]] = nxt[id];
            if (nxt[id] != n) prv[nxt[id]] = n;
        } else {
            nxt[prv[id]] = nxt[id];
            if (nxt[id] != n) prv[nxt[id]] = prv[id];
        }
    };
    for (int i = 0; i < n; i++) insert(i, int(g->deg[i]));
    for (int i = 0; i < n; i++) {
        while (head[min_key] == n) ++min_key;
        const int u = head[min_key];
        // the key at the pop is u's number of later neighbours
        if (max_later && uint32_t(min_key) > *max_later) *max_later = uint32_t(min_key);
        head[min_key] = nxt[u];
        if (head[min_key] != n) prv[head[min_key]] = n;
        dpos[u] = i;
        dseq[i] = u;
        for (uint64_t p = g->off[u]; p < g->off[u] + g->deg[u]; p++) {
            const int v = int(g->nbr[p]);
            if (dpos[v] == INT_MAX) {
                const int nk = key[v] - 1;
                remove(v);
                insert(v, nk);
            }
        }
    }
}


int preprocess(const kp_graph *g, int k, int q, int truss, Prep *out)
{
    std::vector<uint32_t> o1;
    const double t0 = now_ms();
    int rc = core_peel(g, q - k, &out->g, o1);
    if (rc) return rc;
    out->core_ms = now_ms() - t0;
    return finish_preprocess(k, q, truss, o1, nullptr, out);
}

int finish_preprocess(int k, int q, int truss, std::vector<uint32_t> &o1, std::vector<int32_t> *tri, Prep *out)
{
    const int bd = q - k;
    int rc;
    out->pn_first = out->g.n;
    if (truss) {
        double t = now_ms();
        std::vector<int32_t> own;
        if (!tri) {
            truss_supports(&out->g, own);
            tri = &own;
        }
        out->support_ms += now_ms() - t;
        t = now_ms();
        rc = truss_peel(&out->g, q, k, *tri);
        if (rc) return rc;
        out->truss_ms = now_ms() - t;
        kp_graph g2{};
        std::vector<uint32_t> o2;
        rc = core_peel(&out->g, bd, &g2, o2);
        if (rc) return rc;
        for (auto &x : o2) x = o1[x];
        o1.swap(o2);
        kp_graph_free(&out->g);
        out->g = g2;
    }
    out->orig.swap(o1);
    const double t = now_ms();
    out->d1max = 0;
    degeneracy_order(&out->g, out->dseq, out->dpos, &out->d1max);
    out->maxdeg = 0;
    for (uint32_t u = 0; u < out->g.n; u++) out->maxdeg = std::max(out->maxdeg, out->g.deg[u]);
    out->sizes_known = true;
    out->degeneracy_ms = now_ms() - t;
    return KP_OK;
}

} // namespace kp
