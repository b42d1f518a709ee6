// kp_gprep.hip -- preprocessing with its parallel stages on the GPU: the
// (q-k)-core peel, the stable relabel to the survivors' ranks and the
// per-entry truss supports.  The stages whose result depends on a sequential
// order stay on the host: the one-sided truss peel (inc/host_funcs.h:463-559)
// and the LIFO bucket-heap degeneracy order (inc/degeneracy.h:23-203).
// The result is the same Prep as the host preprocessing, bit for bit.
//
//   core_init_kernel   residual degrees, alive flags, first frontier
//   core_peel_kernel   one wave per frontier vertex: decrements each alive
//                      neighbour; the decrement that takes a vertex from bd
//                      to bd-1 kills it and puts it on the next frontier
//                      (the cascade of inc/host_funcs.h:16-63; the bd-core is
//                      unique, so the frontier order does not matter)
//   count_kernel       degree of every survivor among survivors
//   fill_kernel        one wave per survivor: its surviving neighbours, in
//                      list order, as new ids (ballot compaction)
//   support_kernel     one wave per vertex u, for each neighbour v the lanes
//                      binary-search v's list in u's: |N(u) ∩ N(v)|
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "kp_device.h"
#include "kp_engine.h"

#include <algorithm>
#include <cstring>
#include <vector>

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) {                                                        \
            kp::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
            return KP_EHIP;                                                            \
        }                                                                              \
    } while (0)

namespace kpg {

__device__ __forceinline__ uint64_t lt_mask() { return (1ull << __lane_id()) - 1ull; }

__global__ void core_init_kernel(uint32_t n, const uint32_t *deg, int bd, int32_t *res, uint32_t *alive,
                                 uint32_t *frontier, uint32_t *nfront)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    res[v] = int32_t(deg[v]);
    const bool a = int64_t(deg[v]) >= bd;
    alive[v] = a;
    if (!a) frontier[atomicAdd(nfront, 1u)] = v;
}

__global__ void core_peel_kernel(const uint64_t *off, const uint32_t *nbr, int bd, int32_t *res, uint32_t *alive,
                                 const uint32_t *fin, uint32_t nin, uint32_t *fout, uint32_t *nout)
{
    const uint32_t ln = __lane_id();
    const uint32_t nw = gridDim.x * (blockDim.x >> 6);
    for (uint32_t i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < nin; i += nw) {
        const uint32_t u = fin[i];
        for (uint64_t p = off[u] + ln; p < off[u + 1]; p += 64) {
            const uint32_t v = nbr[p];
            if (__hip_atomic_load(&alive[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                const int32_t old = atomicSub(&res[v], 1);
                if (old == bd) {
                    __hip_atomic_store(&alive[v], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    fout[atomicAdd(nout, 1u)] = v;
                }
            }
        }
    }
}

// new degree of every survivor (ndeg indexed by the new id = rank[v])
__global__ void count_kernel(uint32_t n, const uint64_t *off, const uint32_t *nbr, const uint32_t *alive,
                             const uint32_t *rank, uint32_t *ndeg, uint32_t *orig)
{
    const uint32_t ln = __lane_id();
    const uint32_t nw = gridDim.x * (blockDim.x >> 6);
    for (uint32_t v = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); v < n; v += nw) {
        if (!alive[v]) continue;
        uint32_t c = 0;
        for (uint64_t p = off[v] + ln; p < off[v + 1]; p += 64) c += alive[nbr[p]];
        for (int s = 32; s > 0; s >>= 1) c += __shfl_xor(c, s);
        if (ln == 0) {
            ndeg[rank[v]] = c;
            orig[rank[v]] = v;
        }
    }
}

// surviving neighbours of every survivor in list order, as new ids
__global__ void fill_kernel(uint32_t n, const uint64_t *off, const uint32_t *nbr, const uint32_t *alive,
                            const uint32_t *rank, const uint64_t *noff, uint32_t *nnbr)
{
    const uint32_t ln = __lane_id();
    const uint32_t nw = gridDim.x * (blockDim.x >> 6);
    for (uint32_t v = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); v < n; v += nw) {
        if (!alive[v]) continue;
        uint64_t w = noff[rank[v]];
        for (uint64_t base = off[v]; base < off[v + 1]; base += 64) {
            const uint64_t p = base + ln;
            const bool in = p < off[v + 1];
            const uint32_t x = in ? nbr[p] : 0u;
            const bool keep = in && alive[x];
            const uint64_t m = __ballot(keep);
            if (keep) nnbr[w + __popcll(m & lt_mask())] = rank[x];
            w += __popcll(m);
        }
    }
}

// |N(u) ∩ N(v)| for every directed entry (u, v) of a graph with sorted lists
__global__ void support_kernel(uint32_t n, const uint64_t *off, const uint32_t *nbr, int32_t *tri)
{
    const uint32_t ln = __lane_id();
    const uint32_t nw = gridDim.x * (blockDim.x >> 6);
    for (uint32_t u = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); u < n; u += nw) {
        const uint64_t ub = off[u], ue = off[u + 1];
        for (uint64_t p = ub; p < ue; p++) {
            const uint32_t v = nbr[p];
            const uint64_t vb = off[v], ve = off[v + 1];
            uint32_t c = 0;
            for (uint64_t t = vb + ln; t < ve; t += 64) {
                const uint32_t x = nbr[t];
                uint64_t lo = ub, hi = ue;   // lower bound of x in u's list
                while (lo < hi) {
                    const uint64_t mid = (lo + hi) >> 1;
                    if (nbr[mid] < x) lo = mid + 1;
                    else hi = mid;
                }
                c += (lo < ue && nbr[lo] == x);
            }
            for (int s = 32; s > 0; s >>= 1) c += __shfl_xor(c, s);
            if (ln == 0) tri[p] = int32_t(c);
        }
    }
}

} // namespace kpg

namespace kp {

namespace {

template <class T>
int exclusive_sum(const T *in, T *out, uint32_t n, DevBuf &tmp, hipStream_t st)
{
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, st));
    int rc = tmp.ensure(std::max<size_t>(bytes, 256));
    if (rc) return rc;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, n, st));
    return KP_OK;
}

} // namespace

int gpu_preprocess(const kp_graph *g, int k, int q, int truss, Prep *out)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        set_error("no HIP device visible: the HIP backend is required");
        return KP_ENODEV;
    }
    int cus = 0, dev = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const double t0 = now_ms();
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const int bd = q - k;
    if (n == 0) return preprocess(g, k, q, truss, out);
    hipStream_t st = nullptr;
    DevBuf d_off, d_nbr, d_deg, d_res, d_alive, d_f0, d_f1, d_cnt, d_rank, d_orig, d_noff, d_nnbr, d_tri, tmp;
    int rc;
    if ((rc = d_off.ensure(8 * (size_t(n) + 1))) || (rc = d_nbr.ensure(4 * std::max<uint64_t>(m, 1))) ||
        (rc = d_deg.ensure(4 * size_t(n) + 4)) || (rc = d_res.ensure(4 * size_t(n) + 4)) ||
        (rc = d_alive.ensure(4 * size_t(n) + 4)) || (rc = d_f0.ensure(4 * size_t(n) + 4)) ||
        (rc = d_f1.ensure(4 * size_t(n) + 4)) || (rc = d_cnt.ensure(256)) || (rc = d_rank.ensure(4 * size_t(n) + 4)))
        return rc;
    HIPCHK(hipMemcpyAsync(d_off.p, g->off, 8 * (size_t(n) + 1), hipMemcpyHostToDevice, st));
    if (m) HIPCHK(hipMemcpyAsync(d_nbr.p, g->nbr, 4 * m, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_deg.p, g->deg, 4 * size_t(n), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(d_cnt.p, 0, 256, st));
    uint32_t *cnt = d_cnt.as<uint32_t>();
    const uint32_t tb = 256, waves_grid = uint32_t(std::max(1, cus) * 16);
    if (n) {
        hipLaunchKernelGGL(kpg::core_init_kernel, dim3((n + tb - 1) / tb), dim3(tb), 0, st, n, d_deg.as<uint32_t>(), bd,
                           d_res.as<int32_t>(), d_alive.as<uint32_t>(), d_f0.as<uint32_t>(), cnt);
        HIPCHK(hipGetLastError());
    }
    // frontier rounds: cnt[0] holds the current frontier's size, cnt[1] the next one's
    uint32_t nf = 0;
    HIPCHK(hipMemcpyAsync(&nf, cnt, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    DevBuf *fa = &d_f0, *fb = &d_f1;
    while (nf) {
        HIPCHK(hipMemsetAsync(cnt + 1, 0, 4, st));
        const uint32_t grid = std::min<uint32_t>((nf + 3) / 4, waves_grid / 4 + 1);
        hipLaunchKernelGGL(kpg::core_peel_kernel, dim3(grid), dim3(256), 0, st, d_off.as<uint64_t>(), d_nbr.as<uint32_t>(),
                           bd, d_res.as<int32_t>(), d_alive.as<uint32_t>(), fa->as<uint32_t>(), nf, fb->as<uint32_t>(),
                           cnt + 1);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(&nf, cnt + 1, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        std::swap(fa, fb);
    }
    // stable relabel: new id = rank among the survivors
    if ((rc = exclusive_sum(d_alive.as<uint32_t>(), d_rank.as<uint32_t>(), n + 0, tmp, st))) return rc;
    uint32_t last_rank = 0, last_alive = 0;
    if (n) {
        HIPCHK(hipMemcpyAsync(&last_rank, d_rank.as<uint32_t>() + n - 1, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(&last_alive, d_alive.as<uint32_t>() + n - 1, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    const uint32_t pn = n ? last_rank + last_alive : 0;
    if ((rc = d_orig.ensure(4 * size_t(pn) + 4)) || (rc = d_noff.ensure(8 * (size_t(pn) + 1)))) return rc;
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>((n + 3) / 4, waves_grid / 4));
    if (n) {
        // the residual degrees are done with: d_res takes the new degrees
        hipLaunchKernelGGL(kpg::count_kernel, dim3(grid), dim3(256), 0, st, n, d_off.as<uint64_t>(), d_nbr.as<uint32_t>(),
                           d_alive.as<uint32_t>(), d_rank.as<uint32_t>(), d_res.as<uint32_t>(), d_orig.as<uint32_t>());
        HIPCHK(hipGetLastError());
    }
    // new offsets by a host scan of the new degrees
    std::vector<uint32_t> ndeg(pn);
    if (pn) HIPCHK(hipMemcpyAsync(ndeg.data(), d_res.p, 4 * size_t(pn), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<uint64_t> noff(size_t(pn) + 1, 0);
    for (uint32_t i = 0; i < pn; i++) noff[i + 1] = noff[i] + ndeg[i];
    const uint64_t pm = noff[pn];
    if ((rc = d_nnbr.ensure(4 * std::max<uint64_t>(pm, 1)))) return rc;
    HIPCHK(hipMemcpyAsync(d_noff.p, noff.data(), 8 * (size_t(pn) + 1), hipMemcpyHostToDevice, st));
    if (n) {
        hipLaunchKernelGGL(kpg::fill_kernel, dim3(grid), dim3(256), 0, st, n, d_off.as<uint64_t>(), d_nbr.as<uint32_t>(),
                           d_alive.as<uint32_t>(), d_rank.as<uint32_t>(), d_noff.as<uint64_t>(), d_nnbr.as<uint32_t>());
        HIPCHK(hipGetLastError());
    }
    std::vector<int32_t> tri;
    if (truss && pn) {
        if ((rc = d_tri.ensure(4 * std::max<uint64_t>(pm, 1)))) return rc;
        const uint32_t g2 = std::max<uint32_t>(1, std::min<uint32_t>((pn + 3) / 4, waves_grid / 4));
        hipLaunchKernelGGL(kpg::support_kernel, dim3(g2), dim3(256), 0, st, pn, d_noff.as<uint64_t>(),
                           d_nnbr.as<uint32_t>(), d_tri.as<int32_t>());
        HIPCHK(hipGetLastError());
        tri.resize(pm);
    }
    // the first core back on the host, for the sequential stages
    if ((rc = graph_alloc(&out->g, pn, pm))) return rc;
    std::vector<uint32_t> o1(pn);
    std::memcpy(out->g.off, noff.data(), 8 * (size_t(pn) + 1));
    std::memcpy(out->g.deg, ndeg.data(), 4 * size_t(pn));
    if (pm) HIPCHK(hipMemcpyAsync(out->g.nbr, d_nnbr.p, 4 * pm, hipMemcpyDeviceToHost, st));
    if (pn) HIPCHK(hipMemcpyAsync(o1.data(), d_orig.p, 4 * size_t(pn), hipMemcpyDeviceToHost, st));
    if (!tri.empty()) HIPCHK(hipMemcpyAsync(tri.data(), d_tri.p, 4 * pm, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    out->core_ms = now_ms() - t0;
    return finish_preprocess(k, q, truss, o1, truss ? &tri : nullptr, out);
}

} // namespace kp
