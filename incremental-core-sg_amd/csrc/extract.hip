// MST extraction from a device-resident CORE-SG.
//
// Replaces, on device:
//   * KnnGraph::undirected_edges / collect_undirected (src/knn_graph.cpp:127-200):
//     the deduplicated k-NN half, built once and kept resident, sorted by (u,v);
//   * CoreSG::core_distances + edges(mpts) (coresg.cpp:68-70, 391-406;
//     knn_graph.hpp:157-163; knn_graph.cpp:219-235): the MRD reweight, fused
//     into the first Boruvka round;
//   * mst_from_edges (src/mst.cpp:61-156): Boruvka rounds with a 64-bit
//     atomic-min of (weight bits << 32 | rank). Rank is the position in the
//     (u, v)-sorted list, so the packed minimum orders edges exactly as
//     edge_less (types.hpp:38-42) does, and the tree found is the unique MST
//     under that order -- the reference Kruskal tree, edge for edge.
//
// The total weight is the reference's double sum in Kruskal order
// (types.hpp:58-62). It is computed exactly: when every MST weight is a
// multiple of 2^e and the sum is below 2^(53+e), every partial sum of the
// sequential double accumulation is representable, so the sequential result
// equals the exact integer sum, which the device accumulates in u64. The rare
// case that fails the test falls back to a sorted sequential sum on the host.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/coresg_hip.h"
#include "graph.hpp"
#include "mrd_key.hpp"
#include "scan.hpp"

namespace coresg_hip {
namespace {

constexpr unsigned kThreads = 256;

// ---------------------------------------------------------------------------
// Edge-list construction
// ---------------------------------------------------------------------------

// One thread per (p, slot j). Node p owns (p,q) if q > p, or if q < p and p is
// not in q's list; the membership scan is skipped when p's distance exceeds
// q's last stored distance (knn_graph.cpp:143-166).
__global__ void __launch_bounds__(kThreads)
    emit_count_kernel(const std::uint32_t* __restrict__ ids, const float* __restrict__ dist,
                      const std::uint32_t* __restrict__ count, std::uint32_t n, std::uint32_t k,
                      std::uint8_t* __restrict__ emit, std::uint32_t* __restrict__ deg,
                      unsigned long long* __restrict__ total_out) {
    const std::size_t total = std::size_t(n) * k;
    unsigned long long local = 0;
    for (std::size_t gid = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; gid < total;
         gid += std::size_t(gridDim.x) * blockDim.x) {
        const std::uint32_t p = static_cast<std::uint32_t>(gid / k);
        const std::uint32_t j = static_cast<std::uint32_t>(gid - std::size_t(p) * k);
        std::uint8_t e = 0;
        if (j < count[p]) {
            const std::uint32_t q = ids[gid];
            if (q < n && q != p) {
                bool mutual = false;
                if (q < p) {
                    const std::uint32_t cq = count[q];
                    const std::size_t qb = std::size_t(q) * k;
                    if (cq > 0 && !(dist[gid] > dist[qb + cq - 1])) {
                        std::uint32_t t = 0;
                        if ((k & 3u) == 0) {  // rows 16-byte aligned: four ids per load
                            const uint4* row = reinterpret_cast<const uint4*>(ids + qb);
                            for (; t + 4 <= cq && !mutual; t += 4) {
                                const uint4 x = row[t >> 2];
                                mutual = x.x == p || x.y == p || x.z == p || x.w == p;
                            }
                        }
                        for (; t < cq && !mutual; ++t) mutual = ids[qb + t] == p;
                    }
                }
                if (!mutual) {
                    e = 1;
                    ++local;
                    if (deg) atomicAdd(&deg[p < q ? p : q], 1u);
                }
            }
        }
        if (emit) emit[gid] = e;
    }
    if (total_out) {
        local = wave_sum(local);
        if (__lane_id() == 0 && local) atomicAdd(total_out, local);
    }
}

// Bridge edges (bu < bv) counted at both endpoints: deg1 as first endpoint
// (edge list segments), deg2 as second endpoint (the adjacency's rest).
__global__ void bridge_count_kernel(const std::uint32_t* __restrict__ bu,
                                    const std::uint32_t* __restrict__ bv, std::uint32_t m,
                                    std::uint32_t n, std::uint32_t* __restrict__ deg1,
                                    std::uint32_t* __restrict__ deg2) {
    const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m && bu[i] < n) {
        atomicAdd(&deg1[bu[i]], 1u);
        atomicAdd(&deg2[bv[i]], 1u);
    }
}

// Scatter emitted slots into per-u segments (unordered within a segment);
// `src` records a unique origin tag used as the final tie-break. Half a
// wavefront per node p: p's own edges (q > p, u = p) take the first positions
// of p's segment by a ballot prefix count, no atomic; edges that go to q's
// segment (q < p) are placed by `cursor[q]`, which starts at q's own count.
__global__ void __launch_bounds__(kThreads)
    scatter_kernel(const std::uint32_t* __restrict__ ids, const float* __restrict__ dist,
                   std::uint32_t n, std::uint32_t k, const std::uint8_t* __restrict__ emit,
                   const std::uint32_t* __restrict__ off, std::uint32_t* __restrict__ cursor,
                   uint4* __restrict__ tmp) {
    constexpr std::uint32_t kSeg = 32;
    const std::uint32_t lane = __lane_id(), sub = lane % kSeg, half = lane / kSeg;
    const std::uint32_t segs = gridDim.x * (blockDim.x / kSeg);
    const std::uint32_t below = (1u << sub) - 1u;
    for (std::uint32_t p = (blockIdx.x * blockDim.x + threadIdx.x) / kSeg; p < n; p += segs) {
        const std::uint32_t pos0 = off[p];
        std::uint32_t own_before = 0;
        for (std::uint32_t j0 = 0; j0 < k; j0 += kSeg) {
            const std::uint32_t j = j0 + sub;
            const std::size_t gid = std::size_t(p) * k + j;
            const bool e = j < k && emit[gid];
            const std::uint32_t q = e ? ids[gid] : 0u;
            const bool own = e && q > p;
            const std::uint32_t mask =
                static_cast<std::uint32_t>(__ballot(own) >> (half * kSeg));
            if (e) {
                const std::uint32_t pos =
                    own ? pos0 + own_before + __builtin_popcount(mask & below)
                        : off[q] + atomicAdd(&cursor[q], 1u);
                tmp[pos] = make_uint4(own ? q : p, static_cast<std::uint32_t>(gid),
                                      __float_as_uint(dist[gid]), 0u);
            }
            own_before += __builtin_popcount(mask);
        }
    }
}

__global__ void bridge_scatter_kernel(const std::uint32_t* __restrict__ bu,
                                      const std::uint32_t* __restrict__ bv,
                                      const float* __restrict__ bw, std::uint32_t m,
                                      std::uint32_t n, std::uint32_t src_base,
                                      const std::uint32_t* __restrict__ off,
                                      std::uint32_t* __restrict__ cursor,
                                      uint4* __restrict__ tmp) {
    const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || bu[i] >= n) return;
    const std::uint32_t u = bu[i];
    const std::uint32_t pos = off[u] + atomicAdd(&cursor[u], 1u);
    tmp[pos] = make_uint4(bv[i], src_base + i, __float_as_uint(bw[i]), 0u);
}

// One wavefront per u: rank each entry of u's segment by (v, src) and write it
// to its final position e in the sorted list, and to both endpoints' entries of
// the symmetric adjacency: u's first-endpoint entries keep the sorted order
// (position adj_off[u] + rank), v's second-endpoint entries follow v's own
// segment, placed by a cursor (`cursor[v]` starts at adj_off[v] + deg1[v]).
// The scattered entries are one 16-byte record each, (v, src, weight bits, 0),
// so the scatter writes one transaction per entry.
__global__ void __launch_bounds__(kThreads)
    segment_sort_kernel(std::uint32_t n, const std::uint32_t* __restrict__ off,
                        const uint4* __restrict__ tmp, const std::uint32_t* __restrict__ adj_off,
                        std::uint32_t* __restrict__ cursor, std::uint32_t* __restrict__ eu,
                        std::uint32_t* __restrict__ ev, float* __restrict__ ew,
                        AdjEntry* __restrict__ adj) {
    const std::uint32_t waves_per_block = blockDim.x / kWave;
    const std::uint32_t lane = threadIdx.x & (kWave - 1);
    for (std::uint32_t u = blockIdx.x * waves_per_block + threadIdx.x / kWave; u < n;
         u += gridDim.x * waves_per_block) {
        const std::uint32_t s = off[u], L = off[u + 1] - s;
        // Each lane ranks one entry of a 64-entry chunk against every chunk of
        // the segment, broadcast from registers by readlane (segments of
        // hundreds of entries exist: the k-NN graph's hubs).
        for (std::uint32_t i0 = 0; i0 < L; i0 += kWave) {
            const bool has = i0 + lane < L;
            const uint4 t =
                has ? tmp[s + i0 + lane] : make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u);
            const std::uint32_t myv = t.x, mys = t.y;
            std::uint32_t rank = 0;
            for (std::uint32_t j0 = 0; j0 < L; j0 += kWave) {
                const std::uint32_t jn = (L - j0 < kWave) ? L - j0 : kWave;
                std::uint32_t cv = myv, cs = mys;
                if (j0 != i0 && lane < jn) {
                    const uint2 o = *reinterpret_cast<const uint2*>(tmp + s + j0 + lane);
                    cv = o.x;
                    cs = o.y;
                }
                for (std::uint32_t j = 0; j < jn; ++j) {  // j uniform: a readlane each
                    const std::uint32_t ov = static_cast<std::uint32_t>(
                        __builtin_amdgcn_readlane(static_cast<int>(cv), static_cast<int>(j)));
                    const std::uint32_t os = static_cast<std::uint32_t>(
                        __builtin_amdgcn_readlane(static_cast<int>(cs), static_cast<int>(j)));
                    rank += (ov < myv || (ov == myv && os < mys)) ? 1u : 0u;
                }
            }
            if (has) {
                const std::uint32_t e = s + rank;
                const float w = __uint_as_float(t.z);
                eu[e] = u;
                ev[e] = myv;
                ew[e] = w;
                adj[adj_off[u] + rank] = AdjEntry{myv, e, w};
                adj[atomicAdd(&cursor[myv], 1u)] = AdjEntry{u | kSecondEnd, e, w};
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Per-mpts sweep
// ---------------------------------------------------------------------------

// core_mpts(v) = dist[v*k + min(mpts, count) - 1], +inf for an empty list
// (knn_graph.hpp:157-163). Also resets the Boruvka state.
__global__ void core_init_kernel(const float* __restrict__ dist,
                                 const std::uint32_t* __restrict__ count, std::uint32_t n,
                                 std::uint32_t k, std::uint32_t mpts, float* __restrict__ core,
                                 std::uint32_t* __restrict__ comp,
                                 std::uint32_t* __restrict__ parent,
                                 std::uint64_t* __restrict__ best) {
    for (std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += gridDim.x * blockDim.x) {
        const std::uint32_t c = count[v];
        float cd = __builtin_inff();
        if (c > 0) {
            std::uint32_t pos = mpts - 1;
            if (pos >= c) pos = c - 1;
            cd = dist[std::size_t(v) * k + pos];
        }
        core[v] = cd;
        comp[v] = v;
        parent[v] = v;
        best[v] = kNone64;
    }
}

// ---- symmetric adjacency (built once per CORE-SG) --------------------------

// Adds every vertex's degree as first endpoint: its segment of the sorted list.
__global__ void __launch_bounds__(kThreads)
    first_degree_kernel(const std::uint32_t* __restrict__ eoff, std::uint32_t n,
                        std::uint32_t* __restrict__ deg) {
    for (std::uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < n;
         x += gridDim.x * blockDim.x)
        deg[x] += eoff[x + 1] - eoff[x];
}

// Second-endpoint entries of vertex x start after its first-endpoint segment.
__global__ void __launch_bounds__(kThreads)
    adj_base_kernel(const std::uint32_t* __restrict__ eoff, const std::uint32_t* __restrict__ off,
                    std::uint32_t n, std::uint32_t* __restrict__ cursor) {
    for (std::uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < n;
         x += gridDim.x * blockDim.x)
        cursor[x] = off[x] + (eoff[x + 1] - eoff[x]);
}

__global__ void __launch_bounds__(kThreads)
    reweight_kernel(const std::uint32_t* __restrict__ eu, const std::uint32_t* __restrict__ ev,
                    const float* __restrict__ ew, const float* __restrict__ core, std::uint64_t m,
                    float* __restrict__ out) {
    for (std::uint64_t e = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
         e += std::uint64_t(gridDim.x) * blockDim.x)
        out[e] = fmax_ref(ew[e], fmax_ref(core[eu[e]], core[ev[e]]));
}

}  // namespace

void check_mpts(const Graph& g, std::uint32_t* mpts) {
    if (*mpts == 0) *mpts = 1;  // coresg.cpp:392
    if (*mpts > g.k) throw std::runtime_error("CoreSG::edges: mpts exceeds k_max");
}

namespace {
// Ownership flags of every slot (g.emit_flags), optionally their total;
// valid until the k-NN lists change.
void emit_state(Graph& g, std::uint64_t* d_total) {
    const std::size_t slots = std::size_t(g.n) * g.k;
    g.emit_flags.reserve(slots);
    emit_count_kernel<<<capped_blocks(slots, kThreads), kThreads, 0, g.stream>>>(
        g.ids.p, g.dist.p, g.count.p, g.n, g.k, g.emit_flags.p, nullptr,
        reinterpret_cast<unsigned long long*>(d_total));
    CSG_LAUNCH_CHECK();
    g.emit_ready = true;
}

// Owned-edge counts from the ownership flags, per first endpoint u = min(p, q)
// (deg1: the edge list's segments) and per second endpoint v = max(p, q)
// (deg2: the adjacency's second-endpoint entries): one thread per node p
// counts its own side of each owned slot itself and adds the other side to q.
// own[p]: p's edges from its own list with p as first endpoint.
__global__ void __launch_bounds__(kThreads)
    emit_deg_kernel(const std::uint32_t* __restrict__ ids, const std::uint8_t* __restrict__ emit,
                    std::uint32_t n, std::uint32_t k, std::uint32_t* __restrict__ deg1,
                    std::uint32_t* __restrict__ deg2, std::uint32_t* __restrict__ own) {
    for (std::uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += gridDim.x * blockDim.x) {
        std::uint32_t first = 0, second = 0;
        for (std::uint32_t j = 0; j < k; ++j) {
            const std::size_t gid = std::size_t(p) * k + j;
            if (!emit[gid]) continue;
            const std::uint32_t q = ids[gid];
            if (q > p) {
                ++first;
                atomicAdd(&deg2[q], 1u);
            } else {
                ++second;
                atomicAdd(&deg1[q], 1u);
            }
        }
        if (first) atomicAdd(&deg1[p], first);
        if (second) atomicAdd(&deg2[p], second);
        own[p] = first;
    }
}
}  // namespace

void prepare_edges(Graph& g) {
    if (g.edges_ready) return;
    if (!g.knn_set) throw std::runtime_error("prepare_edges: k-NN lists not set");
    const std::uint32_t n = g.n, k = g.k;
    const std::size_t slots = std::size_t(n) * k;
    const std::uint32_t nb = static_cast<std::uint32_t>(g.bu.size());
    if (slots + nb >= 0x7FFFFFFFull)
        throw std::runtime_error("edge list exceeds 2^31 entries");
    if (n >= kSecondEnd) throw std::runtime_error("point count exceeds 2^31");
    hipStream_t s = g.stream;

    // ownership flags from the build's edge count when the lists are unchanged
    if (!g.emit_ready) emit_state(g, nullptr);
    g.eoff.reserve(std::size_t(n) + 1);
    fill_u32(g.eoff.p, std::size_t(n) + 1, 0u, s);
    g.adj_off.reserve(std::size_t(n) + 1);
    fill_u32(g.adj_off.p, std::size_t(n) + 1, 0u, s);
    // cursor: each first endpoint's next free position in its segment, past
    // the edges from its own list
    DBuf<std::uint32_t> cursor;
    cursor.reserve(n);
    emit_deg_kernel<<<capped_blocks(n, kThreads), kThreads, 0, s>>>(
        g.ids.p, g.emit_flags.p, n, k, g.eoff.p, g.adj_off.p, cursor.p);
    CSG_LAUNCH_CHECK();

    DBuf<std::uint32_t> dbu, dbv;
    DBuf<float> dbw;
    if (nb) {
        dbu.reserve(nb);
        dbv.reserve(nb);
        dbw.reserve(nb);
        CSG_CHECK(hipMemcpyAsync(dbu.p, g.bu.data(), nb * 4, hipMemcpyHostToDevice, s));
        CSG_CHECK(hipMemcpyAsync(dbv.p, g.bv.data(), nb * 4, hipMemcpyHostToDevice, s));
        CSG_CHECK(hipMemcpyAsync(dbw.p, g.bw.data(), nb * 4, hipMemcpyHostToDevice, s));
        bridge_count_kernel<<<blocks_for(nb, kThreads), kThreads, 0, s>>>(dbu.p, dbv.p, nb, n,
                                                                           g.eoff.p, g.adj_off.p);
        CSG_LAUNCH_CHECK();
    }

    // CSR offsets; off[n] = |E|.
    g.ensure_scan_tmp(std::size_t(n) + 1);
    g.counters.reserve(16);
    exclusive_scan_u32(g.eoff.p, std::size_t(n) + 1, g.counters.p, g.scan_tmp.p, s);
    std::uint32_t total = 0;
    CSG_CHECK(hipMemcpyAsync(&total, g.counters.p, 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));
    const std::uint64_t m = total;

    DBuf<uint4> tmp;  // (v, src, weight bits, 0) per edge, grouped by u
    tmp.reserve(m);
    scatter_kernel<<<capped_blocks(std::size_t(n) * 32, kThreads), kThreads, 0, s>>>(
        g.ids.p, g.dist.p, n, k, g.emit_flags.p, g.eoff.p, cursor.p, tmp.p);
    CSG_LAUNCH_CHECK();
    if (nb) {
        bridge_scatter_kernel<<<blocks_for(nb, kThreads), kThreads, 0, s>>>(
            dbu.p, dbv.p, dbw.p, nb, n, static_cast<std::uint32_t>(slots), g.eoff.p, cursor.p,
            tmp.p);
        CSG_LAUNCH_CHECK();
    }
    // Symmetric adjacency offsets: each vertex's first-endpoint segment plus
    // its second-endpoint count (already in adj_off); the scatter is done with
    // `cursor`, which now holds each vertex's second-endpoint start.
    first_degree_kernel<<<capped_blocks(n, kThreads), kThreads, 0, s>>>(g.eoff.p, n, g.adj_off.p);
    CSG_LAUNCH_CHECK();
    exclusive_scan_u32(g.adj_off.p, std::size_t(n) + 1, nullptr, g.scan_tmp.p, s);
    adj_base_kernel<<<capped_blocks(n, kThreads), kThreads, 0, s>>>(g.eoff.p, g.adj_off.p, n,
                                                                     cursor.p);
    CSG_LAUNCH_CHECK();
    g.eu.reserve(m);
    g.ev.reserve(m);
    g.ew.reserve(m);
    g.adj.reserve(2 * m);
    segment_sort_kernel<<<capped_blocks(std::size_t(n) * kWave, kThreads, 16384), kThreads, 0,
                          s>>>(n, g.eoff.p, tmp.p, g.adj_off.p, cursor.p, g.eu.p, g.ev.p, g.ew.p,
                               g.adj.p);
    CSG_LAUNCH_CHECK();
    CSG_CHECK(hipStreamSynchronize(s));
    g.num_edges = m;
    std::uint64_t nbv = 0;
    for (std::uint32_t i = 0; i < nb; ++i) nbv += (g.bu[i] < n) ? 1 : 0;
    g.knn_edges = m - nbv;
    g.edges_ready = true;
}

// Undirected k-NN edge count (KnnGraph::undirected_edge_count,
// knn_graph.cpp:202-217) without materialising the list.
std::uint64_t count_knn_edges(Graph& g) {
    g.counters64.reserve(4);
    std::uint64_t* d_total = g.counters64.p + 3;
    fill_u64(d_total, 1, 0ull, g.stream);
    // the ownership state is kept for prepare_edges
    emit_state(g, d_total);
    std::uint64_t total = 0;
    CSG_CHECK(hipMemcpyAsync(&total, d_total, 8, hipMemcpyDeviceToHost, g.stream));
    CSG_CHECK(hipStreamSynchronize(g.stream));
    return total;
}

namespace {
// One block per instance: the sum of its nodes' owned-edge counts.
__global__ void __launch_bounds__(kThreads)
    segment_sum_kernel(const std::uint32_t* __restrict__ deg,
                       const std::uint32_t* __restrict__ off,
                       unsigned long long* __restrict__ out) {
    __shared__ unsigned long long s_w[kThreads / 64];
    const std::uint32_t lo = off[blockIdx.x], hi = off[blockIdx.x + 1];
    unsigned long long acc = 0;
    for (std::uint32_t v = lo + threadIdx.x; v < hi; v += blockDim.x) acc += deg[v];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63u) == 0) s_w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (unsigned w = 0; w < blockDim.x / 64; ++w) t += s_w[w];
        out[blockIdx.x] = t;
    }
}
}  // namespace

std::vector<std::uint64_t> knn_edges_per_instance(Graph& g) {
    const std::uint32_t ninst = g.num_instances();
    if (g.inst_off.empty()) return {count_knn_edges(g)};
    const std::size_t slots = std::size_t(g.n) * g.k;
    DBuf<std::uint32_t> deg;
    DBuf<unsigned long long> out;
    deg.reserve(g.n);
    out.reserve(ninst);
    fill_u32(deg.p, g.n, 0u, g.stream);
    // every owned edge counted at its smaller endpoint, which lies in the
    // same instance as the other one
    emit_count_kernel<<<capped_blocks(slots, kThreads), kThreads, 0, g.stream>>>(
        g.ids.p, g.dist.p, g.count.p, g.n, g.k, nullptr, deg.p, nullptr);
    CSG_LAUNCH_CHECK();
    segment_sum_kernel<<<ninst, kThreads, 0, g.stream>>>(deg.p, g.inst_off_d.p, out.p);
    CSG_LAUNCH_CHECK();
    std::vector<std::uint64_t> h(ninst);
    CSG_CHECK(hipMemcpyAsync(h.data(), out.p, std::size_t(ninst) * 8, hipMemcpyDeviceToHost,
                             g.stream));
    CSG_CHECK(hipStreamSynchronize(g.stream));
    return h;
}

void reweighted_edges(Graph& g, std::uint32_t mpts, std::uint32_t* u, std::uint32_t* v,
                      float* w) {
    check_mpts(g, &mpts);
    prepare_edges(g);
    const std::uint32_t n = g.n;
    hipStream_t s = g.stream;
    g.core.reserve(n);
    g.comp.reserve(n);
    g.parent.reserve(n);
    g.best.reserve(n);
    core_init_kernel<<<capped_blocks(n, kThreads), kThreads, 0, s>>>(
        g.dist.p, g.count.p, n, g.k, mpts, g.core.p, g.comp.p, g.parent.p, g.best.p);
    CSG_LAUNCH_CHECK();
    DBuf<float> w2;
    w2.reserve(g.num_edges);
    reweight_kernel<<<capped_blocks(g.num_edges, kThreads), kThreads, 0, s>>>(
        g.eu.p, g.ev.p, g.ew.p, g.core.p, g.num_edges, w2.p);
    CSG_LAUNCH_CHECK();
    const std::size_t m = g.num_edges;
    if (u) CSG_CHECK(hipMemcpyAsync(u, g.eu.p, m * 4, hipMemcpyDeviceToHost, s));
    if (v) CSG_CHECK(hipMemcpyAsync(v, g.ev.p, m * 4, hipMemcpyDeviceToHost, s));
    if (w) CSG_CHECK(hipMemcpyAsync(w, w2.p, m * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));
}

}  // namespace coresg_hip
