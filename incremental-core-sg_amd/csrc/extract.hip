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
                        for (std::uint32_t t = 0; t < cq; ++t) {
                            if (ids[qb + t] == p) {
                                mutual = true;
                                break;
                            }
                        }
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

__global__ void bridge_count_kernel(const std::uint32_t* __restrict__ bu, std::uint32_t m,
                                    std::uint32_t n, std::uint32_t* __restrict__ deg) {
    const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m && bu[i] < n) atomicAdd(&deg[bu[i]], 1u);
}

// Scatter emitted slots and bridges into per-u segments (unordered within a
// segment); `src` records a unique origin tag used as the final tie-break.
__global__ void __launch_bounds__(kThreads)
    scatter_kernel(const std::uint32_t* __restrict__ ids, const float* __restrict__ dist,
                   std::uint32_t n, std::uint32_t k, const std::uint8_t* __restrict__ emit,
                   const std::uint32_t* __restrict__ off, std::uint32_t* __restrict__ cursor,
                   std::uint32_t* __restrict__ tv, float* __restrict__ tw,
                   std::uint32_t* __restrict__ tsrc) {
    const std::size_t total = std::size_t(n) * k;
    for (std::size_t gid = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; gid < total;
         gid += std::size_t(gridDim.x) * blockDim.x) {
        if (!emit[gid]) continue;
        const std::uint32_t p = static_cast<std::uint32_t>(gid / k);
        const std::uint32_t q = ids[gid];
        const std::uint32_t u = p < q ? p : q, v = p < q ? q : p;
        const std::uint32_t pos = off[u] + atomicAdd(&cursor[u], 1u);
        tv[pos] = v;
        tw[pos] = dist[gid];
        tsrc[pos] = static_cast<std::uint32_t>(gid);
    }
}

__global__ void bridge_scatter_kernel(const std::uint32_t* __restrict__ bu,
                                      const std::uint32_t* __restrict__ bv,
                                      const float* __restrict__ bw, std::uint32_t m,
                                      std::uint32_t n, std::uint32_t src_base,
                                      const std::uint32_t* __restrict__ off,
                                      std::uint32_t* __restrict__ cursor,
                                      std::uint32_t* __restrict__ tv, float* __restrict__ tw,
                                      std::uint32_t* __restrict__ tsrc) {
    const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || bu[i] >= n) return;
    const std::uint32_t u = bu[i];
    const std::uint32_t pos = off[u] + atomicAdd(&cursor[u], 1u);
    tv[pos] = bv[i];
    tw[pos] = bw[i];
    tsrc[pos] = src_base + i;
}

// One wavefront per u: rank each entry of u's segment by (v, src) and write it
// to its final position. Segments of up to 64 entries rank in registers with
// shuffles; longer ones (hubs) compare against the segment in memory.
__global__ void __launch_bounds__(kThreads)
    segment_sort_kernel(std::uint32_t n, const std::uint32_t* __restrict__ off,
                        const std::uint32_t* __restrict__ tv, const float* __restrict__ tw,
                        const std::uint32_t* __restrict__ tsrc, std::uint32_t* __restrict__ eu,
                        std::uint32_t* __restrict__ ev, float* __restrict__ ew) {
    const std::uint32_t waves_per_block = blockDim.x / kWave;
    const std::uint32_t lane = threadIdx.x & (kWave - 1);
    for (std::uint32_t u = blockIdx.x * waves_per_block + threadIdx.x / kWave; u < n;
         u += gridDim.x * waves_per_block) {
        const std::uint32_t s = off[u], L = off[u + 1] - s;
        if (L == 0) continue;
        if (L <= kWave) {
            const bool has = lane < L;
            const std::uint32_t myv = has ? tv[s + lane] : 0xFFFFFFFFu;
            const std::uint32_t mys = has ? tsrc[s + lane] : 0xFFFFFFFFu;
            std::uint32_t rank = 0;
            for (std::uint32_t j = 0; j < L; ++j) {
                const std::uint32_t ov = __shfl(myv, j, kWave);
                const std::uint32_t os = __shfl(mys, j, kWave);
                rank += (ov < myv || (ov == myv && os < mys)) ? 1u : 0u;
            }
            if (has) {
                eu[s + rank] = u;
                ev[s + rank] = myv;
                ew[s + rank] = tw[s + lane];
            }
        } else {
            for (std::uint32_t i = lane; i < L; i += kWave) {
                const std::uint32_t myv = tv[s + i], mys = tsrc[s + i];
                std::uint32_t rank = 0;
                for (std::uint32_t j = 0; j < L; ++j) {
                    const std::uint32_t ov = tv[s + j], os = tsrc[s + j];
                    rank += (ov < myv || (ov == myv && os < mys)) ? 1u : 0u;
                }
                eu[s + rank] = u;
                ev[s + rank] = myv;
                ew[s + rank] = tw[s + i];
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

__device__ __forceinline__ std::uint64_t mrd_key(float w, float cu, float cv, std::uint32_t e) {
    // mrd = max(w, max(core[u], core[v])) exactly as std::max composes it.
    const float wr = fmax_ref(w, fmax_ref(cu, cv));
    return (std::uint64_t(fbits(wr)) << 32) | e;
}

// ---- symmetric adjacency (built once per CORE-SG) --------------------------

// Set on an adjacency entry seen from the edge's second endpoint, so the MRD
// key composes max(core[u], core[v]) in the edge's own operand order.
constexpr std::uint32_t kSecondEnd = 0x80000000u;

__device__ __forceinline__ std::uint64_t adj_key(std::uint32_t raw, float w, float cx,
                                                 const float* __restrict__ core,
                                                 std::uint32_t rank) {
    const std::uint32_t q = raw & ~kSecondEnd;
    const float cq = core[q];
    return (raw & kSecondEnd) ? mrd_key(w, cq, cx, rank) : mrd_key(w, cx, cq, rank);
}

__global__ void __launch_bounds__(kThreads)
    adj_degree_kernel(const std::uint32_t* __restrict__ eu, const std::uint32_t* __restrict__ ev,
                      std::uint64_t m, std::uint32_t* __restrict__ deg) {
    for (std::uint64_t e = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
         e += std::uint64_t(gridDim.x) * blockDim.x) {
        atomicAdd(&deg[eu[e]], 1u);
        atomicAdd(&deg[ev[e]], 1u);
    }
}

__global__ void __launch_bounds__(kThreads)
    adj_fill_kernel(const std::uint32_t* __restrict__ eu, const std::uint32_t* __restrict__ ev,
                    const float* __restrict__ ew, std::uint64_t m,
                    const std::uint32_t* __restrict__ off, std::uint32_t* __restrict__ cursor,
                    std::uint32_t* __restrict__ nbr, std::uint32_t* __restrict__ rank,
                    float* __restrict__ w) {
    for (std::uint64_t e = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
         e += std::uint64_t(gridDim.x) * blockDim.x) {
        const std::uint32_t u = eu[e], v = ev[e];
        const float we = ew[e];
        const std::uint32_t pu = off[u] + atomicAdd(&cursor[u], 1u);
        nbr[pu] = v;
        rank[pu] = static_cast<std::uint32_t>(e);
        w[pu] = we;
        const std::uint32_t pv = off[v] + atomicAdd(&cursor[v], 1u);
        nbr[pv] = u | kSecondEnd;
        rank[pv] = static_cast<std::uint32_t>(e);
        w[pv] = we;
    }
}

// ---- Boruvka rounds over vertices ------------------------------------------
//
// kLanesPerVertex lanes of a wavefront scan one vertex's adjacency and reduce
// the packed minimum with shuffles.
constexpr std::uint32_t kLanesPerVertex = 16;

__device__ __forceinline__ std::uint64_t group_min(std::uint64_t v) {
#pragma unroll
    for (std::uint32_t off = kLanesPerVertex / 2; off > 0; off >>= 1) {
        const std::uint64_t o = __shfl_xor(v, off, kWave);
        v = o < v ? o : v;
    }
    return v;
}

// Round 1: every vertex is its own component; its best edge is the minimum
// MRD key over its adjacency (the reweight happens here), no atomics.
// active[v] records whether v has any edge at all.
__global__ void __launch_bounds__(kThreads)
    vertex_round1_kernel(std::uint32_t n, const std::uint32_t* __restrict__ off,
                         const std::uint32_t* __restrict__ nbr,
                         const std::uint32_t* __restrict__ rank, const float* __restrict__ w,
                         const float* __restrict__ core, std::uint64_t* __restrict__ best,
                         std::uint8_t* __restrict__ active) {
    const std::uint32_t lane = __lane_id(), sub = lane % kLanesPerVertex;
    const std::uint32_t per_wave = kWave / kLanesPerVertex;
    const std::uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const std::uint32_t waves = (gridDim.x * blockDim.x) / kWave;
    for (std::uint32_t vb = wave * per_wave; vb < n; vb += waves * per_wave) {
        const std::uint32_t v = vb + lane / kLanesPerVertex;
        std::uint64_t b = kNone64;
        if (v < n) {
            const float cv = core[v];
            const std::uint32_t e1 = off[v + 1];
            for (std::uint32_t j = off[v] + sub; j < e1; j += kLanesPerVertex) {
                const std::uint64_t key = adj_key(nbr[j], w[j], cv, core, rank[j]);
                b = key < b ? key : b;
            }
        }
        b = group_min(b);
        if (sub == 0 && v < n) {
            best[v] = b;
            active[v] = b != kNone64;
        }
    }
}

// Later rounds: an active vertex's best edge leaving its component, folded
// into the component's best with a 64-bit atomic-min (skipped when the stored
// best is already smaller, which removes most of the contention on large
// components). A vertex with no edge leaving its component stays inside
// forever (components only merge), so it is deactivated.
__global__ void __launch_bounds__(kThreads)
    vertex_round_kernel(std::uint32_t n, const std::uint32_t* __restrict__ off,
                        const std::uint32_t* __restrict__ nbr,
                        const std::uint32_t* __restrict__ rank, const float* __restrict__ w,
                        const float* __restrict__ core, const std::uint32_t* __restrict__ comp,
                        std::uint64_t* __restrict__ best, std::uint8_t* __restrict__ active) {
    const std::uint32_t lane = __lane_id(), sub = lane % kLanesPerVertex;
    const std::uint32_t per_wave = kWave / kLanesPerVertex;
    const std::uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const std::uint32_t waves = (gridDim.x * blockDim.x) / kWave;
    for (std::uint32_t vb = wave * per_wave; vb < n; vb += waves * per_wave) {
        const std::uint32_t v = vb + lane / kLanesPerVertex;
        const bool live = v < n && active[v];
        std::uint32_t cv = 0;
        std::uint64_t b = kNone64;
        if (live) {
            cv = comp[v];
            const float corev = core[v];
            const std::uint32_t e1 = off[v + 1];
            for (std::uint32_t j = off[v] + sub; j < e1; j += kLanesPerVertex) {
                const std::uint32_t raw = nbr[j];
                if (comp[raw & ~kSecondEnd] == cv) continue;
                const std::uint64_t key = adj_key(raw, w[j], corev, core, rank[j]);
                b = key < b ? key : b;
            }
        }
        b = group_min(b);
        if (sub == 0 && live) {
            if (b == kNone64) {
                active[v] = 0;
            } else if (b < __builtin_nontemporal_load(&best[cv])) {
                atomicMin(reinterpret_cast<unsigned long long*>(&best[cv]), b);
            }
        }
    }
}

// Every component root with a best edge hooks onto the component across that
// edge, except the smaller root of a mutual pair. Keys are unique, so hooks
// form a forest; every hook adds its edge to the tree exactly once.
// counters[0] = MST edges so far; counters[kErrorSlot] collects kErr* bits
// (checked by the host after every round; none is set on a well-formed input).
constexpr std::uint32_t kErrorSlot = 12;
constexpr std::uint32_t kErrEdgeIndex = 1u, kErrHookCycle = 2u;

__global__ void __launch_bounds__(kThreads)
    hook_kernel(std::uint32_t n, std::uint64_t m, const std::uint32_t* __restrict__ eu,
                const std::uint32_t* __restrict__ ev, const std::uint32_t* __restrict__ comp,
                const std::uint64_t* __restrict__ best, std::uint32_t* __restrict__ parent,
                std::uint64_t* __restrict__ mst_keys, std::uint8_t* __restrict__ mst_flag,
                std::uint32_t* __restrict__ counters) {
    for (std::uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += gridDim.x * blockDim.x) {
        if (comp[c] != c) continue;
        const std::uint64_t b = best[c];
        if (b == kNone64) continue;
        const std::uint32_t e = static_cast<std::uint32_t>(b);
        if (e >= m) {
            atomicOr(&counters[kErrorSlot], kErrEdgeIndex);
            continue;
        }
        const std::uint32_t a = comp[eu[e]], bb = comp[ev[e]];
        const std::uint32_t other = (a == c) ? bb : a;
        if (best[other] == b && c < other) continue;
        parent[c] = other;
        const std::uint32_t slot = atomicAdd(&counters[0], 1u);
        if (slot < n) mst_keys[slot] = b;
        if (mst_flag) mst_flag[e] = 1;
    }
}

// Final root of every hooked root (chains end at a root that did not hook).
__global__ void __launch_bounds__(kThreads)
    jump_kernel(std::uint32_t n, const std::uint32_t* __restrict__ comp,
                const std::uint32_t* __restrict__ parent, std::uint32_t* __restrict__ newroot,
                std::uint32_t* __restrict__ counters) {
    for (std::uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += gridDim.x * blockDim.x) {
        if (comp[c] != c) continue;
        // Hooks form a forest (unique keys admit no cycle longer than a
        // mutual pair, which the hook rule breaks); the bound only guards
        // against a corrupted parent array.
        std::uint32_t r = c, it = 0;
        for (; parent[r] != r && it < n; ++it) r = parent[r];
        if (it == n) atomicOr(&counters[kErrorSlot], kErrHookCycle);
        newroot[c] = r;
    }
}

__global__ void __launch_bounds__(kThreads)
    relabel_kernel(std::uint32_t n, std::uint32_t* __restrict__ comp,
                   const std::uint32_t* __restrict__ newroot, std::uint32_t* __restrict__ parent,
                   std::uint64_t* __restrict__ best) {
    for (std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += gridDim.x * blockDim.x) {
        comp[v] = newroot[comp[v]];
        parent[v] = v;
        best[v] = kNone64;
    }
}

// ---- exact total weight ----------------------------------------------------

// Decompose a non-negative finite float as M * 2^e with M odd.
__device__ __forceinline__ void odd_decompose(std::uint32_t bits, std::uint32_t* M, int* e) {
    const std::uint32_t E = (bits >> 23) & 0xFFu;
    std::uint32_t m = bits & 0x7FFFFFu;
    int ex;
    if (E == 0) {
        ex = -149;
    } else {
        m |= 0x800000u;
        ex = int(E) - 150;
    }
    const int tz = __builtin_ctz(m);
    *M = m >> tz;
    *e = ex + tz;
}

// stats64[0] = min exponent + 1024 (as u64 for atomicMin), stats[1] = inf flag
__global__ void __launch_bounds__(kThreads)
    weight_exp_kernel(const std::uint64_t* __restrict__ keys,
                      const std::uint32_t* __restrict__ counters, std::uint32_t n,
                      std::uint32_t* __restrict__ stats) {
    const std::uint32_t m = min(counters[0], n);  // a forest has at most n-1 edges
    std::uint32_t local = 0xFFFFFFFFu;
    bool inf = false;
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += gridDim.x * blockDim.x) {
        const std::uint32_t bits = static_cast<std::uint32_t>(keys[i] >> 32);
        if ((bits & 0x7FFFFFFFu) == 0) continue;
        if ((bits & 0x7F800000u) == 0x7F800000u) {
            inf = true;
            continue;
        }
        std::uint32_t M;
        int e;
        odd_decompose(bits, &M, &e);
        const std::uint32_t enc = static_cast<std::uint32_t>(e + 1024);
        local = enc < local ? enc : local;
    }
    if (local != 0xFFFFFFFFu) atomicMin(&stats[0], local);
    if (inf) atomicOr(&stats[1], 1u);
}

// Sum of M << (e - e_min) in u64; stats[2] = overflow flag.
__global__ void __launch_bounds__(kThreads)
    weight_sum_kernel(const std::uint64_t* __restrict__ keys,
                      const std::uint32_t* __restrict__ counters, std::uint32_t n,
                      std::uint32_t* __restrict__ stats, unsigned long long* __restrict__ sum) {
    const std::uint32_t m = min(counters[0], n);
    if (stats[0] == 0xFFFFFFFFu) return;  // no positive finite weight: sum is 0 or inf
    const int emin = int(stats[0]) - 1024;
    unsigned long long local = 0;
    bool over = false;
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += gridDim.x * blockDim.x) {
        const std::uint32_t bits = static_cast<std::uint32_t>(keys[i] >> 32);
        if ((bits & 0x7FFFFFFFu) == 0 || (bits & 0x7F800000u) == 0x7F800000u) continue;
        std::uint32_t M;
        int e;
        odd_decompose(bits, &M, &e);
        const int sh = e - emin;
        const int len = 32 - __builtin_clz(M);
        if (sh + len > 53) {
            over = true;
            continue;
        }
        const unsigned long long t = static_cast<unsigned long long>(M) << sh;
        const unsigned long long s = local + t;
        if (s < local || s >= (1ull << 53)) over = true;
        local = s;
    }
    // wave reduce (values < 2^53 each, so a 64-lane sum cannot wrap u64)
    for (int off = 32; off > 0; off >>= 1) {
        local += __shfl_xor(local, off, kWave);
        const int other = __shfl_xor(static_cast<int>(over), off, kWave);
        over = over || other != 0;
    }
    if (__lane_id() == 0) {
        if (local) {
            const unsigned long long old = atomicAdd(sum, local);
            if (old + local < old || old + local >= (1ull << 53)) over = true;
        }
        if (over) atomicOr(&stats[2], 1u);
    }
}

// ---- outputs ---------------------------------------------------------------

__global__ void __launch_bounds__(kThreads)
    minv_kernel(std::uint32_t n, const std::uint32_t* __restrict__ comp,
                std::uint32_t* __restrict__ minv) {
    for (std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += gridDim.x * blockDim.x)
        atomicMin(&minv[comp[v]], v);
}

__global__ void __launch_bounds__(kThreads)
    first_kernel(std::uint32_t n, const std::uint32_t* __restrict__ comp,
                 const std::uint32_t* __restrict__ minv, std::uint32_t* __restrict__ first) {
    for (std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += gridDim.x * blockDim.x)
        first[v] = (minv[comp[v]] == v) ? 1u : 0u;
}

// Dense labels by first appearance: the label of v is the number of
// components whose smallest member precedes v's component's smallest member.
__global__ void __launch_bounds__(kThreads)
    label_kernel(std::uint32_t n, const std::uint32_t* __restrict__ comp,
                 const std::uint32_t* __restrict__ minv,
                 const std::uint32_t* __restrict__ first_scan, std::uint32_t* __restrict__ labels) {
    for (std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += gridDim.x * blockDim.x)
        labels[v] = first_scan[minv[comp[v]]];
}

__global__ void __launch_bounds__(kThreads)
    flag_to_u32_kernel(const std::uint8_t* __restrict__ flag, std::uint64_t m,
                       std::uint32_t* __restrict__ out) {
    for (std::uint64_t i = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m;
         i += std::uint64_t(gridDim.x) * blockDim.x)
        out[i] = flag[i];
}

// Ordered compaction of the flagged edges with their reweighted weight.
__global__ void __launch_bounds__(kThreads)
    gather_mst_kernel(const std::uint8_t* __restrict__ flag, const std::uint32_t* __restrict__ pos,
                      std::uint64_t m, const std::uint32_t* __restrict__ eu,
                      const std::uint32_t* __restrict__ ev, const float* __restrict__ ew,
                      const float* __restrict__ core, std::uint32_t* __restrict__ ou,
                      std::uint32_t* __restrict__ ov, float* __restrict__ ow) {
    for (std::uint64_t e = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
         e += std::uint64_t(gridDim.x) * blockDim.x) {
        if (!flag[e]) continue;
        const std::uint32_t p = pos[e];
        const std::uint32_t u = eu[e], v = ev[e];
        ou[p] = u;
        ov[p] = v;
        ow[p] = fmax_ref(ew[e], fmax_ref(core[u], core[v]));
    }
}

__global__ void __launch_bounds__(kThreads)
    reweight_kernel(const std::uint32_t* __restrict__ eu, const std::uint32_t* __restrict__ ev,
                    const float* __restrict__ ew, const float* __restrict__ core, std::uint64_t m,
                    float* __restrict__ out) {
    for (std::uint64_t e = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
         e += std::uint64_t(gridDim.x) * blockDim.x)
        out[e] = fmax_ref(ew[e], fmax_ref(core[eu[e]], core[ev[e]]));
}

void check_mpts(const Graph& g, std::uint32_t* mpts) {
    if (*mpts == 0) *mpts = 1;  // coresg.cpp:392
    if (*mpts > g.k) throw std::runtime_error("CoreSG::edges: mpts exceeds k_max");
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

    DBuf<std::uint8_t> emit;
    emit.reserve(slots);
    g.eoff.reserve(std::size_t(n) + 1);
    fill_u32(g.eoff.p, std::size_t(n) + 1, 0u, s);
    emit_count_kernel<<<capped_blocks(slots, kThreads), kThreads, 0, s>>>(
        g.ids.p, g.dist.p, g.count.p, n, k, emit.p, g.eoff.p, nullptr);
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
        bridge_count_kernel<<<blocks_for(nb, kThreads), kThreads, 0, s>>>(dbu.p, nb, n,
                                                                           g.eoff.p);
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

    DBuf<std::uint32_t> cursor, tv, tsrc;
    DBuf<float> tw;
    cursor.reserve(n);
    fill_u32(cursor.p, n, 0u, s);
    tv.reserve(m);
    tw.reserve(m);
    tsrc.reserve(m);
    scatter_kernel<<<capped_blocks(slots, kThreads), kThreads, 0, s>>>(
        g.ids.p, g.dist.p, n, k, emit.p, g.eoff.p, cursor.p, tv.p, tw.p, tsrc.p);
    CSG_LAUNCH_CHECK();
    if (nb) {
        bridge_scatter_kernel<<<blocks_for(nb, kThreads), kThreads, 0, s>>>(
            dbu.p, dbv.p, dbw.p, nb, n, static_cast<std::uint32_t>(slots), g.eoff.p, cursor.p,
            tv.p, tw.p, tsrc.p);
        CSG_LAUNCH_CHECK();
    }
    g.eu.reserve(m);
    g.ev.reserve(m);
    g.ew.reserve(m);
    segment_sort_kernel<<<capped_blocks(std::size_t(n) * kWave, kThreads, 16384), kThreads, 0,
                          s>>>(n, g.eoff.p, tv.p, tw.p, tsrc.p, g.eu.p, g.ev.p, g.ew.p);
    CSG_LAUNCH_CHECK();

    // Symmetric adjacency for the vertex-centric Boruvka rounds.
    g.adj_off.reserve(std::size_t(n) + 1);
    fill_u32(g.adj_off.p, std::size_t(n) + 1, 0u, s);
    if (m) {
        adj_degree_kernel<<<capped_blocks(m, kThreads), kThreads, 0, s>>>(g.eu.p, g.ev.p, m,
                                                                          g.adj_off.p);
        CSG_LAUNCH_CHECK();
    }
    exclusive_scan_u32(g.adj_off.p, std::size_t(n) + 1, nullptr, g.scan_tmp.p, s);
    g.adj_nbr.reserve(2 * m);
    g.adj_rank.reserve(2 * m);
    g.adj_w.reserve(2 * m);
    if (m) {
        fill_u32(cursor.p, n, 0u, s);
        adj_fill_kernel<<<capped_blocks(m, kThreads), kThreads, 0, s>>>(
            g.eu.p, g.ev.p, g.ew.p, m, g.adj_off.p, cursor.p, g.adj_nbr.p, g.adj_rank.p,
            g.adj_w.p);
        CSG_LAUNCH_CHECK();
    }
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
    const std::size_t slots = std::size_t(g.n) * g.k;
    g.counters64.reserve(4);
    std::uint64_t* d_total = g.counters64.p + 3;
    fill_u64(d_total, 1, 0ull, g.stream);
    emit_count_kernel<<<capped_blocks(slots, kThreads), kThreads, 0, g.stream>>>(
        g.ids.p, g.dist.p, g.count.p, g.n, g.k, nullptr, nullptr,
        reinterpret_cast<unsigned long long*>(d_total));
    CSG_LAUNCH_CHECK();
    std::uint64_t total = 0;
    CSG_CHECK(hipMemcpyAsync(&total, d_total, 8, hipMemcpyDeviceToHost, g.stream));
    CSG_CHECK(hipStreamSynchronize(g.stream));
    return total;
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

void extract_msts(Graph& g, const std::uint32_t* mpts_list, std::uint32_t count, void* info_v,
                  std::uint32_t* out_u, std::uint32_t* out_v, float* out_w,
                  std::uint32_t* out_labels) {
    auto* info = static_cast<coresg_hip_mst_info*>(info_v);
    prepare_edges(g);
    const std::uint32_t n = g.n;
    const std::uint64_t m = g.num_edges;
    hipStream_t s = g.stream;
    g.core.reserve(n);
    g.comp.reserve(n);
    g.parent.reserve(n);
    g.newroot.reserve(n);
    g.best.reserve(n);
    g.mst_keys.reserve(n ? n : 1);
    g.counters.reserve(16);
    g.counters64.reserve(4);
    g.active.reserve(n ? n : 1);
    const bool want_edges = out_u || out_v || out_w;
    if (want_edges) g.mst_flag.reserve(m ? m : 1);
    if (out_labels) {
        g.minv.reserve(n);
        g.labels_tmp.reserve(std::size_t(n) + 1);
        g.ensure_scan_tmp(std::size_t(n) + 1);
    }
    if (want_edges) {
        g.out_pos.reserve(m + 1);
        g.ensure_scan_tmp(m + 1);
    }
    std::vector<std::uint32_t> hc(8);

    for (std::uint32_t qi = 0; qi < count; ++qi) {
        std::uint32_t mpts = mpts_list[qi];
        check_mpts(g, &mpts);
        info[qi].mpts = mpts;
        const unsigned nblk = capped_blocks(n, kThreads);
        const unsigned vblk =
            capped_blocks(std::size_t(n) * kLanesPerVertex, kThreads, 16384);
        fill_u32(g.counters.p, 16, 0u, s);
        fill_u64(g.counters64.p, 4, 0ull, s);
        if (want_edges) fill_u8(g.mst_flag.p, m, 0, s);
        core_init_kernel<<<nblk, kThreads, 0, s>>>(g.dist.p, g.count.p, n, g.k, mpts, g.core.p,
                                                   g.comp.p, g.parent.p, g.best.p);
        CSG_LAUNCH_CHECK();

        std::uint32_t rounds = 0, prev_edges = 0;
        while (true) {
            ++rounds;
            if (rounds == 1) {
                if (g.profiling) g.add_work("boruvka_round1", {2 * m, n});
                ScopedKernelTimer t(g, "boruvka_round1");
                vertex_round1_kernel<<<vblk, kThreads, 0, s>>>(n, g.adj_off.p, g.adj_nbr.p,
                                                               g.adj_rank.p, g.adj_w.p, g.core.p,
                                                               g.best.p, g.active.p);
                CSG_LAUNCH_CHECK();
            } else {
                ScopedKernelTimer t(g, rounds == 2 ? "boruvka_round2" : "boruvka_round_live");
                vertex_round_kernel<<<vblk, kThreads, 0, s>>>(n, g.adj_off.p, g.adj_nbr.p,
                                                              g.adj_rank.p, g.adj_w.p, g.core.p,
                                                              g.comp.p, g.best.p, g.active.p);
                CSG_LAUNCH_CHECK();
            }
            hook_kernel<<<nblk, kThreads, 0, s>>>(n, m, g.eu.p, g.ev.p, g.comp.p, g.best.p,
                                                  g.parent.p, g.mst_keys.p,
                                                  want_edges ? g.mst_flag.p : nullptr,
                                                  g.counters.p);
            CSG_LAUNCH_CHECK();
            jump_kernel<<<nblk, kThreads, 0, s>>>(n, g.comp.p, g.parent.p, g.newroot.p,
                                                  g.counters.p);
            CSG_LAUNCH_CHECK();
            relabel_kernel<<<nblk, kThreads, 0, s>>>(n, g.comp.p, g.newroot.p, g.parent.p,
                                                     g.best.p);
            CSG_LAUNCH_CHECK();
            std::uint32_t hc[kErrorSlot + 1];
            CSG_CHECK(hipMemcpyAsync(hc, g.counters.p, sizeof(hc), hipMemcpyDeviceToHost, s));
            CSG_CHECK(hipStreamSynchronize(s));
            const std::uint32_t now = hc[0];
            if (hc[kErrorSlot] || now >= n)
                throw std::runtime_error("extract: Boruvka state corrupted (error bits " +
                                         std::to_string(hc[kErrorSlot]) + ", " +
                                         std::to_string(now) + " hooks for " +
                                         std::to_string(n) + " points)");
            if (now == prev_edges || now + 1 >= n) break;
            prev_edges = now;
        }

        // exact weight
        {
            std::uint32_t* stats = g.counters.p + 8;  // [8] emin, [9] inf, [10] overflow
            fill_u32(stats, 1, 0xFFFFFFFFu, s);
            const unsigned wb = capped_blocks(n, kThreads, 1024);
            weight_exp_kernel<<<wb, kThreads, 0, s>>>(g.mst_keys.p, g.counters.p, n, stats);
            CSG_LAUNCH_CHECK();
            weight_sum_kernel<<<wb, kThreads, 0, s>>>(
                g.mst_keys.p, g.counters.p, n, stats,
                reinterpret_cast<unsigned long long*>(g.counters64.p));
            CSG_LAUNCH_CHECK();
        }
        std::uint32_t hstat[11];
        unsigned long long hsum = 0;
        CSG_CHECK(hipMemcpyAsync(hstat, g.counters.p, sizeof(hstat), hipMemcpyDeviceToHost, s));
        CSG_CHECK(hipMemcpyAsync(&hsum, g.counters64.p, 8, hipMemcpyDeviceToHost, s));
        CSG_CHECK(hipStreamSynchronize(s));
        const std::uint32_t nmst = hstat[0];
        double weight;
        if (hstat[9]) {
            weight = HUGE_VAL;
        } else if (hstat[8] == 0xFFFFFFFFu) {
            weight = 0.0;
        } else if (!hstat[10]) {
            weight = std::ldexp(static_cast<double>(hsum), int(hstat[8]) - 1024);
        } else {
            // Kruskal order is ascending (weight bits, rank): sort and sum.
            std::vector<std::uint64_t> keys(nmst);
            CSG_CHECK(hipMemcpyAsync(keys.data(), g.mst_keys.p, std::size_t(nmst) * 8,
                                     hipMemcpyDeviceToHost, s));
            CSG_CHECK(hipStreamSynchronize(s));
            std::sort(keys.begin(), keys.end());
            weight = 0.0;
            for (std::uint64_t key : keys) {
                const std::uint32_t bits = static_cast<std::uint32_t>(key >> 32);
                float w;
                std::memcpy(&w, &bits, 4);
                weight += static_cast<double>(w);
            }
        }
        info[qi].num_edges = nmst;
        info[qi].num_components = n - nmst;
        info[qi].rounds = rounds;
        info[qi].weight = weight;

        if (out_labels) {
            fill_u32(g.minv.p, n, 0xFFFFFFFFu, s);
            minv_kernel<<<nblk, kThreads, 0, s>>>(n, g.comp.p, g.minv.p);
            CSG_LAUNCH_CHECK();
            first_kernel<<<nblk, kThreads, 0, s>>>(n, g.comp.p, g.minv.p, g.labels_tmp.p);
            CSG_LAUNCH_CHECK();
            exclusive_scan_u32(g.labels_tmp.p, n, nullptr, g.scan_tmp.p, s);
            DBuf<std::uint32_t> lab;
            lab.reserve(n);
            label_kernel<<<nblk, kThreads, 0, s>>>(n, g.comp.p, g.minv.p, g.labels_tmp.p, lab.p);
            CSG_LAUNCH_CHECK();
            CSG_CHECK(hipMemcpyAsync(out_labels + std::size_t(qi) * n, lab.p, std::size_t(n) * 4,
                                     hipMemcpyDeviceToHost, s));
            CSG_CHECK(hipStreamSynchronize(s));
        }
        if (want_edges && n > 0) {
            flag_to_u32_kernel<<<capped_blocks(m, kThreads), kThreads, 0, s>>>(g.mst_flag.p, m,
                                                                               g.out_pos.p);
            CSG_LAUNCH_CHECK();
            exclusive_scan_u32(g.out_pos.p, m, nullptr, g.scan_tmp.p, s);
            DBuf<std::uint32_t> ou, ov;
            DBuf<float> ow;
            ou.reserve(n);
            ov.reserve(n);
            ow.reserve(n);
            gather_mst_kernel<<<capped_blocks(m, kThreads), kThreads, 0, s>>>(
                g.mst_flag.p, g.out_pos.p, m, g.eu.p, g.ev.p, g.ew.p, g.core.p, ou.p, ov.p, ow.p);
            CSG_LAUNCH_CHECK();
            const std::size_t stride = n - 1;
            if (out_u)
                CSG_CHECK(hipMemcpyAsync(out_u + qi * stride, ou.p, std::size_t(nmst) * 4,
                                         hipMemcpyDeviceToHost, s));
            if (out_v)
                CSG_CHECK(hipMemcpyAsync(out_v + qi * stride, ov.p, std::size_t(nmst) * 4,
                                         hipMemcpyDeviceToHost, s));
            if (out_w)
                CSG_CHECK(hipMemcpyAsync(out_w + qi * stride, ow.p, std::size_t(nmst) * 4,
                                         hipMemcpyDeviceToHost, s));
            CSG_CHECK(hipStreamSynchronize(s));
        }
    }
}

}  // namespace coresg_hip
