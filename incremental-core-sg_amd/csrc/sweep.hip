// The mpts sweep: core distances, MRD reweight and MST for a batch of mpts
// values in one pass over the resident CORE-SG.
//
// Replaces, on device, the reference's per-mpts loop (CoreSG::mst / all_msts,
// coresg.cpp:408-417; coresg_cli.cpp:553-570), each iteration of which
// rebuilds the reweighted edge list (coresg.cpp:391-406, knn_graph.cpp:219-235)
// and runs mst_from_edges (mst.cpp:61-156) on it.
//
// Up to kMaxSlots mpts values ("slots") run together. Per-slot state is
// interleaved by vertex -- comp[v*B + b], core[v*B + b] -- so the values a
// Boruvka round needs for one neighbour and every slot sit in one cache line,
// and each adjacency entry read from HBM serves all B slots.
//
// Boruvka with the key (weight bits << 32 | rank): rank is the position in the
// (u, v)-sorted list, so keys are unique and order edges exactly as edge_less
// (types.hpp:38-42); the tree found is the unique MST under that order, the
// reference's Kruskal tree edge for edge. Every component root hooks along
// its minimum outgoing edge (the smaller root of a mutual pair stays), each
// vertex hooks at most once in its lifetime, and the key it hooked with is
// kept per (vertex, slot): the MST is exactly the set of those keys.
//
// The total weight is the reference's double sum in Kruskal order
// (types.hpp:58-62), computed exactly: when every MST weight is M * 2^e with
// M odd and the sum of M * 2^(e - e_min) stays below 2^53, every partial sum of
// the sequential double accumulation is representable, so the sequential
// result equals that integer sum. The rare batch that fails the test is summed
// on the host in Kruskal order.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/coresg_hip.h"
#include "graph.hpp"
#include "mrd_key.hpp"
#include "scan.hpp"

namespace coresg_hip {
namespace {

constexpr unsigned kThreads = 256;
constexpr std::uint32_t kMaxSlots = 32;
constexpr unsigned kReduceBlocks = 1024;  // blocks of the per-slot reducing kernels

constexpr std::uint32_t kErrEdgeIndex = 1u, kErrHookCycle = 2u, kErrHookCount = 4u,
                        kErrTooManyEdges = 8u;
constexpr std::uint32_t kGraphRounds = 4;  // Boruvka rounds per replay of the captured graph
// adjacency entries a round-kernel lane has in flight (each with the
// neighbour's core distance and component behind it); 6 and 8 measured no
// faster at n = 1M (fewer waves per SIMD)
constexpr int kSweepInFlight = 4;
// Largest slot count whose sweeps keep each vertex's best outgoing edge. At
// n = 1M it pays for one slot over many rounds (mpts 1: 5.87 -> 4.29 ms; mpts
// 16, four rounds, unchanged); with 2 to 32 slots a vertex skips its scan
// only when every slot's edge still leaves, and the checks cost more than
// they save (1-6 % slower).
constexpr std::uint32_t kCacheSlots = 1;
constexpr std::uint64_t kOver = 1ull << 63;  // saturated checked sum
constexpr std::uint64_t kExpInf = 1ull << 32;

struct Sweep {
    std::uint32_t n, k, B, Bp;  // points, list length, slots, slots rounded up to a power of 2
    std::uint64_t m;            // CORE-SG edges
    const std::uint32_t* live;  // device word, bit b: slot b still running
    const std::uint32_t* off;   // adjacency
    const AdjEntry* adj;
    const std::uint32_t* eu;  // edge list (by rank)
    const std::uint32_t* ev;
    float* core;  // n*B, interleaved
    std::uint32_t* comp;
    std::uint32_t* parent;
    std::uint32_t* newroot;
    std::uint64_t* best;
    std::uint64_t* mkey;  // key v hooked with as a root, per slot
    std::uint8_t* active;
    std::uint64_t* part;  // kReduceBlocks*B per-block partials
    std::uint32_t* err;
    // per (vertex, slot): the key of the vertex's best outgoing edge at its
    // last scan and that edge's far endpoint (nullptr: no cache)
    std::uint64_t* vkey;
    std::uint32_t* vfar;
};

// Elements i = v*B + b are visited with a grid stride that is a multiple of B,
// so a thread keeps one slot for the whole kernel and can accumulate in
// registers; threads beyond that stride idle.
struct SlotStride {
    std::size_t first, stride;
    bool idle;
};

__device__ __forceinline__ SlotStride slot_stride(std::uint32_t B) {
    const std::size_t total_threads = std::size_t(gridDim.x) * blockDim.x;
    const std::size_t stride = total_threads - total_threads % B;
    const std::size_t t = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    return SlotStride{t, stride, t >= stride};
}

__device__ __forceinline__ std::uint64_t checked_add(std::uint64_t a, std::uint64_t b) {
    if ((a | b) & kOver) return kOver;
    const std::uint64_t s = a + b;
    return (s >= (1ull << 53)) ? kOver : s;
}

struct OpAdd {
    __device__ std::uint64_t operator()(std::uint64_t a, std::uint64_t b) const {
        return checked_add(a, b);
    }
};
// low 32 bits: minimum; bit 32: any-inf flag
struct OpExp {
    __device__ std::uint64_t operator()(std::uint64_t a, std::uint64_t b) const {
        const std::uint64_t la = a & 0xFFFFFFFFull, lb = b & 0xFFFFFFFFull;
        return (la < lb ? la : lb) | ((a | b) & kExpInf);
    }
};

// Per-slot reduction of one value per thread into part[blockIdx.x*B + b].
// Every thread of the block must call it.
template <class Op>
__device__ void slot_block_reduce(std::uint64_t v, std::uint32_t B, std::uint64_t identity,
                                  Op op, std::uint64_t* part) {
    __shared__ std::uint64_t s_val[kThreads];
    s_val[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x < B) {
        const std::size_t base = std::size_t(blockIdx.x) * blockDim.x;
        std::uint64_t acc = identity;
        for (std::uint32_t t = 0; t < blockDim.x; ++t)
            if ((base + t) % B == threadIdx.x) acc = op(acc, s_val[t]);
        part[std::size_t(blockIdx.x) * B + threadIdx.x] = acc;
    }
    __syncthreads();
}

// One block per slot: folds the per-block partials.
template <class Op>
__global__ void __launch_bounds__(kThreads)
    slot_final_kernel(const std::uint64_t* __restrict__ part, std::uint32_t nblocks,
                      std::uint32_t B, std::uint64_t identity, std::uint64_t* __restrict__ out) {
    __shared__ std::uint64_t s_val[kThreads];
    const std::uint32_t b = blockIdx.x;
    const Op op;
    std::uint64_t acc = identity;
    for (std::uint32_t i = threadIdx.x; i < nblocks; i += blockDim.x)
        acc = op(acc, part[std::size_t(i) * B + b]);
    s_val[threadIdx.x] = acc;
    __syncthreads();
    for (std::uint32_t h = blockDim.x / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) s_val[threadIdx.x] = op(s_val[threadIdx.x], s_val[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = s_val[0];
}

// core_mpts(v) = dist[v*k + min(mpts, count) - 1], +inf for an empty list
// (knn_graph.hpp:157-163); every vertex starts as its own component.
__global__ void __launch_bounds__(kThreads)
    sweep_init_kernel(Sweep s, const float* __restrict__ dist,
                      const std::uint32_t* __restrict__ count,
                      const std::uint32_t* __restrict__ mpts) {
    const std::size_t total = std::size_t(s.n) * s.B;
    for (std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += std::size_t(gridDim.x) * blockDim.x) {
        const std::uint32_t v = static_cast<std::uint32_t>(i / s.B);
        const std::uint32_t b = static_cast<std::uint32_t>(i - std::size_t(v) * s.B);
        const std::uint32_t c = count[v];
        float cd = __builtin_inff();
        if (c > 0) {
            const std::uint32_t pos = (mpts[b] < c ? mpts[b] : c) - 1;
            cd = dist[std::size_t(v) * s.k + pos];
        }
        s.core[i] = cd;
        s.comp[i] = v;
        s.parent[i] = v;
        s.best[i] = kNone64;
        s.mkey[i] = kNone64;
        s.active[i] = 1;
    }
}

// Key of adjacency entry `raw` (bit 31: seen from the second endpoint) for a
// slot, with the reweight max(w, max(core[u], core[v])) in the edge's order.
__device__ __forceinline__ std::uint64_t entry_key(std::uint32_t raw, float w, float cx, float cq,
                                                   std::uint32_t rank) {
    return (raw & kSecondEnd) ? mrd_key(w, cq, cx, rank) : mrd_key(w, cx, cq, rank);
}

// The far endpoint, seen from v, of the edge a key names (its rank).
__device__ __forceinline__ std::uint32_t key_far(const Sweep& s, std::uint64_t key,
                                                 std::uint32_t v) {
    const std::uint32_t e = static_cast<std::uint32_t>(key);
    const std::uint32_t a = s.eu[e];
    return a == v ? s.ev[e] : a;
}

// One Boruvka round. A group of W lanes serves one vertex: lane sub handles
// slot sub % Bp and every (W/Bp)-th adjacency entry from sub / Bp, so each
// entry's neighbour, rank and weight are read once for all slots. The first
// round stores each vertex's minimum directly (every vertex is a component);
// later rounds skip neighbours in the same component and fold into the
// component's best with a read-filtered 64-bit atomic-min. A vertex with no
// edge leaving its component stays inside for good (components only merge)
// and is deactivated.
// With the cache (vkey/vfar, few slots), a vertex whose best outgoing edge at
// its last scan still leaves its component skips the scan: its outgoing
// edges are a subset of those it had (components only merge), so that edge
// is still its minimum. Only when it has turned internal is the vertex's
// adjacency read again.
template <std::uint32_t W, bool FIRST>
__global__ void __launch_bounds__(kThreads) sweep_round_kernel(Sweep s) {
    const std::uint32_t lane = __lane_id();
    const std::uint32_t grp = lane / W, sub = lane % W;
    const std::uint32_t b = sub % s.Bp, ph = sub / s.Bp, P = W / s.Bp;
    constexpr std::uint32_t per_wave = kWave / W;
    const std::uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const std::uint32_t waves = (gridDim.x * blockDim.x) / kWave;
    const std::uint64_t gmask = (W == 64) ? ~0ull : (((1ull << W) - 1ull) << (grp * W));
    const std::uint32_t live = *s.live;
    if (live == 0) return;
    for (std::uint32_t vb = wave * per_wave; vb < s.n; vb += waves * per_wave) {
        const std::uint32_t v = vb + grp;
        const bool slot = v < s.n && b < s.B && ((live >> b) & 1u);
        const std::size_t i = std::size_t(v) * s.B + b;
        const bool act = slot && (FIRST || s.active[i] != 0);
        std::uint64_t key_min = kNone64;
        std::uint32_t cv = v;
        bool need = act;  // the vertex's adjacency is read for this slot
        std::uint64_t cached = kNone64;
        if (!FIRST && act) {
            cv = s.comp[i];
            if (s.vkey) {
                cached = s.vkey[i];
                if (cached != kNone64 && s.comp[std::size_t(s.vfar[i]) * s.B + b] != cv)
                    need = false;
            }
        }
        if (__ballot(need) & gmask) {
            float cx = 0.0f;
            if (need) cx = s.core[i];
            const std::uint32_t e1 = s.off[v + 1];
            std::uint32_t j = s.off[v] + ph;
            if (need) {
                // kSweepInFlight entries in flight per lane
                constexpr int F = kSweepInFlight;
                for (; j + (F - 1) * P < e1; j += F * P) {
                    AdjEntry a[F];
#pragma unroll
                    for (int u = 0; u < F; ++u) a[u] = s.adj[j + u * P];
                    std::uint32_t cq_comp[F];
                    float cq[F];
#pragma unroll
                    for (int u = 0; u < F; ++u) {
                        const std::size_t iq = std::size_t(a[u].nbr & ~kSecondEnd) * s.B + b;
                        cq[u] = s.core[iq];
                        cq_comp[u] = FIRST ? ~0u : s.comp[iq];
                    }
#pragma unroll
                    for (int u = 0; u < F; ++u) {
                        if (!FIRST && cq_comp[u] == cv) continue;
                        const std::uint64_t key = entry_key(a[u].nbr, a[u].w, cx, cq[u], a[u].rank);
                        key_min = key < key_min ? key : key_min;
                    }
                }
                for (; j < e1; j += P) {
                    const AdjEntry a = s.adj[j];
                    const std::size_t iq = std::size_t(a.nbr & ~kSecondEnd) * s.B + b;
                    if (!FIRST && s.comp[iq] == cv) continue;
                    const std::uint64_t key = entry_key(a.nbr, a.w, cx, s.core[iq], a.rank);
                    key_min = key < key_min ? key : key_min;
                }
            }
        }
        for (std::uint32_t o = s.Bp; o < W; o <<= 1) {
            const std::uint64_t x = __shfl_xor(key_min, o, kWave);
            key_min = x < key_min ? x : key_min;
        }
        if (ph == 0 && act) {
            if (s.vkey && need) {
                s.vkey[i] = key_min;
                if (key_min != kNone64) s.vfar[i] = key_far(s, key_min, v);
            }
            if (!need) key_min = cached;
            if (FIRST) {
                s.best[i] = key_min;
                s.active[i] = key_min != kNone64;
            } else if (key_min == kNone64) {
                s.active[i] = 0;
            } else {
                std::uint64_t* dst = &s.best[std::size_t(cv) * s.B + b];
                if (key_min < __builtin_nontemporal_load(dst))
                    atomicMin(reinterpret_cast<unsigned long long*>(dst), key_min);
            }
        }
    }
}

// Every root with a best edge hooks onto the component across it, except the
// smaller root of a mutual pair. Keys are unique, so hooks form a forest and
// every hook adds its edge to the tree exactly once. Hooks per slot are
// counted into per-block partials.
__global__ void __launch_bounds__(kThreads) sweep_hook_kernel(Sweep s) {
    const std::uint32_t live = *s.live;
    if (live == 0) return;  // block-uniform; the partials are read for live slots only
    const SlotStride st = slot_stride(s.B);
    std::uint64_t hooks = 0;
    const std::size_t total = std::size_t(s.n) * s.B;
    if (!st.idle) {
        for (std::size_t i = st.first; i < total; i += st.stride) {
            const std::uint32_t v = static_cast<std::uint32_t>(i / s.B);
            const std::uint32_t b = static_cast<std::uint32_t>(i - std::size_t(v) * s.B);
            if (!((live >> b) & 1u) || s.comp[i] != v) continue;
            const std::uint64_t key = s.best[i];
            if (key == kNone64) continue;
            const std::uint32_t e = static_cast<std::uint32_t>(key);
            if (e >= s.m) {
                atomicOr(s.err, kErrEdgeIndex);
                continue;
            }
            const std::uint32_t a = s.comp[std::size_t(s.eu[e]) * s.B + b];
            const std::uint32_t c = s.comp[std::size_t(s.ev[e]) * s.B + b];
            const std::uint32_t other = (a == v) ? c : a;
            if (s.best[std::size_t(other) * s.B + b] == key && v < other) continue;
            s.parent[i] = other;
            s.mkey[i] = key;
            ++hooks;
        }
    }
    slot_block_reduce(st.idle ? 0ull : hooks, s.B, 0ull, OpAdd(), s.part);
}

// Final root of every hooked root (chains end at a root that did not hook).
__global__ void __launch_bounds__(kThreads) sweep_jump_kernel(Sweep s) {
    const std::uint32_t live = *s.live;
    if (live == 0) return;
    const std::size_t total = std::size_t(s.n) * s.B;
    for (std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += std::size_t(gridDim.x) * blockDim.x) {
        const std::uint32_t v = static_cast<std::uint32_t>(i / s.B);
        const std::uint32_t b = static_cast<std::uint32_t>(i - std::size_t(v) * s.B);
        if (!((live >> b) & 1u) || s.comp[i] != v) continue;
        // The bound only guards against a corrupted parent array.
        std::uint32_t r = v, it = 0;
        for (; it < s.n; ++it) {
            const std::uint32_t p = s.parent[std::size_t(r) * s.B + b];
            if (p == r) break;
            r = p;
        }
        if (it == s.n) atomicOr(s.err, kErrHookCycle);
        s.newroot[i] = r;
    }
}

__global__ void __launch_bounds__(kThreads) sweep_relabel_kernel(Sweep s) {
    const std::uint32_t live = *s.live;
    if (live == 0) return;
    const std::size_t total = std::size_t(s.n) * s.B;
    for (std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += std::size_t(gridDim.x) * blockDim.x) {
        const std::uint32_t v = static_cast<std::uint32_t>(i / s.B);
        const std::uint32_t b = static_cast<std::uint32_t>(i - std::size_t(v) * s.B);
        if (!((live >> b) & 1u)) continue;
        s.comp[i] = s.newroot[std::size_t(s.comp[i]) * s.B + b];
        s.parent[i] = v;
        s.best[i] = kNone64;
    }
}

// End of a round, one wavefront: per live slot add its hooks to the total,
// record the round, and retire the slot when it hooked nothing or holds n - 1
// edges (a spanning tree). Keeping this on the device lets several rounds run
// as one captured graph without a host read in between.
__global__ void __launch_bounds__(64)
    sweep_live_kernel(const std::uint64_t* __restrict__ cnt, std::uint32_t B, std::uint32_t n,
                      std::uint32_t* __restrict__ live, std::uint64_t* __restrict__ total,
                      std::uint32_t* __restrict__ rounds, std::uint32_t* __restrict__ round_no,
                      std::uint32_t* __restrict__ err) {
    const std::uint32_t b = threadIdx.x;
    const std::uint32_t mask = *live;
    if (mask == 0) return;  // wave-uniform
    const std::uint32_t r = *round_no + 1;
    bool keep = false;
    if (b < B && ((mask >> b) & 1u)) {
        const std::uint64_t c = cnt[b];
        if (c & kOver) atomicOr(err, kErrHookCount);
        const std::uint64_t t = total[b] + c;
        total[b] = t;
        rounds[b] = r;
        if (t >= n) atomicOr(err, kErrTooManyEdges);
        keep = !(c == 0 || t + 1 >= n);
    }
    const std::uint64_t m = __ballot(keep);
    if (b == 0) {
        *live = static_cast<std::uint32_t>(m);
        *round_no = r;
    }
}

// Decompose a non-negative finite float as M * 2^e with M odd.
__device__ __forceinline__ void odd_decompose(std::uint32_t bits, std::uint32_t* M, int* e) {
    const std::uint32_t E = (bits >> 23) & 0xFFu;
    std::uint32_t mant = bits & 0x7FFFFFu;
    int ex;
    if (E == 0) {
        ex = -149;
    } else {
        mant |= 0x800000u;
        ex = int(E) - 150;
    }
    const int tz = __builtin_ctz(mant);
    *M = mant >> tz;
    *e = ex + tz;
}

// Per slot: smallest exponent of a positive finite MST weight (+1024), and
// whether any weight is infinite.
__global__ void __launch_bounds__(kThreads) sweep_wexp_kernel(Sweep s) {
    const SlotStride st = slot_stride(s.B);
    std::uint64_t acc = 0xFFFFFFFFull;
    const std::size_t total = std::size_t(s.n) * s.B;
    if (!st.idle) {
        for (std::size_t i = st.first; i < total; i += st.stride) {
            const std::uint64_t key = s.mkey[i];
            if (key == kNone64) continue;
            const std::uint32_t bits = static_cast<std::uint32_t>(key >> 32);
            if ((bits & 0x7FFFFFFFu) == 0) continue;
            if ((bits & 0x7F800000u) == 0x7F800000u) {
                acc |= kExpInf;
                continue;
            }
            std::uint32_t M;
            int e;
            odd_decompose(bits, &M, &e);
            const std::uint64_t enc = static_cast<std::uint32_t>(e + 1024);
            acc = OpExp()(acc, enc);
        }
    }
    slot_block_reduce(acc, s.B, 0xFFFFFFFFull, OpExp(), s.part);
}

// Per slot: sum of M << (e - e_min) with saturation at 2^53 (kOver).
__global__ void __launch_bounds__(kThreads)
    sweep_wsum_kernel(Sweep s, const std::uint64_t* __restrict__ emin) {
    const SlotStride st = slot_stride(s.B);
    std::uint64_t acc = 0;
    const std::size_t total = std::size_t(s.n) * s.B;
    if (!st.idle) {
        const std::uint32_t b = static_cast<std::uint32_t>(st.first % s.B);
        const std::uint64_t ex = emin[b];
        const bool none = (ex & 0xFFFFFFFFull) == 0xFFFFFFFFull || (ex & kExpInf);
        const int e_min = int(ex & 0xFFFFFFFFull) - 1024;
        if (!none) {
            for (std::size_t i = st.first; i < total; i += st.stride) {
                const std::uint64_t key = s.mkey[i];
                if (key == kNone64) continue;
                const std::uint32_t bits = static_cast<std::uint32_t>(key >> 32);
                if ((bits & 0x7FFFFFFFu) == 0) continue;
                std::uint32_t M;
                int e;
                odd_decompose(bits, &M, &e);
                const int sh = e - e_min;
                const int len = 32 - __builtin_clz(M);
                acc = (sh + len > 53) ? kOver
                                      : checked_add(acc, static_cast<std::uint64_t>(M) << sh);
            }
        }
    }
    slot_block_reduce(acc, s.B, 0ull, OpAdd(), s.part);
}

// ---- outputs of one slot ----------------------------------------------------

__global__ void __launch_bounds__(kThreads)
    slot_flag_kernel(const std::uint64_t* __restrict__ mkey, std::uint32_t n, std::uint32_t B,
                     std::uint32_t b, std::uint64_t m, std::uint8_t* __restrict__ flag) {
    for (std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += gridDim.x * blockDim.x) {
        const std::uint64_t key = mkey[std::size_t(v) * B + b];
        if (key != kNone64 && static_cast<std::uint32_t>(key) < m)
            flag[static_cast<std::uint32_t>(key)] = 1;
    }
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
                      const float* __restrict__ core, std::uint32_t B, std::uint32_t b,
                      std::uint32_t* __restrict__ ou, std::uint32_t* __restrict__ ov,
                      float* __restrict__ ow) {
    for (std::uint64_t e = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
         e += std::uint64_t(gridDim.x) * blockDim.x) {
        if (!flag[e]) continue;
        const std::uint32_t p = pos[e];
        const std::uint32_t u = eu[e], v = ev[e];
        ou[p] = u;
        ov[p] = v;
        ow[p] = fmax_ref(ew[e], fmax_ref(core[std::size_t(u) * B + b], core[std::size_t(v) * B + b]));
    }
}

// Smallest member of every component of one slot (read-filtered atomic-min:
// members are visited in ascending order, so most find a smaller one stored).
__global__ void __launch_bounds__(kThreads)
    slot_minv_kernel(const std::uint32_t* __restrict__ comp, std::uint32_t n, std::uint32_t B,
                     std::uint32_t b, std::uint32_t* __restrict__ minv) {
    for (std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += gridDim.x * blockDim.x) {
        const std::uint32_t c = comp[std::size_t(v) * B + b];
        if (v < __builtin_nontemporal_load(&minv[c])) atomicMin(&minv[c], v);
    }
}

__global__ void __launch_bounds__(kThreads)
    slot_first_kernel(const std::uint32_t* __restrict__ comp, std::uint32_t n, std::uint32_t B,
                      std::uint32_t b, const std::uint32_t* __restrict__ minv,
                      std::uint32_t* __restrict__ first) {
    for (std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += gridDim.x * blockDim.x)
        first[v] = (minv[comp[std::size_t(v) * B + b]] == v) ? 1u : 0u;
}

// Dense labels by first appearance (mst.cpp:48-59): the label of v is the
// number of components whose smallest member precedes that of v's component.
__global__ void __launch_bounds__(kThreads)
    slot_label_kernel(const std::uint32_t* __restrict__ comp, std::uint32_t n, std::uint32_t B,
                      std::uint32_t b, const std::uint32_t* __restrict__ minv,
                      const std::uint32_t* __restrict__ first_scan,
                      std::uint32_t* __restrict__ labels) {
    for (std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += gridDim.x * blockDim.x)
        labels[v] = first_scan[minv[comp[std::size_t(v) * B + b]]];
}

// ---- per-instance summaries (independent instances sharing the graph) -------

// One block per instance; a thread keeps one slot (threads beyond the largest
// multiple of B idle) and folds the instance's per-(vertex, slot) hook keys.
// Pass 0: MST edge count and (smallest exponent | any-inf) per slot; pass 1:
// the exact integer sum against the instance's smallest exponent.
template <int PASS>
__global__ void __launch_bounds__(kThreads)
    inst_fold_kernel(Sweep s, const std::uint32_t* __restrict__ off,
                     const std::uint64_t* __restrict__ emin, std::uint64_t* __restrict__ out_a,
                     std::uint64_t* __restrict__ out_b) {
    __shared__ std::uint64_t s_a[kThreads], s_b[kThreads];
    const std::uint32_t inst = blockIdx.x, B = s.B;
    const std::uint32_t T = blockDim.x - blockDim.x % B;
    const std::uint32_t lo = off[inst], hi = off[inst + 1];
    const std::uint32_t b = threadIdx.x % B;
    std::uint64_t a = 0, x = PASS == 0 ? 0xFFFFFFFFull : 0ull;
    bool none = false;
    int e_min = 0;
    if (PASS == 1 && threadIdx.x < T) {
        const std::uint64_t ex = emin[std::size_t(inst) * B + b];
        none = (ex & 0xFFFFFFFFull) == 0xFFFFFFFFull || (ex & kExpInf);
        e_min = int(ex & 0xFFFFFFFFull) - 1024;
    }
    if (threadIdx.x < T && !none) {
        for (std::size_t i = std::size_t(lo) * B + threadIdx.x; i < std::size_t(hi) * B; i += T) {
            const std::uint64_t key = s.mkey[i];
            if (key == kNone64) continue;
            const std::uint32_t bits = static_cast<std::uint32_t>(key >> 32);
            if (PASS == 0) {
                ++a;
                if ((bits & 0x7FFFFFFFu) == 0) continue;
                if ((bits & 0x7F800000u) == 0x7F800000u) {
                    x |= kExpInf;
                    continue;
                }
                std::uint32_t M;
                int e;
                odd_decompose(bits, &M, &e);
                x = OpExp()(x, static_cast<std::uint32_t>(e + 1024));
            } else {
                if ((bits & 0x7FFFFFFFu) == 0) continue;
                std::uint32_t M;
                int e;
                odd_decompose(bits, &M, &e);
                const int sh = e - e_min;
                const int len = 32 - __builtin_clz(M);
                x = (sh + len > 53) ? kOver : checked_add(x, static_cast<std::uint64_t>(M) << sh);
            }
        }
    }
    s_a[threadIdx.x] = a;
    s_b[threadIdx.x] = x;
    __syncthreads();
    if (threadIdx.x < B) {
        std::uint64_t ra = 0, rx = PASS == 0 ? 0xFFFFFFFFull : 0ull;
        for (std::uint32_t t = threadIdx.x; t < T; t += B) {
            ra += s_a[t];
            rx = PASS == 0 ? OpExp()(rx, s_b[t]) : OpAdd()(rx, s_b[t]);
        }
        if (PASS == 0) out_a[std::size_t(inst) * B + threadIdx.x] = ra;
        out_b[std::size_t(inst) * B + threadIdx.x] = rx;
    }
}

std::uint32_t pow2_at_least(std::uint32_t x) {
    std::uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Lanes per vertex group (W >= Bp: a group holds every slot of its vertex).
template <std::uint32_t W>
void launch_round_w(const Sweep& sw, bool first, unsigned blocks, hipStream_t s) {
    if (first)
        sweep_round_kernel<W, true><<<blocks, kThreads, 0, s>>>(sw);
    else
        sweep_round_kernel<W, false><<<blocks, kThreads, 0, s>>>(sw);
}

void launch_round(const Sweep& sw, bool first, std::uint32_t W, unsigned blocks, hipStream_t s) {
    switch (W) {
        case 4: launch_round_w<4>(sw, first, blocks, s); break;
        case 8: launch_round_w<8>(sw, first, blocks, s); break;
        case 16: launch_round_w<16>(sw, first, blocks, s); break;
        case 32: launch_round_w<32>(sw, first, blocks, s); break;
        case 64: launch_round_w<64>(sw, first, blocks, s); break;
        default: throw std::runtime_error("sweep: lanes per vertex must be 4..64");
    }
    CSG_LAUNCH_CHECK();
}

// Exact double total of one slot's MST from the device statistics, or the
// host's Kruskal-order sum when the integer sum saturated.
double slot_weight(Graph& g, const Sweep& sw, std::uint32_t b, std::uint64_t ex,
                   std::uint64_t sum, std::uint32_t lo = 0, std::uint32_t hi = kNone) {
    if (hi == kNone) hi = sw.n;
    if (ex & kExpInf) return HUGE_VAL;
    if ((ex & 0xFFFFFFFFull) == 0xFFFFFFFFull) return 0.0;
    if (!(sum & kOver)) return std::ldexp(static_cast<double>(sum), int(ex & 0xFFFFFFFFull) - 1024);
    std::vector<std::uint64_t> keys(hi - lo);
    CSG_CHECK(hipMemcpy2DAsync(keys.data(), 8, sw.mkey + std::size_t(lo) * sw.B + b,
                               std::size_t(sw.B) * 8, 8, hi - lo, hipMemcpyDeviceToHost, g.stream));
    CSG_CHECK(hipStreamSynchronize(g.stream));
    keys.erase(std::remove(keys.begin(), keys.end(), kNone64), keys.end());
    std::sort(keys.begin(), keys.end());  // Kruskal order: (weight bits, rank)
    double total = 0.0;
    for (std::uint64_t key : keys) {
        const std::uint32_t bits = static_cast<std::uint32_t>(key >> 32);
        float w;
        std::memcpy(&w, &bits, 4);
        total += static_cast<double>(w);
    }
    return total;
}

void run_batch(Graph& g, const std::uint32_t* mpts, std::uint32_t B, coresg_hip_mst_info* info,
               std::uint32_t* out_u, std::uint32_t* out_v, float* out_w,
               std::uint32_t* out_labels, coresg_hip_mst_info* inst_info = nullptr,
               std::uint32_t inst_stride = 0) {
    const std::uint32_t n = g.n;
    const std::uint64_t m = g.num_edges;
    hipStream_t s = g.stream;
    const std::size_t nb = std::size_t(n) * B;
    g.core.reserve(nb);
    g.comp.reserve(nb);
    g.parent.reserve(nb);
    g.newroot.reserve(nb);
    g.best.reserve(nb);
    g.mst_keys.reserve(nb);
    g.active.reserve(nb);
    g.slot_part.reserve(std::size_t(kReduceBlocks) * kMaxSlots);
    g.slot_out.reserve(4 * kMaxSlots);  // counts, exponents, sums, totals
    g.slot_mpts.reserve(kMaxSlots);
    g.counters.reserve(16);

    std::uint32_t hm[kMaxSlots];
    for (std::uint32_t b = 0; b < B; ++b) {
        hm[b] = mpts[b];
        check_mpts(g, &hm[b]);
        info[b].mpts = hm[b];
    }
    CSG_CHECK(hipMemcpyAsync(g.slot_mpts.p, hm, B * 4, hipMemcpyHostToDevice, s));
    fill_u32(g.counters.p, 16, 0u, s);

    Sweep sw{};
    sw.n = n;
    sw.k = g.k;
    sw.B = B;
    sw.Bp = pow2_at_least(B);
    sw.m = m;
    // device state of the round loop: live mask, round number, per-slot
    // totals and rounds (slot_out: [0,32) hook counts, [32,64) exponents,
    // [64,96) sums, [96,128) totals; counters: 0 error bits, 1 live, 2 round)
    std::uint32_t* d_live = g.counters.p + 1;
    std::uint32_t* d_round = g.counters.p + 2;
    std::uint64_t* d_total = g.slot_out.p + 3 * kMaxSlots;
    DBuf<std::uint32_t>& d_rounds = g.slot_rounds;
    d_rounds.reserve(kMaxSlots);
    fill_u32(d_live, 1, (B == 32) ? 0xFFFFFFFFu : ((1u << B) - 1u), s);
    fill_u64(d_total, B, 0ull, s);
    fill_u32(d_rounds.p, B, 0u, s);
    sw.live = d_live;
    sw.off = g.adj_off.p;
    sw.adj = g.adj.p;
    sw.eu = g.eu.p;
    sw.ev = g.ev.p;
    sw.core = g.core.p;
    sw.comp = g.comp.p;
    sw.parent = g.parent.p;
    sw.newroot = g.newroot.p;
    sw.best = g.best.p;
    sw.mkey = g.mst_keys.p;
    sw.active = g.active.p;
    sw.part = g.slot_part.p;
    sw.err = g.counters.p;
    // the best-edge cache pays where a vertex's check (one random component
    // read per slot) costs less than its shared adjacency scan: few slots
    if (B <= kCacheSlots) {
        g.sweep_vkey.reserve(nb);
        g.sweep_vfar.reserve(nb);
        sw.vkey = g.sweep_vkey.p;
        sw.vfar = g.sweep_vfar.p;
    }

    const unsigned eblk = capped_blocks(nb, kThreads);
    // launch shape of the round kernel: the autotuned one for this slot
    // count, else 16 lanes per vertex (32 above 16 slots) and at most 16384
    // blocks
    SweepShape shape = sweep_shape(g, sw.Bp);
    const unsigned vlanes = shape.lanes;
    const unsigned vblk = capped_blocks(std::size_t(n) * vlanes, kThreads, shape.max_blocks);
    sweep_init_kernel<<<eblk, kThreads, 0, s>>>(sw, g.dist.p, g.count.p, g.slot_mpts.p);
    CSG_LAUNCH_CHECK();

    std::uint64_t* d_cnt = g.slot_out.p;
    // the rest of a round after its scan: hooks, their per-slot count, pointer
    // jumping, relabelling, liveness
    auto round_tail = [&]() {
        sweep_hook_kernel<<<kReduceBlocks, kThreads, 0, s>>>(sw);
        CSG_LAUNCH_CHECK();
        slot_final_kernel<OpAdd><<<B, kThreads, 0, s>>>(sw.part, kReduceBlocks, B, 0ull, d_cnt);
        CSG_LAUNCH_CHECK();
        sweep_jump_kernel<<<eblk, kThreads, 0, s>>>(sw);
        CSG_LAUNCH_CHECK();
        sweep_relabel_kernel<<<eblk, kThreads, 0, s>>>(sw);
        CSG_LAUNCH_CHECK();
        sweep_live_kernel<<<1, 64, 0, s>>>(d_cnt, B, n, d_live, d_total, d_rounds.p, d_round,
                                           g.counters.p);
        CSG_LAUNCH_CHECK();
    };
    // round 1 (each vertex its own component) directly
    {
        if (g.profiling) g.add_work("boruvka_round1", {2 * m, n, B});
        ScopedKernelTimer t(g, "boruvka_round1");
        launch_round(sw, true, vlanes, vblk, s);
    }
    round_tail();
    if (g.plain) {
        // the plain form: every later round launched on its own and its live
        // mask read back before the next
        for (std::uint32_t round = 2;; ++round) {
            std::uint32_t st[2] = {0, 0};  // error bits, live mask
            CSG_CHECK(hipMemcpyAsync(st, g.counters.p, 8, hipMemcpyDeviceToHost, s));
            CSG_CHECK(hipStreamSynchronize(s));
            if (st[0])
                throw std::runtime_error("extract: Boruvka state corrupted (error bits " +
                                         std::to_string(st[0]) + ")");
            if (st[1] == 0) break;
            if (round > 64) throw std::runtime_error("extract: Boruvka did not converge");
            launch_round(sw, false, vlanes, vblk, s);
            round_tail();
        }
    }
    // later rounds: kGraphRounds of them per replay of a captured graph (the
    // same launches every round; kernels of retired slots return at once), one
    // host read of the live mask per replay
    if (!g.plain) {
        std::vector<unsigned char> key(sizeof(Sweep) + 4 * sizeof(unsigned));
        const unsigned dims[4] = {vblk, eblk, B, kGraphRounds * 256u + vlanes};
        std::memcpy(key.data(), &sw, sizeof(Sweep));
        std::memcpy(key.data() + sizeof(Sweep), dims, sizeof(dims));
        if (!g.sweep_exec || g.sweep_key != key) {
            if (g.sweep_exec) {
                std::lock_guard<std::mutex> lk(device_api_mutex());
                CSG_CHECK(hipGraphExecDestroy(g.sweep_exec));
            }
            g.sweep_exec = nullptr;
            hipGraph_t graph = nullptr;
            {
                CaptureGuard cap(s);
                for (std::uint32_t r = 0; r < kGraphRounds; ++r) {
                    launch_round(sw, false, vlanes, vblk, s);
                    round_tail();
                }
                graph = cap.end();
            }
            hipError_t e;
            {
                std::lock_guard<std::mutex> lk(device_api_mutex());
                e = hipGraphInstantiate(&g.sweep_exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
            }
            CSG_CHECK(e);
            g.sweep_key = key;
        }
    }
    for (std::uint32_t replay = 0; !g.plain; ++replay) {
        std::uint32_t st[2] = {0, 0};  // error bits, live mask
        CSG_CHECK(hipMemcpyAsync(st, g.counters.p, 8, hipMemcpyDeviceToHost, s));
        CSG_CHECK(hipStreamSynchronize(s));
        if (st[0])
            throw std::runtime_error("extract: Boruvka state corrupted (error bits " +
                                     std::to_string(st[0]) + ")");
        if (st[1] == 0) break;
        if (replay * kGraphRounds > 64)
            throw std::runtime_error("extract: Boruvka did not converge");
        {
            ScopedKernelTimer t(g, "boruvka_rounds_graph");
            CSG_CHECK(hipGraphLaunch(g.sweep_exec, s));
        }
    }
    std::uint64_t total[kMaxSlots] = {};
    std::uint32_t rounds[kMaxSlots] = {};
    CSG_CHECK(hipMemcpyAsync(total, d_total, B * 8, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipMemcpyAsync(rounds, d_rounds.p, B * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));

    if (inst_info) {
        // per instance: edges, components, exact weight (rounds: the batch's)
        const std::uint32_t ninst = g.num_instances();
        const std::size_t m = std::size_t(ninst) * B;
        DBuf<std::uint64_t> icnt, iexp, isum;
        icnt.reserve(m);
        iexp.reserve(m);
        isum.reserve(m);
        inst_fold_kernel<0><<<ninst, kThreads, 0, s>>>(sw, g.inst_off_d.p, nullptr, icnt.p, iexp.p);
        CSG_LAUNCH_CHECK();
        inst_fold_kernel<1><<<ninst, kThreads, 0, s>>>(sw, g.inst_off_d.p, iexp.p, nullptr, isum.p);
        CSG_LAUNCH_CHECK();
        std::vector<std::uint64_t> hc(m), he(m), hs(m);
        CSG_CHECK(hipMemcpyAsync(hc.data(), icnt.p, m * 8, hipMemcpyDeviceToHost, s));
        CSG_CHECK(hipMemcpyAsync(he.data(), iexp.p, m * 8, hipMemcpyDeviceToHost, s));
        CSG_CHECK(hipMemcpyAsync(hs.data(), isum.p, m * 8, hipMemcpyDeviceToHost, s));
        CSG_CHECK(hipStreamSynchronize(s));
        for (std::uint32_t i = 0; i < ninst; ++i) {
            const std::uint32_t lo = g.inst_off[i], hi = g.inst_off[i + 1];
            for (std::uint32_t b = 0; b < B; ++b) {
                coresg_hip_mst_info& r = inst_info[std::size_t(i) * inst_stride + b];
                const std::size_t j = std::size_t(i) * B + b;
                r.mpts = hm[b];
                r.num_edges = static_cast<std::uint32_t>(hc[j]);
                r.num_components = (hi - lo) - static_cast<std::uint32_t>(hc[j]);
                r.rounds = rounds[b];
                r.weight = slot_weight(g, sw, b, he[j], hs[j], lo, hi);
            }
        }
        return;
    }

    // exact totals
    std::uint64_t* d_exp = g.slot_out.p + kMaxSlots;
    std::uint64_t* d_sum = g.slot_out.p + 2 * kMaxSlots;
    sweep_wexp_kernel<<<kReduceBlocks, kThreads, 0, s>>>(sw);
    CSG_LAUNCH_CHECK();
    slot_final_kernel<OpExp><<<B, kThreads, 0, s>>>(sw.part, kReduceBlocks, B, 0xFFFFFFFFull,
                                                    d_exp);
    CSG_LAUNCH_CHECK();
    sweep_wsum_kernel<<<kReduceBlocks, kThreads, 0, s>>>(sw, d_exp);
    CSG_LAUNCH_CHECK();
    slot_final_kernel<OpAdd><<<B, kThreads, 0, s>>>(sw.part, kReduceBlocks, B, 0ull, d_sum);
    CSG_LAUNCH_CHECK();
    std::uint64_t hexp[kMaxSlots], hsum[kMaxSlots];
    CSG_CHECK(hipMemcpyAsync(hexp, d_exp, B * 8, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipMemcpyAsync(hsum, d_sum, B * 8, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));
    for (std::uint32_t b = 0; b < B; ++b) {
        info[b].num_edges = static_cast<std::uint32_t>(total[b]);
        info[b].num_components = n - static_cast<std::uint32_t>(total[b]);
        info[b].rounds = rounds[b];
        info[b].weight = slot_weight(g, sw, b, hexp[b], hsum[b]);
    }

    const bool want_edges = out_u || out_v || out_w;
    if (!want_edges && !out_labels) return;
    const unsigned nblk = capped_blocks(n, kThreads);
    if (out_labels) {
        g.minv.reserve(n);
        g.labels_tmp.reserve(std::size_t(n) + 1);
        g.label_buf.reserve(n);
        g.ensure_scan_tmp(std::size_t(n) + 1);
    }
    if (want_edges) {
        g.mst_flag.reserve(m ? m : 1);
        g.out_pos.reserve(m + 1);
        g.out_u.reserve(n);
        g.out_v.reserve(n);
        g.out_w.reserve(n);
        g.ensure_scan_tmp(m + 1);
    }
    const std::size_t estride = n ? n - 1 : 0;
    for (std::uint32_t b = 0; b < B; ++b) {
        if (out_labels) {
            fill_u32(g.minv.p, n, 0xFFFFFFFFu, s);
            slot_minv_kernel<<<nblk, kThreads, 0, s>>>(g.comp.p, n, B, b, g.minv.p);
            CSG_LAUNCH_CHECK();
            slot_first_kernel<<<nblk, kThreads, 0, s>>>(g.comp.p, n, B, b, g.minv.p,
                                                        g.labels_tmp.p);
            CSG_LAUNCH_CHECK();
            exclusive_scan_u32(g.labels_tmp.p, n, nullptr, g.scan_tmp.p, s);
            slot_label_kernel<<<nblk, kThreads, 0, s>>>(g.comp.p, n, B, b, g.minv.p,
                                                        g.labels_tmp.p, g.label_buf.p);
            CSG_LAUNCH_CHECK();
            CSG_CHECK(hipMemcpyAsync(out_labels + std::size_t(b) * n, g.label_buf.p,
                                     std::size_t(n) * 4, hipMemcpyDeviceToHost, s));
        }
        if (want_edges && n > 0 && m > 0) {
            const std::size_t nmst = total[b];
            fill_u8(g.mst_flag.p, m, 0, s);
            slot_flag_kernel<<<nblk, kThreads, 0, s>>>(g.mst_keys.p, n, B, b, m, g.mst_flag.p);
            CSG_LAUNCH_CHECK();
            flag_to_u32_kernel<<<capped_blocks(m, kThreads), kThreads, 0, s>>>(g.mst_flag.p, m,
                                                                               g.out_pos.p);
            CSG_LAUNCH_CHECK();
            exclusive_scan_u32(g.out_pos.p, m, nullptr, g.scan_tmp.p, s);
            gather_mst_kernel<<<capped_blocks(m, kThreads), kThreads, 0, s>>>(
                g.mst_flag.p, g.out_pos.p, m, g.eu.p, g.ev.p, g.ew.p, g.core.p, B, b, g.out_u.p,
                g.out_v.p, g.out_w.p);
            CSG_LAUNCH_CHECK();
            if (out_u)
                CSG_CHECK(hipMemcpyAsync(out_u + b * estride, g.out_u.p, nmst * 4,
                                         hipMemcpyDeviceToHost, s));
            if (out_v)
                CSG_CHECK(hipMemcpyAsync(out_v + b * estride, g.out_v.p, nmst * 4,
                                         hipMemcpyDeviceToHost, s));
            if (out_w)
                CSG_CHECK(hipMemcpyAsync(out_w + b * estride, g.out_w.p, nmst * 4,
                                         hipMemcpyDeviceToHost, s));
        }
        // the per-slot scratch above is reused by the next slot
        CSG_CHECK(hipStreamSynchronize(s));
    }
}

}  // namespace

// ---- the sweep's autotuner ---------------------------------------------------

namespace {

struct TuneKey {
    std::uint32_t log2n, k, Bp;
    bool operator<(const TuneKey& o) const {
        return std::tie(log2n, k, Bp) < std::tie(o.log2n, o.k, o.Bp);
    }
};

std::mutex g_tune_mu;
std::map<TuneKey, SweepShape> g_tune;  // process-wide, by size bucket

TuneKey tune_key(const Graph& g, std::uint32_t Bp) {
    std::uint32_t l = 0;
    while ((std::uint64_t(1) << (l + 1)) <= g.n) ++l;
    return TuneKey{l, g.k, Bp};
}

}  // namespace

// The graph's own tuned shape for Bp slots, else the table's for its bucket,
// else the table's for the nearest bucket with the same (k, Bp) at most two
// size doublings away (the smaller one on a tie), else the default.
SweepShape sweep_shape(const Graph& g, std::uint32_t Bp) {
    auto it = g.sweep_tune.find(Bp);
    if (it != g.sweep_tune.end()) return it->second;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        const TuneKey key = tune_key(g, Bp);
        for (std::uint32_t dist = 0; dist <= 2; ++dist) {
            for (const std::uint32_t l : {key.log2n - dist, key.log2n + dist}) {
                if (l > 63) continue;  // below bucket 0
                auto jt = g_tune.find(TuneKey{l, key.k, key.Bp});
                if (jt != g_tune.end()) return jt->second;
            }
        }
    }
    return SweepShape{Bp > 16 ? 32u : 16u, 16384u};
}

void autotune_sweep(Graph& g, const std::uint32_t* mpts, std::uint32_t count, void* res_v) {
    auto* res = static_cast<coresg_hip_tune_result*>(res_v);
    if (count == 0) throw std::runtime_error("autotune: no mpts values");
    const std::uint32_t B = std::min<std::uint32_t>(count, kMaxSlots);
    const std::uint32_t Bp = pow2_at_least(B);
    prepare_edges(g);
    std::vector<coresg_hip_mst_info> info(B);
    auto time_one = [&]() {
        double best = 1e300;
        for (int rep = 0; rep < 3; ++rep) {  // the first also warms up
            const auto t0 = std::chrono::steady_clock::now();
            extract_msts(g, mpts, B, info.data(), nullptr, nullptr, nullptr, nullptr);
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                    .count();
            if (rep > 0) best = std::min(best, ms);
        }
        return best;
    };
    g.sweep_tune[Bp] = SweepShape{Bp > 16 ? 32u : 16u, 16384u};
    const double default_ms = time_one();
    SweepShape best = g.sweep_tune[Bp];
    double best_ms = default_ms;
    std::uint32_t tried = 1;
    for (std::uint32_t lanes = 4; lanes <= 64; lanes <<= 1) {
        if (lanes < Bp) continue;
        for (std::uint32_t cap : {4096u, 8192u, 16384u, 32768u}) {
            const SweepShape c{lanes, cap};
            if (c.lanes == (Bp > 16 ? 32u : 16u) && c.max_blocks == 16384u) continue;
            g.sweep_tune[Bp] = c;
            const double ms = time_one();
            ++tried;
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
    }
    g.sweep_tune[Bp] = best;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[tune_key(g, Bp)] = best;
    }
    if (res) {
        res->slots = B;
        res->lanes = best.lanes;
        res->max_blocks = best.max_blocks;
        res->candidates = tried;
        res->best_ms = best_ms;
        res->default_ms = default_ms;
    }
}

std::string tune_table_text() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::string out;
    for (const auto& kv : g_tune) {
        out += std::to_string(kv.first.log2n) + " " + std::to_string(kv.first.k) + " " +
               std::to_string(kv.first.Bp) + " " + std::to_string(kv.second.lanes) + " " +
               std::to_string(kv.second.max_blocks) + "\n";
    }
    return out;
}

void tune_table_load(const char* text) {
    std::map<TuneKey, SweepShape> add;
    const char* p = text;
    while (p && *p) {
        unsigned a, b, c, d, e;
        int used = 0;
        if (std::sscanf(p, "%u %u %u %u %u%n", &a, &b, &c, &d, &e, &used) != 5)
            throw std::runtime_error("tune table: expected 'log2n k slots lanes max_blocks' lines");
        if (d < 4 || d > 64 || (d & (d - 1)) || d < c || c == 0 || c > 32 || (c & (c - 1)) ||
            e == 0)
            throw std::runtime_error("tune table: invalid entry");
        add[TuneKey{a, b, c}] = SweepShape{d, e};
        p += used;
        while (*p == '\n' || *p == ' ' || *p == '\r' || *p == '\t') ++p;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto& kv : add) g_tune[kv.first] = kv.second;
}

void extract_msts(Graph& g, const std::uint32_t* mpts_list, std::uint32_t count, void* info_v,
                  std::uint32_t* out_u, std::uint32_t* out_v, float* out_w,
                  std::uint32_t* out_labels) {
    auto* info = static_cast<coresg_hip_mst_info*>(info_v);
    prepare_edges(g);
    const std::size_t n = g.n, estride = n ? n - 1 : 0;
    // the plain form sweeps one mpts value per pass
    const std::uint32_t per_pass = g.plain ? 1u : kMaxSlots;
    for (std::uint32_t q0 = 0; q0 < count; q0 += per_pass) {
        const std::uint32_t B = std::min<std::uint32_t>(per_pass, count - q0);
        run_batch(g, mpts_list + q0, B, info + q0, out_u ? out_u + q0 * estride : nullptr,
                  out_v ? out_v + q0 * estride : nullptr, out_w ? out_w + q0 * estride : nullptr,
                  out_labels ? out_labels + q0 * n : nullptr);
    }
}

void extract_instances(Graph& g, const std::uint32_t* mpts_list, std::uint32_t count,
                       void* info_v) {
    auto* info = static_cast<coresg_hip_mst_info*>(info_v);
    if (g.inst_off.empty()) throw std::runtime_error("extract_instances: no instances set");
    prepare_edges(g);
    std::vector<coresg_hip_mst_info> union_info(kMaxSlots);
    for (std::uint32_t q0 = 0; q0 < count; q0 += kMaxSlots) {
        const std::uint32_t B = std::min<std::uint32_t>(kMaxSlots, count - q0);
        run_batch(g, mpts_list + q0, B, union_info.data(), nullptr, nullptr, nullptr, nullptr,
                  info + q0, count);
    }
}

}  // namespace coresg_hip
