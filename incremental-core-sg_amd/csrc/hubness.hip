// Hubness, components and anti-hub selection on device.
//
// Replaces the host code of
//   * ConnectivityTracker::rebuild (src/connectivity.cpp:99-139, labels :64-74):
//     lock-free union-find linking the larger root under the smaller, so every
//     root is the smallest id of its component and the dense labels (by first
//     appearance) are the rank of the root among roots;
//   * HubnessIndex::build's reverse degrees (src/hubness.cpp:86-100);
//   * compute_score / apply_scores (src/hubness.cpp:307-352, 404-422): int64
//     neighbour sums and the double division, bit for bit;
//   * select (src/hubness.cpp:424-478): the `budget` smallest nodes under the
//     total order (score, reverse degree, id) by an MSB radix select over the
//     128-bit key, plus the per-component quota of the `per` smallest.
#include <algorithm>
#include <vector>

#include "hubness.hpp"
#include "scan.hpp"

namespace coresg_hip {
namespace {

constexpr unsigned kThreads = 256;

__device__ __forceinline__ std::uint32_t uf_find(std::uint32_t* parent, std::uint32_t x) {
    // Path halving; parents only ever move to ancestors, so a plain store is
    // safe under concurrent unions.
    while (true) {
        const std::uint32_t p = __atomic_load_n(&parent[x], __ATOMIC_RELAXED);
        if (p == x) return x;
        const std::uint32_t gp = __atomic_load_n(&parent[p], __ATOMIC_RELAXED);
        if (gp != p) __atomic_store_n(&parent[x], gp, __ATOMIC_RELAXED);
        x = gp;
    }
}

__device__ __forceinline__ void uf_unite(std::uint32_t* parent, std::uint32_t a,
                                         std::uint32_t b) {
    while (true) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a > b) {
            const std::uint32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&parent[b], b, a) == b) return;
    }
}

// Root of x read through plain (cached) loads. A stale pointer still names an
// ancestor that x had at some point (pointers only move to smaller ancestors),
// so equal results for two nodes prove them connected; different results
// prove nothing and go to uf_unite.
__device__ __forceinline__ std::uint32_t uf_find_cached(const std::uint32_t* parent,
                                                        std::uint32_t x) {
    while (true) {
        const std::uint32_t p = parent[x];
        if (p == x) return x;
        x = p;
    }
}

__global__ void iota_kernel(std::uint32_t* __restrict__ p, std::uint32_t n) {
    for (std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x)
        p[i] = i;
}

// Components, phase 1 (a heuristic in the spirit of Afforest, Sutton et al.
// 2018): a few rounds of Shiloach-Vishkin hooking along each node's first
// kLinkFirst listed neighbours -- the larger of two roots is hooked under the
// smaller by atomicMin -- each followed by compressing every path to its
// root. Hooks follow real edges and pointers only decrease, so the forest
// stays a valid union-find state; a hook lost to a concurrent smaller one is
// simply redone by the full pass below, which visits every slot (the lists
// are directed; no edge is skipped) and is exact. The point is depth: the
// full pass then mostly finds both roots in one or two cached loads.
constexpr std::uint32_t kLinkFirst = 2;
constexpr int kHookRounds = 3;

__global__ void __launch_bounds__(kThreads)
    hook_kernel(const std::uint32_t* __restrict__ ids, const std::uint32_t* __restrict__ count,
                std::uint32_t n, std::uint32_t k, std::uint32_t* parent) {
    for (std::uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += gridDim.x * blockDim.x) {
        const std::uint32_t c = count[p] < kLinkFirst ? count[p] : kLinkFirst;
        for (std::uint32_t j = 0; j < c; ++j) {
            const std::uint32_t q = ids[std::size_t(p) * k + j];
            if (q >= n) continue;
            const std::uint32_t rp = uf_find_cached(parent, p), rq = uf_find_cached(parent, q);
            if (rp != rq) atomicMin(&parent[rp > rq ? rp : rq], rp < rq ? rp : rq);
        }
    }
}

// Every node to its root (no concurrent hooks: plain loads see valid
// ancestors).
__global__ void __launch_bounds__(kThreads)
    compress_kernel(std::uint32_t n, std::uint32_t* parent) {
    for (std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += gridDim.x * blockDim.x)
        parent[v] = uf_find_cached(parent, v);
}

// One pass over every stored slot: reverse-degree count and union. A
// segment of SEG >= k lanes serves one node (lane j its slot j).
template <std::uint32_t SEG>
__global__ void __launch_bounds__(kThreads)
    slot_pass_kernel(const std::uint32_t* __restrict__ ids,
                     const std::uint32_t* __restrict__ count, std::uint32_t n, std::uint32_t k,
                     std::int32_t* __restrict__ rdeg, std::uint32_t* parent) {
    constexpr std::uint32_t P = 64 / SEG;
    const std::uint32_t lane = threadIdx.x & 63u, sub = lane % SEG, seg = lane / SEG;
    const std::uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const std::uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    for (std::uint32_t v0 = wave * P; v0 < n; v0 += waves * P) {
        const std::uint32_t p = v0 + seg;
        if (p >= n || sub >= count[p]) continue;
        const std::uint32_t q = ids[std::size_t(p) * k + sub];
        if (q >= n) continue;
        atomicAdd(&rdeg[q], 1);
        if (q == p) continue;
        if (uf_find_cached(parent, p) == uf_find_cached(parent, q)) continue;
        uf_unite(parent, p, q);
    }
}

__global__ void __launch_bounds__(kThreads)
    root_flag_kernel(std::uint32_t n, std::uint32_t* __restrict__ parent,
                     std::uint32_t* __restrict__ flag) {
    for (std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += gridDim.x * blockDim.x) {
        const std::uint32_t r = uf_find(parent, v);
        flag[v] = (r == v) ? 1u : 0u;
    }
}

__global__ void __launch_bounds__(kThreads)
    label_from_root_kernel(std::uint32_t n, std::uint32_t* __restrict__ parent,
                           const std::uint32_t* __restrict__ root_scan,
                           std::uint32_t* __restrict__ labels) {
    for (std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += gridDim.x * blockDim.x)
        labels[v] = root_scan[uf_find(parent, v)];
}

// Scores: `lanes` lanes per node (power of two <= 64) sum the neighbours'
// reverse degrees, then one lane applies the degenerate rules and divides.
__global__ void __launch_bounds__(kThreads)
    score_kernel(const std::uint32_t* __restrict__ ids, const std::uint32_t* __restrict__ count,
                 std::uint32_t n, std::uint32_t k, std::uint32_t k_for_score,
                 const std::int32_t* __restrict__ rdeg, std::uint32_t lanes,
                 double* __restrict__ score, std::uint32_t* __restrict__ tagged) {
    const std::uint32_t lane = threadIdx.x & 63u;
    const std::uint32_t sub = lane & (lanes - 1u);
    const std::uint32_t per_wave = 64u / lanes;
    const std::uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const std::uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    std::uint32_t local_tagged = 0;
    for (std::uint32_t base = wave * per_wave; base < n; base += waves * per_wave) {
        const std::uint32_t v = base + lane / lanes;
        std::int64_t sum = 0;
        std::uint32_t deg = 0;
        if (v < n) {
            deg = count[v];
            for (std::uint32_t j = sub; j < deg; j += lanes) {
                const std::uint32_t q = ids[std::size_t(v) * k + j];
                if (q < n) sum += rdeg[q];
            }
        }
        for (std::uint32_t off = 1; off < lanes; off <<= 1) sum += __shfl_xor(sum, off, 64);
        if (v < n && sub == 0) {
            const std::int32_t rd = rdeg[v];
            double s;
            if (rd > 0 && sum <= 0) {
                s = double(k_for_score);
            } else if (rd <= 0 && sum > 0) {
                s = 0.0;
            } else if (rd <= 0 && sum <= 0) {
                s = 1.0;
            } else if (deg == 0) {
                s = 1.0;
            } else {
                const double denom = double(sum) / double(deg);
                s = (denom > 0.0) ? double(rd) / denom : double(k_for_score);
            }
            score[v] = s;
            local_tagged += (s < 1.0) ? 1u : 0u;
        }
    }
    local_tagged = wave_sum(local_tagged);
    if (lane == 0 && local_tagged) atomicAdd(tagged, local_tagged);
}

// ---- radix select over the 128-bit key (score bits, rdeg, id) -------------

struct SelectState {
    unsigned long long pref_hi, mask_hi, pref_lo, mask_lo;
    unsigned int rank;  // 0-based rank still to find inside the matching set
    unsigned int pad;
};

__device__ __forceinline__ void node_key(const double* score, const std::int32_t* rdeg,
                                         std::uint32_t v, unsigned long long* hi,
                                         unsigned long long* lo) {
    *hi = static_cast<unsigned long long>(__double_as_longlong(score[v]));
    *lo = (static_cast<unsigned long long>(static_cast<std::uint32_t>(rdeg[v])) << 32) | v;
}

__global__ void __launch_bounds__(kThreads)
    radix_hist_kernel(std::uint32_t n, const double* __restrict__ score,
                      const std::int32_t* __restrict__ rdeg, const SelectState* __restrict__ st,
                      int digit, std::uint32_t* __restrict__ hist) {
    __shared__ std::uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const SelectState s = *st;
    const int shift = (digit < 8) ? (56 - 8 * digit) : (56 - 8 * (digit - 8));
    for (std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += gridDim.x * blockDim.x) {
        unsigned long long hi, lo;
        node_key(score, rdeg, v, &hi, &lo);
        if ((hi & s.mask_hi) != s.pref_hi || (lo & s.mask_lo) != s.pref_lo) continue;
        const unsigned b = static_cast<unsigned>(((digit < 8) ? hi : lo) >> shift) & 0xFFu;
        atomicAdd(&h[b], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

__global__ void __launch_bounds__(256)
    radix_pick_kernel(SelectState* __restrict__ st, int digit, std::uint32_t* __restrict__ hist) {
    __shared__ std::uint32_t c[256];
    c[threadIdx.x] = hist[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        SelectState s = *st;
        std::uint32_t cum = 0, b = 0;
        for (b = 0; b < 256; ++b) {
            if (s.rank < cum + c[b]) break;
            cum += c[b];
        }
        if (b == 256) b = 255;  // unreachable when rank < |matching set|
        s.rank -= cum;
        const int shift = (digit < 8) ? (56 - 8 * digit) : (56 - 8 * (digit - 8));
        if (digit < 8) {
            s.pref_hi |= static_cast<unsigned long long>(b) << shift;
            s.mask_hi |= 0xFFull << shift;
        } else {
            s.pref_lo |= static_cast<unsigned long long>(b) << shift;
            s.mask_lo |= 0xFFull << shift;
        }
        *st = s;
    }
    __syncthreads();
    hist[threadIdx.x] = 0;
}

// Append every node whose key is <= the threshold key (exactly `budget` of them).
__global__ void __launch_bounds__(kThreads)
    collect_le_kernel(std::uint32_t n, const double* __restrict__ score,
                      const std::int32_t* __restrict__ rdeg, const SelectState* __restrict__ st,
                      std::uint32_t cap, std::uint32_t* __restrict__ out,
                      std::uint32_t* __restrict__ out_count) {
    const SelectState s = *st;
    for (std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += gridDim.x * blockDim.x) {
        unsigned long long hi, lo;
        node_key(score, rdeg, v, &hi, &lo);
        if (hi < s.pref_hi || (hi == s.pref_hi && lo <= s.pref_lo)) {
            const std::uint32_t pos = atomicAdd(out_count, 1u);
            if (pos < cap) out[pos] = v;
        }
    }
}

// One 1024-thread block per instance: the `budgets[i]` lowest nodes of
// instance i by (score, rdeg, id) -- the MSB radix select of radix_hist/pick/
// collect_le with the instance-local index as the id, which orders ties
// exactly as the global index does -- written to out[out_off[i] ...].
constexpr unsigned kSegThreads = 1024;
__global__ void __launch_bounds__(kSegThreads)
    segmented_select_kernel(const double* __restrict__ score, const std::int32_t* __restrict__ rdeg,
                            const std::uint32_t* __restrict__ inst_off,
                            const std::uint32_t* __restrict__ budgets,
                            const std::uint32_t* __restrict__ out_off,
                            std::uint32_t* __restrict__ out) {
    __shared__ std::uint32_t h[256];
    __shared__ SelectState st;
    __shared__ std::uint32_t filled;
    const std::uint32_t lo = inst_off[blockIdx.x], cnt = inst_off[blockIdx.x + 1] - lo;
    const std::uint32_t budget = budgets[blockIdx.x];
    std::uint32_t* dst = out + out_off[blockIdx.x];
    const double* sc = score + lo;
    const std::int32_t* rd = rdeg + lo;
    if (budget == 0) return;
    if (budget >= cnt) {
        for (std::uint32_t v = threadIdx.x; v < cnt; v += blockDim.x) dst[v] = lo + v;
        return;
    }
    if (threadIdx.x == 0) {
        st = SelectState{0, 0, 0, 0, budget - 1, 0};
        filled = 0;
    }
    for (int digit = 0; digit < 16; ++digit) {
        if (threadIdx.x < 256) h[threadIdx.x] = 0;
        __syncthreads();
        const SelectState s = st;
        const int shift = (digit < 8) ? (56 - 8 * digit) : (56 - 8 * (digit - 8));
        for (std::uint32_t v = threadIdx.x; v < cnt; v += blockDim.x) {
            unsigned long long khi, klo;
            node_key(sc, rd, v, &khi, &klo);
            if ((khi & s.mask_hi) != s.pref_hi || (klo & s.mask_lo) != s.pref_lo) continue;
            atomicAdd(&h[static_cast<unsigned>(((digit < 8) ? khi : klo) >> shift) & 0xFFu], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            SelectState t = st;
            std::uint32_t cum = 0, b = 0;
            for (b = 0; b < 256; ++b) {
                if (t.rank < cum + h[b]) break;
                cum += h[b];
            }
            if (b == 256) b = 255;  // unreachable when rank < |matching set|
            t.rank -= cum;
            if (digit < 8) {
                t.pref_hi |= static_cast<unsigned long long>(b) << shift;
                t.mask_hi |= 0xFFull << shift;
            } else {
                t.pref_lo |= static_cast<unsigned long long>(b) << shift;
                t.mask_lo |= 0xFFull << shift;
            }
            st = t;
        }
        __syncthreads();
    }
    const SelectState s = st;
    for (std::uint32_t v = threadIdx.x; v < cnt; v += blockDim.x) {
        unsigned long long khi, klo;
        node_key(sc, rd, v, &khi, &klo);
        if (khi < s.pref_hi || (khi == s.pref_hi && klo <= s.pref_lo)) {
            const std::uint32_t pos = atomicAdd(&filled, 1u);
            if (pos < budget) dst[pos] = lo + v;
        }
    }
}

// ---- per-component quota: `per` rounds of a three-stage lexicographic min ---

// Most nodes share a handful of components, so candidates are first reduced
// per wavefront and component (one atomic per distinct component in the
// wave, skipped when the stored minimum is already smaller).
__global__ void __launch_bounds__(kThreads)
    quota_stage_kernel(std::uint32_t n, const double* __restrict__ score,
                       const std::int32_t* __restrict__ rdeg,
                       const std::uint32_t* __restrict__ comp, std::uint32_t ncomp,
                       const std::uint8_t* __restrict__ excluded, int stage,
                       unsigned long long* __restrict__ qs, std::uint32_t* __restrict__ qr,
                       std::uint32_t* __restrict__ qi) {
    const std::uint32_t lane = __lane_id();
    const std::uint32_t stride = gridDim.x * blockDim.x;
    for (std::uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < n;
         base += stride) {
        const std::uint32_t v = base + lane;
        bool act = false;
        std::uint32_t c = 0;
        std::uint64_t val = ~0ull;
        if (v < n && !excluded[v]) {
            c = comp[v];
            if (c < ncomp) {
                const unsigned long long sb =
                    static_cast<unsigned long long>(__double_as_longlong(score[v]));
                if (stage == 0) {
                    act = true;
                    val = sb;
                } else if (stage == 1) {
                    act = sb == qs[c];
                    val = static_cast<std::uint32_t>(rdeg[v]);
                } else {
                    act = sb == qs[c] && static_cast<std::uint32_t>(rdeg[v]) == qr[c];
                    val = v;
                }
            }
        }
        std::uint64_t pending = __ballot(act);
        while (pending) {
            const std::uint32_t leader = static_cast<std::uint32_t>(__builtin_ctzll(pending));
            const std::uint32_t cl = __shfl(c, leader, 64);
            const bool mine = act && c == cl;
            const std::uint64_t m = wave_min_u64(mine ? val : ~0ull);
            if (lane == leader) {
                if (stage == 0) {
                    if (m < qs[cl]) atomicMin(&qs[cl], static_cast<unsigned long long>(m));
                } else if (stage == 1) {
                    if (m < qr[cl]) atomicMin(&qr[cl], static_cast<std::uint32_t>(m));
                } else {
                    if (m < qi[cl]) atomicMin(&qi[cl], static_cast<std::uint32_t>(m));
                }
            }
            act = act && !mine;
            pending = __ballot(act);
        }
    }
}

__global__ void quota_commit_kernel(std::uint32_t ncomp, std::uint32_t round,
                                    unsigned long long* __restrict__ qs,
                                    std::uint32_t* __restrict__ qr,
                                    std::uint32_t* __restrict__ qi,
                                    std::uint8_t* __restrict__ excluded,
                                    std::uint32_t* __restrict__ winners) {
    for (std::uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < ncomp;
         c += gridDim.x * blockDim.x) {
        const std::uint32_t w = qi[c];
        winners[std::size_t(round) * ncomp + c] = w;
        if (w != kNone) excluded[w] = 1;
        qs[c] = ~0ull;
        qr[c] = kNone;
        qi[c] = kNone;
    }
}

template <typename T>
__global__ void gather_kernel(const T* __restrict__ src, const std::uint32_t* __restrict__ idx,
                              std::uint32_t m, T* __restrict__ dst) {
    const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) dst[i] = src[idx[i]];
}

}  // namespace

void hubness_compute(Graph& g, std::uint32_t* tagged_out) {
    const std::uint32_t n = g.n, k = g.k;
    hipStream_t s = g.stream;
    g.rdeg.reserve(n);
    g.score.reserve(n);
    g.conn.reserve(n);
    DBuf<std::uint32_t> parent, flag;
    parent.reserve(n);
    flag.reserve(std::size_t(n) + 1);
    g.counters.reserve(16);
    fill_u32(reinterpret_cast<std::uint32_t*>(g.rdeg.p), n, 0u, s);
    iota_kernel<<<capped_blocks(n, kThreads), kThreads, 0, s>>>(parent.p, n);
    CSG_LAUNCH_CHECK();
    {
        ScopedKernelTimer t(g, "hub_slot_pass");
        for (int r = 0; r < kHookRounds; ++r) {
            hook_kernel<<<capped_blocks(n, kThreads), kThreads, 0, s>>>(g.ids.p, g.count.p, n, k,
                                                                       parent.p);
            CSG_LAUNCH_CHECK();
            compress_kernel<<<capped_blocks(n, kThreads), kThreads, 0, s>>>(n, parent.p);
            CSG_LAUNCH_CHECK();
        }
        const auto pass = k <= 16 ? slot_pass_kernel<16>
                          : k <= 32 ? slot_pass_kernel<32> : slot_pass_kernel<64>;
        pass<<<capped_blocks(std::size_t(n) * 64, kThreads, 16384), kThreads, 0, s>>>(
            g.ids.p, g.count.p, n, k, g.rdeg.p, parent.p);
        CSG_LAUNCH_CHECK();
    }
    root_flag_kernel<<<capped_blocks(n, kThreads), kThreads, 0, s>>>(n, parent.p, flag.p);
    CSG_LAUNCH_CHECK();
    g.ensure_scan_tmp(std::size_t(n) + 1);
    std::uint32_t* d_ncomp = g.counters.p + 12;
    exclusive_scan_u32(flag.p, n, d_ncomp, g.scan_tmp.p, s);
    label_from_root_kernel<<<capped_blocks(n, kThreads), kThreads, 0, s>>>(n, parent.p, flag.p,
                                                                          g.conn.p);
    CSG_LAUNCH_CHECK();

    std::uint32_t lanes = 1;
    while (lanes < k && lanes < 64) lanes <<= 1;
    std::uint32_t* d_tagged = g.counters.p + 13;
    fill_u32(d_tagged, 1, 0u, s);
    {
        ScopedKernelTimer t(g, "hub_score");
        score_kernel<<<capped_blocks(std::size_t(n) * lanes, kThreads), kThreads, 0, s>>>(
            g.ids.p, g.count.p, n, k, g.score_k ? g.score_k : k, g.rdeg.p, lanes, g.score.p,
            d_tagged);
        CSG_LAUNCH_CHECK();
    }
    std::uint32_t h[2] = {0, 0};
    CSG_CHECK(hipMemcpyAsync(h, d_ncomp, 8, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));
    g.num_components = h[0];
    if (tagged_out) *tagged_out = h[1];
    g.hub_ready = true;
}

std::vector<std::uint32_t> select_budget(Graph& g, std::uint32_t lo, std::uint32_t cnt,
                                         std::size_t budget) {
    hipStream_t s = g.stream;
    std::vector<std::uint32_t> sel;
    if (budget > cnt) budget = cnt;
    if (budget == cnt) {
        sel.resize(cnt);
        for (std::uint32_t i = 0; i < cnt; ++i) sel[i] = lo + i;
        return sel;
    }
    if (budget == 0) return sel;
    // the range's arrays from `lo`: keys (score, rdeg, local index) order
    // ties by id exactly as the global index does
    const double* score = g.score.p + lo;
    const std::int32_t* rdeg = g.rdeg.p + lo;
    DBuf<std::uint32_t> picked;
    picked.reserve(budget);
    g.counters.reserve(16);
    std::uint32_t* d_cnt = g.counters.p + 14;
    fill_u32(d_cnt, 1, 0u, s);
    DBuf<SelectState> st;
    DBuf<std::uint32_t> hist;
    st.reserve(1);
    hist.reserve(256);
    SelectState init{0, 0, 0, 0, static_cast<unsigned int>(budget - 1), 0};
    CSG_CHECK(hipMemcpyAsync(st.p, &init, sizeof(init), hipMemcpyHostToDevice, s));
    fill_u32(hist.p, 256, 0u, s);
    const unsigned nb = capped_blocks(cnt, kThreads, 2048);
    for (int d = 0; d < 16; ++d) {
        radix_hist_kernel<<<nb, kThreads, 0, s>>>(cnt, score, rdeg, st.p, d, hist.p);
        CSG_LAUNCH_CHECK();
        radix_pick_kernel<<<1, 256, 0, s>>>(st.p, d, hist.p);
        CSG_LAUNCH_CHECK();
    }
    collect_le_kernel<<<capped_blocks(cnt, kThreads), kThreads, 0, s>>>(
        cnt, score, rdeg, st.p, static_cast<std::uint32_t>(budget), picked.p, d_cnt);
    CSG_LAUNCH_CHECK();
    std::uint32_t got = 0;
    CSG_CHECK(hipMemcpyAsync(&got, d_cnt, 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));
    if (got != budget) throw std::runtime_error("select: radix select miscounted");
    sel.resize(budget);
    CSG_CHECK(hipMemcpyAsync(sel.data(), picked.p, budget * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));
    for (std::uint32_t& v : sel) v += lo;
    return sel;
}

std::vector<std::vector<std::uint32_t>> select_budget_instances(
    Graph& g, const std::vector<std::size_t>& budgets) {
    const std::uint32_t ninst = g.num_instances();
    hipStream_t s = g.stream;
    std::vector<std::uint32_t> b32(ninst), off(ninst + 1, 0);
    for (std::uint32_t i = 0; i < ninst; ++i) {
        const std::uint32_t cnt = g.inst_off[i + 1] - g.inst_off[i];
        b32[i] = static_cast<std::uint32_t>(std::min<std::size_t>(budgets[i], cnt));
        off[i + 1] = off[i] + b32[i];
    }
    std::vector<std::vector<std::uint32_t>> sel(ninst);
    if (off[ninst] == 0) return sel;
    DBuf<std::uint32_t> d_b, d_off, d_out;
    d_b.reserve(ninst);
    d_off.reserve(std::size_t(ninst) + 1);
    d_out.reserve(off[ninst]);
    CSG_CHECK(hipMemcpyAsync(d_b.p, b32.data(), std::size_t(ninst) * 4, hipMemcpyHostToDevice, s));
    CSG_CHECK(hipMemcpyAsync(d_off.p, off.data(), (std::size_t(ninst) + 1) * 4,
                             hipMemcpyHostToDevice, s));
    {
        ScopedKernelTimer t(g, "hub_select_instances");
        segmented_select_kernel<<<ninst, kSegThreads, 0, s>>>(g.score.p, g.rdeg.p, g.inst_off_d.p,
                                                              d_b.p, d_off.p, d_out.p);
        CSG_LAUNCH_CHECK();
    }
    std::vector<std::uint32_t> all(off[ninst]);
    CSG_CHECK(hipMemcpyAsync(all.data(), d_out.p, all.size() * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));
    for (std::uint32_t i = 0; i < ninst; ++i) sel[i].assign(all.begin() + off[i], all.begin() + off[i + 1]);
    return sel;
}

std::vector<std::uint32_t> quota_winners(Graph& g, std::size_t per) {
    const std::uint32_t n = g.n;
    hipStream_t s = g.stream;
    std::vector<std::uint32_t> quota;
    const std::uint32_t ncomp = g.num_components;
    if (per == 0 || ncomp == 0) return quota;
    DBuf<unsigned long long> qs;
    DBuf<std::uint32_t> qr, qi, winners;
    DBuf<std::uint8_t> excl;
    qs.reserve(ncomp);
    qr.reserve(ncomp);
    qi.reserve(ncomp);
    winners.reserve(std::size_t(ncomp) * per);
    excl.reserve(n);
    fill_u64(reinterpret_cast<std::uint64_t*>(qs.p), ncomp, ~0ull, s);
    fill_u32(qr.p, ncomp, kNone, s);
    fill_u32(qi.p, ncomp, kNone, s);
    fill_u8(excl.p, n, 0, s);
    const unsigned nb = capped_blocks(n, kThreads);
    for (std::size_t r = 0; r < per; ++r) {
        for (int stage = 0; stage < 3; ++stage) {
            quota_stage_kernel<<<nb, kThreads, 0, s>>>(n, g.score.p, g.rdeg.p, g.conn.p, ncomp,
                                                      excl.p, stage, qs.p, qr.p, qi.p);
            CSG_LAUNCH_CHECK();
        }
        quota_commit_kernel<<<capped_blocks(ncomp, kThreads), kThreads, 0, s>>>(
            ncomp, static_cast<std::uint32_t>(r), qs.p, qr.p, qi.p, excl.p, winners.p);
        CSG_LAUNCH_CHECK();
    }
    quota.resize(std::size_t(ncomp) * per);
    CSG_CHECK(hipMemcpyAsync(quota.data(), winners.p, quota.size() * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));
    return quota;
}

std::vector<std::uint32_t> merge_selection(Graph& g, std::vector<std::uint32_t> sel,
                                           const std::vector<std::uint32_t>& quota) {
    hipStream_t s = g.stream;
    // Union of the budget set and the quota winners (distinct by
    // construction: one component per node, earlier winners excluded), then
    // the reference's final order (score, rdeg, id).
    std::vector<std::uint32_t> all = sel;
    std::sort(sel.begin(), sel.end());
    for (std::uint32_t w : quota) {
        if (w == kNone) continue;
        if (std::binary_search(sel.begin(), sel.end(), w)) continue;
        all.push_back(w);
    }
    if (all.empty()) return all;
    DBuf<std::uint32_t> d_ids;
    DBuf<double> d_sc;
    DBuf<std::int32_t> d_rd;
    const std::uint32_t m = static_cast<std::uint32_t>(all.size());
    d_ids.reserve(m);
    d_sc.reserve(m);
    d_rd.reserve(m);
    CSG_CHECK(hipMemcpyAsync(d_ids.p, all.data(), std::size_t(m) * 4, hipMemcpyHostToDevice, s));
    gather_kernel<double><<<blocks_for(m, 256), 256, 0, s>>>(g.score.p, d_ids.p, m, d_sc.p);
    gather_kernel<std::int32_t><<<blocks_for(m, 256), 256, 0, s>>>(g.rdeg.p, d_ids.p, m, d_rd.p);
    CSG_LAUNCH_CHECK();
    std::vector<double> sc(m);
    std::vector<std::int32_t> rd(m);
    CSG_CHECK(hipMemcpyAsync(sc.data(), d_sc.p, std::size_t(m) * 8, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipMemcpyAsync(rd.data(), d_rd.p, std::size_t(m) * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));
    std::vector<std::uint32_t> order(m);
    for (std::uint32_t i = 0; i < m; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](std::uint32_t a, std::uint32_t b) {
        if (sc[a] != sc[b]) return sc[a] < sc[b];
        if (rd[a] != rd[b]) return rd[a] < rd[b];
        return all[a] < all[b];
    });
    std::vector<std::uint32_t> out(m);
    for (std::uint32_t i = 0; i < m; ++i) out[i] = all[order[i]];
    return out;
}

std::vector<std::vector<std::uint32_t>> merge_selections(
    Graph& g, std::vector<std::vector<std::uint32_t>> sels,
    const std::vector<std::vector<std::uint32_t>>& quotas) {
    hipStream_t s = g.stream;
    // every instance's union of budget set and quota winners, then one
    // gather of all their scores and reverse degrees
    std::vector<std::vector<std::uint32_t>> alls(sels.size());
    std::vector<std::uint32_t> flat;
    for (std::size_t i = 0; i < sels.size(); ++i) {
        std::vector<std::uint32_t> all = sels[i];
        std::sort(sels[i].begin(), sels[i].end());
        for (std::uint32_t w : quotas[i]) {
            if (w == kNone || std::binary_search(sels[i].begin(), sels[i].end(), w)) continue;
            all.push_back(w);
        }
        flat.insert(flat.end(), all.begin(), all.end());
        alls[i] = std::move(all);
    }
    const std::uint32_t m = static_cast<std::uint32_t>(flat.size());
    if (m == 0) return alls;
    DBuf<std::uint32_t> d_ids;
    DBuf<double> d_sc;
    DBuf<std::int32_t> d_rd;
    d_ids.reserve(m);
    d_sc.reserve(m);
    d_rd.reserve(m);
    CSG_CHECK(hipMemcpyAsync(d_ids.p, flat.data(), std::size_t(m) * 4, hipMemcpyHostToDevice, s));
    gather_kernel<double><<<blocks_for(m, 256), 256, 0, s>>>(g.score.p, d_ids.p, m, d_sc.p);
    gather_kernel<std::int32_t><<<blocks_for(m, 256), 256, 0, s>>>(g.rdeg.p, d_ids.p, m, d_rd.p);
    CSG_LAUNCH_CHECK();
    std::vector<double> sc(m);
    std::vector<std::int32_t> rd(m);
    CSG_CHECK(hipMemcpyAsync(sc.data(), d_sc.p, std::size_t(m) * 8, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipMemcpyAsync(rd.data(), d_rd.p, std::size_t(m) * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));
    std::size_t base = 0;
    for (auto& all : alls) {
        const std::size_t c = all.size();
        std::vector<std::uint32_t> order(c);
        for (std::size_t i = 0; i < c; ++i) order[i] = static_cast<std::uint32_t>(i);
        std::sort(order.begin(), order.end(), [&](std::uint32_t a, std::uint32_t b) {
            if (sc[base + a] != sc[base + b]) return sc[base + a] < sc[base + b];
            if (rd[base + a] != rd[base + b]) return rd[base + a] < rd[base + b];
            return all[a] < all[b];
        });
        std::vector<std::uint32_t> out(c);
        for (std::size_t i = 0; i < c; ++i) out[i] = all[order[i]];
        all = std::move(out);
        base += c;
    }
    return alls;
}

std::vector<std::uint32_t> select_antihubs(Graph& g, std::size_t budget, std::size_t per) {
    return merge_selection(g, select_budget(g, 0, g.n, budget), quota_winners(g, per));
}

}  // namespace coresg_hip
