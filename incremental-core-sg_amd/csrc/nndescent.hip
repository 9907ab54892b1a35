// Dynamic NN-Descent on device, lock free.
//
// Reference behaviour (src/nndescent.cpp):
//   * init_random (:211-240) with the per-node splitmix64 stream NodeRng
//     (:12-29). Every node only writes its own list, so the result does not
//     depend on the schedule; it is reproduced bit for bit here: each lane of a
//     wavefront computes one draw (splitmix64's state advances by a constant,
//     so draw t is independent of the others), the first `needed` distinct
//     non-self candidates are accepted in draw order, and the list is merged
//     stably by distance exactly as repeated sorted insertion would.
//   * run_descent/local_join (:52-209): sample <= smp "new" entries (clearing
//     their flag) and <= 3*smp "old" ones from the sorted list, publish into
//     bounded reverse lists, join new x (new u old) and insert improvements.
//
// The reference's device kernel serialises list updates with per-node spin
// locks, which need independent thread scheduling. Here a round has three
// lock-free phases:
//   1. sample: one wavefront per node, ballots give the new/old ranks;
//      reverse lists are bounded atomic appends.
//   2. join: a gather pass lists each node's distinct candidates (new u
//      reverse-new, then old u reverse-old) with their thresholds; one
//      128-thread group per node stages the candidate rows in LDS, computes
//      every new x (new u old) distance in the reference's exact summation
//      order, and appends every candidate that beats a member's current k-th
//      distance to that member's bounded proposal buffer as a packed
//      (distance bits << 32 | id) key.
//   3. merge: one 64-thread group per node drops duplicates, self and ids
//      already listed, ranks the buffer by counting, and merges the best into
//      the sorted list (new entries after equal distances, flagged new).
// The early stop is the reference's: updates <= delta * k * n.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <vector>

#include "build.hpp"
#include "mfma.hpp"
#include "scan.hpp"

namespace coresg_hip {
namespace {

constexpr std::uint32_t kMaxK = 64;        // list length handled by one wavefront
constexpr std::uint32_t kMaxSample = 32;   // reference cap (cuda_common.cuh:32)
constexpr std::uint32_t kSmallMerge = 16;  // merge survivors ranked by counting
constexpr std::uint32_t kJoinThreads = 128;
// reverse lists by target buckets: kRevTargets targets per bucket, at most
// kRevMaxBuckets buckets (n <= 2^25; larger graphs use one device atomic per
// entry), sampling and scatter on a grid of kRevBlocks blocks. Every block
// writes one run per bucket: 4096-target buckets and 2048 blocks keep these
// runs long (about 100 entries at n = 1M, k = 32), so their partly written
// lines stay few.
constexpr std::uint32_t kRevShift = 12;
constexpr std::uint32_t kRevTargets = 1u << kRevShift;
constexpr std::uint32_t kRevMaxBuckets = 8192;
constexpr unsigned kRevBlocks = 2048;
constexpr unsigned kRevThreads = 1024;
constexpr std::uint32_t kRevBatch = 4;  // pairs in flight per thread of the bucket pass
constexpr std::uint32_t kJoinRegStride = 64;  // widest row a join thread keeps in registers
// matrix-core filter join: threads per node, and the members whose rows a
// first pass stages per node for rows of at least 64 floats (nodes with more
// members -- a few percent -- go to a second pass that stages all S): the
// smaller LDS footprint lets four nodes share a CU at d = 128 instead of two
constexpr std::uint32_t mfma_join_threads(std::uint32_t) { return 256u; }
constexpr std::uint32_t mfma_row_cap(std::uint32_t stride) { return stride >= 64 ? 48u : 0u; }
constexpr std::uint32_t kMfmaList = 1536;     // pairs listed for exact distances per batch
constexpr unsigned kDeferBlocks = 2048;        // grid of the filter join's second pass

// ---------------------------------------------------------------------------
// init_random
// ---------------------------------------------------------------------------

__device__ __forceinline__ std::uint64_t splitmix_mix(std::uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// One wavefront per node. Requires count <= k <= 64 and attempts <= 64 per
// chunk (looped for larger k).
__global__ void __launch_bounds__(256)
    init_random_kernel(const float* __restrict__ pts, std::uint32_t n, std::uint32_t stride,
                       std::uint32_t k, std::uint64_t seed, std::uint32_t* __restrict__ ids,
                       float* __restrict__ dist, std::uint8_t* __restrict__ flags,
                       std::uint32_t* __restrict__ count, const std::uint32_t* __restrict__ inst_of,
                       const std::uint32_t* __restrict__ inst_off, bool cosine) {
    __shared__ std::uint32_t s_stage_id[4][64];
    __shared__ float s_stage_d[4][64];
    const std::uint32_t lane = threadIdx.x & 63u, wib = threadIdx.x >> 6;
    const std::uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const std::uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    for (std::uint32_t node = wave; node < n; node += waves) {
        const std::size_t base = std::size_t(node) * k;
        const std::uint32_t cnt0 = count[node];
        const std::uint32_t needed = k - cnt0;
        const std::uint32_t attempts = needed + needed / 2 + 4;
        // With instances the node draws inside its own instance, with its
        // instance-local index: exactly the init of the instance built alone.
        std::uint32_t lo = 0, span = n;
        if (inst_of) {
            const std::uint32_t i = inst_of[node];
            lo = inst_off[i];
            span = inst_off[i + 1] - lo;
        }
        // NodeRng(seed, node, 0): state s0, one discarded draw; draw t uses
        // state s0 + (t + 2) * phi.
        const std::uint64_t s0 = seed ^ (0x9E3779B97F4A7C15ull * (std::uint64_t(node - lo) + 1)) ^
                                 (0xC2B2AE3D27D4EB4Full * 1ull);
        // existing entries
        const std::uint32_t my_id = lane < cnt0 ? ids[base + lane] : kNone;
        const float my_d = lane < cnt0 ? dist[base + lane] : __builtin_inff();
        const std::uint8_t my_f = lane < cnt0 ? flags[base + lane] : 0;
        std::uint32_t accepted = 0;
        // accepted candidates so far, held by lanes [0, accepted) of acc_*
        std::uint32_t acc_id = kNone;
        float acc_d = __builtin_inff();
        for (std::uint32_t t0 = 0; t0 < attempts && accepted < needed; t0 += 64) {
            const std::uint32_t t = t0 + lane;
            const bool live = t < attempts;
            std::uint32_t cand = kNone;
            if (live) {
                const std::uint64_t z =
                    splitmix_mix(s0 + (std::uint64_t(t) + 2ull) * 0x9E3779B97F4A7C15ull);
                cand = static_cast<std::uint32_t>(z >> 32) % span + lo;
            }
            bool ok = live && cand != node;
            // not already in the list (existing or accepted in earlier chunks)
            // (every lane executes every shuffle: a shuffle skipped by
            // short-circuit evaluation would read from inactive lanes)
            for (std::uint32_t j = 0; j < cnt0; ++j) {
                const std::uint32_t oid = __shfl(my_id, j, 64);
                ok = ok && (oid != cand);
            }
            for (std::uint32_t j = 0; j < accepted; ++j) {
                const std::uint32_t oid = __shfl(acc_id, j, 64);
                ok = ok && (oid != cand);
            }
            // first occurrence in this chunk
#pragma unroll
            for (std::uint32_t j = 0; j < 64; ++j) {
                const std::uint32_t oc = static_cast<std::uint32_t>(
                    __builtin_amdgcn_readlane(static_cast<int>(cand), static_cast<int>(j)));
                ok = ok && !(j < lane && oc == cand);
            }
            const std::uint64_t mask = __ballot(ok);
            const std::uint32_t rank = accepted + __popcll(mask & ((1ull << lane) - 1ull));
            const bool take = ok && rank < needed;
            float d = __builtin_inff();
            if (take)
                d = metric_finalize(l2sqr_exact(pts + std::size_t(node) * stride,
                                                pts + std::size_t(cand) * stride, stride),
                                    cosine);
            // move taken candidates into lanes [accepted, accepted + taken):
            // each writes its slot (its rank) in the wave's staging area
            const std::uint64_t tmask = __ballot(take);
            const std::uint32_t taken = __popcll(tmask);
            if (take) {
                s_stage_id[wib][rank] = cand;
                s_stage_d[wib][rank] = d;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane >= accepted && lane < accepted + taken) {
                acc_id = s_stage_id[wib][lane];
                acc_d = s_stage_d[wib][lane];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            accepted += taken;
        }
        // stable merge: existing entries (sorted) first among equals, then the
        // accepted ones in draw order.
        std::uint32_t pos_e = lane;  // for existing lane < cnt0
        std::uint32_t pos_a = 0;     // for accepted lane < accepted
        for (std::uint32_t j = 0; j < accepted; ++j) {
            const float od = __shfl(acc_d, j, 64);
            pos_e += (od < my_d) ? 1u : 0u;
            // accepted before this one in the order (distance, draw index)
            pos_a += (od < acc_d || (j < lane && od == acc_d)) ? 1u : 0u;
        }
        for (std::uint32_t j = 0; j < cnt0; ++j) {
            const float od = __shfl(my_d, j, 64);
            pos_a += (od <= acc_d) ? 1u : 0u;
        }
        if (lane < cnt0) {
            ids[base + pos_e] = my_id;
            dist[base + pos_e] = my_d;
            flags[base + pos_e] = my_f;
        }
        if (lane < accepted) {
            ids[base + pos_a] = acc_id;
            dist[base + pos_a] = acc_d;
            flags[base + pos_a] = 1;
        }
        if (lane == 0) count[node] = cnt0 + accepted;
    }
}

// ---------------------------------------------------------------------------
// sampling
// ---------------------------------------------------------------------------

// One segment of SEG lanes per node (SEG >= k, a power of two), 64 / SEG
// nodes per wavefront: the nodes' dependent loads and atomics are in flight
// together. Lane `sub` of a segment holds list entry `sub`; ballots are masked
// to the segment. With BUCKETS the reverse lists are left to the bucket
// passes below and the block only counts its sampled entries per target
// bucket (hist, bucket-major: hist[b * gridDim.x + block]); otherwise every
// entry takes its reverse-list slot with a device atomic. Each node's join
// threshold goes to thr_node (n floats, small enough to stay in cache for the
// gather's random reads).
template <std::uint32_t SEG, bool BUCKETS>
__global__ void __launch_bounds__(256)
    sample_kernel(std::uint32_t n, std::uint32_t k, const std::uint32_t* __restrict__ ids,
                  std::uint8_t* __restrict__ flags, const std::uint32_t* __restrict__ count,
                  std::uint32_t smp, std::uint32_t* __restrict__ new_list,
                  std::uint32_t* __restrict__ new_cnt, std::uint32_t* __restrict__ old_list,
                  std::uint32_t* __restrict__ old_cnt, std::uint32_t* __restrict__ rev_new,
                  std::uint32_t* __restrict__ rev_new_cnt, std::uint32_t* __restrict__ rev_old,
                  std::uint32_t* __restrict__ rev_old_cnt, std::uint32_t* __restrict__ hist,
                  std::uint32_t nb, const float* __restrict__ dist,
                  float* __restrict__ thr_node, const std::uint32_t* __restrict__ halt,
                  bool cosine) {
    extern __shared__ std::uint32_t s_hist[];  // nb buckets (BUCKETS only)
    if (*halt) return;
    constexpr std::uint32_t P = 64 / SEG;  // nodes per wavefront
    const std::uint32_t lane = threadIdx.x & 63u;
    const std::uint32_t sub = lane % SEG, seg = lane / SEG;
    const std::uint64_t segmask = (SEG == 64 ? ~0ull : ((1ull << SEG) - 1ull)) << (seg * SEG);
    const std::uint64_t lt = ((1ull << lane) - 1ull) & segmask;
    const std::uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const std::uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    const std::uint32_t max_new = smp, max_old = 3u * smp;
    if (BUCKETS) {
        for (std::uint32_t x = threadIdx.x; x < nb; x += blockDim.x) s_hist[x] = 0;
        __syncthreads();
    }
    // two of the wavefront's nodes (v and v + waves P, the nodes the scatter
    // pass walks in the same block) per iteration: both nodes' counts and
    // list slots are loaded before either is sampled (a node's k slots are
    // read whatever its count, so these loads do not wait for it)
    for (std::uint32_t v00 = wave * P; v00 < n; v00 += 2 * waves * P) {
      std::uint32_t c2[2], id2[2], fl2[2];
      float last2[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const std::uint32_t v = v00 + h * waves * P + seg;
        const bool live = v < n;
        const std::size_t base = std::size_t(v) * k;
        c2[h] = live ? count[v] : 0u;
        id2[h] = live && sub < k ? ids[base + sub] : kNone;
        fl2[h] = live && sub < k ? flags[base + sub] : 0u;
        last2[h] = live && sub == 0 ? dist[base + k - 1] : 0.0f;
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const std::uint32_t v = v00 + h * waves * P + seg;
        const bool live = v < n;
        const std::size_t base = std::size_t(v) * k;
        const std::uint32_t c = c2[h];
        const bool has = sub < c;
        const std::uint32_t id = has ? id2[h] : kNone;
        const bool isnew = has && fl2[h] != 0;
        const bool isold = has && !isnew;
        const std::uint64_t mnew = __ballot(isnew), mold = __ballot(isold);
        const std::uint32_t rn = __popcll(mnew & lt), ro = __popcll(mold & lt);
        const bool take_new = isnew && rn < max_new && id < n;
        const bool take_old = isold && ro < max_old && id < n;
        if (isnew && rn < max_new) {
            flags[base + sub] = 0;
            new_list[std::size_t(v) * max_new + rn] = id < n ? id : kNone;
        }
        if (isold && ro < max_old) old_list[std::size_t(v) * max_old + ro] = id < n ? id : kNone;
        if (BUCKETS) {
            if (take_new || take_old) atomicAdd(&s_hist[id >> kRevShift], 1u);
        } else {
            // bounded reverse lists: a full list takes no more (the reference
            // checks the count before its fetch_add, nndescent.cpp ReverseCache)
            if (take_new && __atomic_load_n(&rev_new_cnt[id], __ATOMIC_RELAXED) < smp) {
                const std::uint32_t pos = atomicAdd(&rev_new_cnt[id], 1u);
                if (pos < smp) rev_new[std::size_t(id) * smp + pos] = v;
            }
            if (take_old && __atomic_load_n(&rev_old_cnt[id], __ATOMIC_RELAXED) < smp) {
                const std::uint32_t pos = atomicAdd(&rev_old_cnt[id], 1u);
                if (pos < smp) rev_old[std::size_t(id) * smp + pos] = v;
            }
        }
        if (sub == 0 && live) {
            const std::uint32_t tn = __popcll(mnew & segmask), to = __popcll(mold & segmask);
            new_cnt[v] = tn < max_new ? tn : max_new;
            old_cnt[v] = to < max_old ? to : max_old;
            // the node's join threshold D[k-1]^2 for this round (to_rank of
            // its current k-th distance, nndescent.cpp:41-42)
            thr_node[v] = metric_to_rank(last2[h], cosine);
        }
      }
    }
    if (BUCKETS) {
        __syncthreads();
        for (std::uint32_t x = threadIdx.x; x < nb; x += blockDim.x)
            hist[std::size_t(x) * gridDim.x + blockIdx.x] = s_hist[x];
    }
}

// Reverse lists by target buckets (kRevTargets consecutive targets each), in
// place of one device atomic per sampled entry. After sample_kernel<.., true>
// and an exclusive scan of its histograms, the scatter pass walks the same
// nodes in the same blocks and writes each sampled entry (target, source,
// old bit) into its bucket's range at the block's offset plus an LDS cursor;
// then one block per bucket hands out the reverse-list slots with LDS
// counters and writes every target's counts. Which sources a full reverse
// list keeps is arrival order, as with the reference's fetch_add.
template <std::uint32_t SEG>
__global__ void __launch_bounds__(256)
    rev_scatter_kernel(std::uint32_t n, std::uint32_t smp,
                       const std::uint32_t* __restrict__ new_list,
                       const std::uint32_t* __restrict__ new_cnt,
                       const std::uint32_t* __restrict__ old_list,
                       const std::uint32_t* __restrict__ old_cnt,
                       const std::uint32_t* __restrict__ off, std::uint32_t nb,
                       unsigned long long* __restrict__ pairs,
                       const std::uint32_t* __restrict__ halt) {
    extern __shared__ std::uint32_t s_cur[];  // nb bucket cursors
    if (*halt) return;
    constexpr std::uint32_t P = 64 / SEG;
    const std::uint32_t lane = threadIdx.x & 63u;
    const std::uint32_t sub = lane % SEG, seg = lane / SEG;
    const std::uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const std::uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    for (std::uint32_t x = threadIdx.x; x < nb; x += blockDim.x)
        s_cur[x] = off[std::size_t(x) * gridDim.x + blockIdx.x];
    __syncthreads();
    // two of the wavefront's nodes (the sampling pass's v and v + waves P) per
    // iteration, their counts and then all their entries loaded before any is
    // placed: two memory waits per two nodes instead of per node
    constexpr std::uint32_t RN = (kMaxSample + SEG - 1) / SEG;
    constexpr std::uint32_t RO = (3 * kMaxSample + SEG - 1) / SEG;
    for (std::uint32_t v0 = wave * P; v0 < n; v0 += 2 * waves * P) {
        const std::uint32_t vv[2] = {v0 + seg, v0 + waves * P + seg};
        std::uint32_t nc[2], oc[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            nc[h] = vv[h] < n ? new_cnt[vv[h]] : 0u;
            oc[h] = vv[h] < n ? old_cnt[vv[h]] : 0u;
        }
        std::uint32_t ne[2][RN], oe[2][RO];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (std::uint32_t r = 0; r < RN; ++r) {
                const std::uint32_t e = sub + r * SEG;
                ne[h][r] = e < nc[h] ? new_list[std::size_t(vv[h]) * smp + e] : kNone;
            }
#pragma unroll
            for (std::uint32_t r = 0; r < RO; ++r) {
                const std::uint32_t e = sub + r * SEG;
                oe[h][r] = e < oc[h] ? old_list[std::size_t(vv[h]) * 3 * smp + e] : kNone;
            }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (std::uint32_t r = 0; r < RN; ++r) {
                const std::uint32_t id = ne[h][r];
                if (id < n) {
                    const std::uint32_t pos = atomicAdd(&s_cur[id >> kRevShift], 1u);
                    pairs[pos] = (static_cast<unsigned long long>(id) << 32) | vv[h];
                }
            }
#pragma unroll
            for (std::uint32_t r = 0; r < RO; ++r) {
                const std::uint32_t id = oe[h][r];
                if (id < n) {
                    const std::uint32_t pos = atomicAdd(&s_cur[id >> kRevShift], 1u);
                    pairs[pos] = (static_cast<unsigned long long>(id) << 32) | vv[h] | 0x80000000u;
                }
            }
        }
    }
}

__global__ void __launch_bounds__(kRevThreads)
    rev_bucket_kernel(std::uint32_t n, std::uint32_t smp,
                      const unsigned long long* __restrict__ pairs,
                      const std::uint32_t* __restrict__ off, std::uint32_t nblk,
                      const std::uint32_t* __restrict__ total, std::uint32_t* __restrict__ rev_new,
                      std::uint32_t* __restrict__ rev_new_cnt, std::uint32_t* __restrict__ rev_old,
                      std::uint32_t* __restrict__ rev_old_cnt,
                      const std::uint32_t* __restrict__ halt) {
    __shared__ std::uint32_t s_cn[kRevTargets], s_co[kRevTargets];
    if (*halt) return;
    const std::uint32_t b = blockIdx.x, lo = b << kRevShift;
    for (std::uint32_t t = threadIdx.x; t < kRevTargets; t += kRevThreads) {
        s_cn[t] = 0;
        s_co[t] = 0;
    }
    __syncthreads();
    const std::uint32_t start = off[std::size_t(b) * nblk];
    const std::uint32_t end = b + 1 < gridDim.x ? off[std::size_t(b + 1) * nblk] : *total;
    // kRevBatch pairs loaded per thread before any is placed: one memory wait
    // per batch instead of one per pair
    for (std::uint32_t e0 = start + threadIdx.x; e0 < end; e0 += kRevBatch * kRevThreads) {
        unsigned long long pb[kRevBatch];
#pragma unroll
        for (std::uint32_t u = 0; u < kRevBatch; ++u) {
            const std::uint32_t e = e0 + u * kRevThreads;
            pb[u] = e < end ? pairs[e] : ~0ull;
        }
#pragma unroll
        for (std::uint32_t u = 0; u < kRevBatch; ++u) {
            if (e0 + u * kRevThreads >= end) break;
            const unsigned long long pr = pb[u];
            const std::uint32_t id = static_cast<std::uint32_t>(pr >> 32);
            const std::uint32_t src = static_cast<std::uint32_t>(pr) & 0x7FFFFFFFu;
            const bool is_old = (static_cast<std::uint32_t>(pr) & 0x80000000u) != 0;
            const std::uint32_t pos = atomicAdd(is_old ? &s_co[id - lo] : &s_cn[id - lo], 1u);
            if (pos < smp) (is_old ? rev_old : rev_new)[std::size_t(id) * smp + pos] = src;
        }
    }
    __syncthreads();
    for (std::uint32_t t = threadIdx.x; t < kRevTargets && lo + t < n; t += kRevThreads) {
        rev_new_cnt[lo + t] = s_cn[t];
        rev_old_cnt[lo + t] = s_co[t];
    }
}

// ---------------------------------------------------------------------------
// local join
// ---------------------------------------------------------------------------

// Gather: each node builds its deduplicated candidate list -- A = sampled new
// u reverse new (first), then old u reverse old -- keeping the first copy of
// every id, plus each member's threshold D[k-1]^2 (to_rank of its current
// k-th distance, nndescent.cpp:41-42, from the sampling pass's thr_node). cnt = na | ns << 16. One segment of
// SEG lanes per node, 64 / SEG nodes per wavefront (their dependent loads in
// flight together); ballots are masked to the segment. First copies are
// found in a per-node LDS table (id -> smallest position) instead of
// comparing every entry with all earlier ones.
template <std::uint32_t SEG>
__global__ void __launch_bounds__(256)
    gather_kernel(std::uint32_t n, std::uint32_t smp, const std::uint32_t* __restrict__ new_list,
                  const std::uint32_t* __restrict__ new_cnt,
                  const std::uint32_t* __restrict__ old_list,
                  const std::uint32_t* __restrict__ old_cnt,
                  const std::uint32_t* __restrict__ rev_new,
                  const std::uint32_t* __restrict__ rev_new_cnt,
                  const std::uint32_t* __restrict__ rev_old,
                  const std::uint32_t* __restrict__ rev_old_cnt,
                  std::uint32_t* __restrict__ out_ids, float* __restrict__ out_thr,
                  std::uint32_t* __restrict__ out_cnt, const float* __restrict__ thr_node,
                  const std::uint32_t* __restrict__ halt) {
    if (*halt) return;
    constexpr std::uint32_t P = 64 / SEG;  // nodes per wavefront
    constexpr std::uint32_t R = (6 * kMaxSample + SEG - 1) / SEG;  // entries per lane
    // per node: open-addressing table id -> smallest position (load <= 3/4)
    constexpr std::uint32_t H = 256;
    static_assert(4 * 6 * kMaxSample <= 3 * H, "table load");
    __shared__ std::uint32_t s_hkey[4][P][H], s_hidx[4][P][H];
    const std::uint32_t lane = threadIdx.x & 63u, wib = threadIdx.x >> 6;
    const std::uint32_t sub = lane % SEG, seg = lane / SEG;
    const std::uint64_t segmask = (SEG == 64 ? ~0ull : ((1ull << SEG) - 1ull)) << (seg * SEG);
    const std::uint64_t lt = ((1ull << lane) - 1ull) & segmask;
    const std::uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const std::uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    const std::uint32_t S = 6u * smp;
    std::uint32_t* hkey = s_hkey[wib][seg];
    std::uint32_t* hidx = s_hidx[wib][seg];
    for (std::uint32_t v0 = wave * P; v0 < n; v0 += waves * P) {
        const std::uint32_t v = v0 + seg;
        const bool live = v < n;
        const std::uint32_t nn = live ? new_cnt[v] : 0u;
        const std::uint32_t rn = live ? min(rev_new_cnt[v], smp) : 0u;
        const std::uint32_t no = live ? old_cnt[v] : 0u;
        const std::uint32_t ro = live ? min(rev_old_cnt[v], smp) : 0u;
        const std::uint32_t na0 = nn + rn;
        // a node without new members has no pairs to join: nothing to gather
        const std::uint32_t ns0 = na0 ? na0 + no + ro : 0u;
        if (ns0)
            for (std::uint32_t h = sub; h < H; h += SEG) {
                hkey[h] = kNone;
                hidx[h] = kNone;
            }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // entry e = sub + r SEG: its id, its threshold and its slot in the
        // table; all ids are loaded, then all thresholds, before the table's
        // LDS atomics, so each kind of load waits for memory once
        std::uint32_t id[R], slot[R];
        float th[R];
#pragma unroll
        for (std::uint32_t r = 0; r < R; ++r) {
            const std::uint32_t e = sub + r * SEG;
            std::uint32_t x = kNone;
            if (e < ns0) {
                if (e < nn) x = new_list[std::size_t(v) * smp + e];
                else if (e < na0) x = rev_new[std::size_t(v) * smp + (e - nn)];
                else if (e < na0 + no) x = old_list[std::size_t(v) * 3 * smp + (e - na0)];
                else x = rev_old[std::size_t(v) * smp + (e - na0 - no)];
            }
            id[r] = x < n ? x : kNone;
        }
#pragma unroll
        for (std::uint32_t r = 0; r < R; ++r) th[r] = id[r] != kNone ? thr_node[id[r]] : 0.0f;
#pragma unroll
        for (std::uint32_t r = 0; r < R; ++r) {
            const std::uint32_t e = sub + r * SEG;
            slot[r] = 0;
            if (id[r] != kNone) {
                std::uint32_t h = (id[r] * 0x9E3779B1u) >> 24;
                for (;;) {
                    const std::uint32_t prev = atomicCAS(&hkey[h], kNone, id[r]);
                    if (prev == kNone || prev == id[r]) break;
                    h = (h + 1) & (H - 1);
                }
                atomicMin(&hidx[h], e);
                slot[r] = h;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // first copies, in order; members from the A part (positions < na0)
        std::uint32_t na = 0, ns = 0;
#pragma unroll
        for (std::uint32_t r = 0; r < R; ++r) {
            const std::uint32_t e = sub + r * SEG;
            const bool keep = id[r] != kNone && hidx[slot[r]] == e;
            const std::uint64_t m = __ballot(keep) & segmask;
            const std::uint32_t pos = ns + __popcll(m & lt);
            if (keep) {
                out_ids[std::size_t(v) * S + pos] = id[r];
                out_thr[std::size_t(v) * S + pos] = th[r];
            }
            na += __popcll(__ballot(keep && e < na0) & segmask);
            ns += __popcll(m);
        }
        if (sub == 0 && live) out_cnt[v] = na | (ns << 16);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}


struct JoinArgs {
    const float* pts;
    std::uint32_t n, stride, smp;
    const std::uint32_t* ids;  // gathered members, n x S
    const float* thr;          // their thresholds, n x S
    const std::uint32_t* cnt;  // na | ns << 16
    unsigned long long* cand;  // n x cap proposal buffer
    std::uint32_t* fill;       // proposals offered per node this round
    std::uint32_t cap;
    // Gram-filter join only: work tallies (pairs, exact recomputations) and
    // the audit maxima (fast-distance relative error vs FP64, error / bound),
    // both nullptr when not profiling
    unsigned long long* tally;
    unsigned int* audit;
    const std::uint32_t* halt;  // nonzero: the descent has converged, skip
    bool cosine;                // stored values are r / 2 instead of sqrt(r)
    // filter join only: nodes of more than row_cap members are set aside in
    // `defer` (count in *defer_count) for a second pass that stages up to S
    // rows per node (deferred_pass); 0 = every node in one pass
    std::uint32_t row_cap = 0;
    std::uint32_t* defer = nullptr;
    std::uint32_t* defer_count = nullptr;
    bool deferred_pass = false;
    // member-stationary join only: entries per wavefront of the LDS list of
    // passing pairs (the rest are recomputed in the emit pass)
    std::uint32_t list_cap = 0;
};

// Exact squared distance with `a` held in registers (compile-time stride),
// same summation order as l2sqr_exact.
template <std::uint32_t STRIDE>
__device__ __forceinline__ float l2sqr_reg(const float (&a)[STRIDE],
                                           const float* __restrict__ b) {
    float acc0[8], acc1[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        acc0[j] = 0.0f;
        acc1[j] = 0.0f;
    }
#pragma unroll
    for (std::uint32_t i = 0; i + 16 <= STRIDE; i += 16) {
        const float4* pb = reinterpret_cast<const float4*>(b + i);
        const float4 b0 = pb[0], b1 = pb[1], b2 = pb[2], b3 = pb[3];
        const float bb[16] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w,
                              b2.x, b2.y, b2.z, b2.w, b3.x, b3.y, b3.z, b3.w};
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float d0 = a[i + j] - bb[j];
            const float d1 = a[i + 8 + j] - bb[8 + j];
            acc0[j] = __builtin_fmaf(d0, d0, acc0[j]);
            acc1[j] = __builtin_fmaf(d1, d1, acc1[j]);
        }
    }
    if constexpr (STRIDE % 16 == 8) {
        constexpr std::uint32_t i = STRIDE - 8;
        const float4* pb = reinterpret_cast<const float4*>(b + i);
        const float4 b0 = pb[0], b1 = pb[1];
        const float bb[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float d0 = a[i + j] - bb[j];
            acc0[j] = __builtin_fmaf(d0, d0, acc0[j]);
        }
    }
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = acc0[j] + acc1[j];
    const float l0 = v[0] + v[4], l1 = v[1] + v[5], l2 = v[2] + v[6], l3 = v[3] + v[7];
    return (l0 + l1) + (l2 + l3);
}

// l2sqr_reg's value from the same operations on each accumulator in the same
// order, computed four accumulator lanes at a time: accumulators j = 4q .. 4q+3
// take coordinates 16c + j and 16c + 8 + j of every 16-wide chunk c, then
// the 8-wide tail, for q = 0 and then q = 1. An empty register fence after
// the first half keeps its sums computed, and the second half's loads
// issued, on either side of it, so only half of the other row is held in
// registers at once.
template <std::uint32_t STRIDE>
__device__ __forceinline__ float l2sqr_reg_halves(const float (&a)[STRIDE],
                                                  const float* __restrict__ b) {
    constexpr std::uint32_t M = STRIDE / 16;
    float v[8];
#pragma unroll
    for (std::uint32_t q = 0; q < 2; ++q) {
        float acc0[4], acc1[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            acc0[u] = 0.0f;
            acc1[u] = 0.0f;
        }
#pragma unroll
        for (std::uint32_t c = 0; c < M; ++c) {
            const float4 b0 = *reinterpret_cast<const float4*>(b + 16 * c + 4 * q);
            const float4 b1 = *reinterpret_cast<const float4*>(b + 16 * c + 8 + 4 * q);
            const float x0[4] = {b0.x, b0.y, b0.z, b0.w};
            const float x1[4] = {b1.x, b1.y, b1.z, b1.w};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float d0 = a[16 * c + 4 * q + u] - x0[u];
                const float d1 = a[16 * c + 8 + 4 * q + u] - x1[u];
                acc0[u] = __builtin_fmaf(d0, d0, acc0[u]);
                acc1[u] = __builtin_fmaf(d1, d1, acc1[u]);
            }
        }
        if constexpr (STRIDE % 16 == 8) {
            const float4 b0 = *reinterpret_cast<const float4*>(b + 16 * M + 4 * q);
            const float x0[4] = {b0.x, b0.y, b0.z, b0.w};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float d0 = a[16 * M + 4 * q + u] - x0[u];
                acc0[u] = __builtin_fmaf(d0, d0, acc0[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) v[4 * q + u] = acc0[u] + acc1[u];
        if (q == 0) asm volatile("" : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3])::"memory");
    }
    const float l0 = v[0] + v[4], l1 = v[1] + v[5], l2 = v[2] + v[6], l3 = v[3] + v[7];
    return (l0 + l1) + (l2 + l3);
}

// The member-stationary join's exact distance: l2sqr_reg_halves where its
// lower register count pays, l2sqr_reg at stride 48 (measured faster there).
template <std::uint32_t STRIDE>
__device__ __forceinline__ float l2sqr_join(const float (&a)[STRIDE], const float* __restrict__ b) {
    if constexpr (STRIDE == 48) return l2sqr_reg<STRIDE>(a, b);
    else return l2sqr_reg_halves<STRIDE>(a, b);
}

// Shared phases of the two local-join kernels. One 128-thread group (two
// wavefronts) per node. LDS: ids[S], thr[S], rows[S*stride] and the distance
// tile D[A][S+1] (odd pitch: rows and columns are both read without bank
// conflicts).

// Stage the node's members (ids, thresholds, rows). Returns false (and touches
// nothing) for a node without new members.
template <std::uint32_t NT = kJoinThreads>
__device__ __forceinline__ bool join_stage(const JoinArgs& a, std::uint32_t v, std::uint32_t S,
                                           std::uint32_t stride, std::uint32_t* s_id,
                                           float* s_thr, float* s_rows, std::uint32_t* na_out,
                                           std::uint32_t* ns_out) {
    const std::uint32_t tid = threadIdx.x;
    const std::uint32_t cnt = a.cnt[v];
    const std::uint32_t na = cnt & 0xFFFFu, ns = cnt >> 16;
    *na_out = na;
    *ns_out = ns;
    if (na == 0) return false;  // block-uniform
    for (std::uint32_t i = tid; i < ns; i += NT) {
        s_id[i] = a.ids[std::size_t(v) * S + i];
        s_thr[i] = a.thr[std::size_t(v) * S + i];
    }
    __syncthreads();
    const std::uint32_t q4 = stride / 4;
    for (std::uint32_t e = tid; e < ns * q4; e += NT) {
        const std::uint32_t i = e / q4, f = e - i * q4;
        reinterpret_cast<float4*>(s_rows)[e] =
            reinterpret_cast<const float4*>(a.pts + std::size_t(s_id[i]) * stride)[f];
    }
    __syncthreads();
    return true;
}

// Per member count ns <= kJoinThreads: ceil(65536 / ns), and kJoinThreads / ns,
// read from constant memory (ns is the same in the whole block, so these are
// scalar loads instead of a division on the vector unit).
struct JoinDivTable {
    std::uint32_t recip16[kJoinThreads + 1];
    std::uint32_t per_member[kJoinThreads + 1];
    constexpr JoinDivTable() : recip16(), per_member() {
        for (std::uint32_t ns = 1; ns <= kJoinThreads; ++ns) {
            recip16[ns] = (65536u + ns - 1) / ns;
            per_member[ns] = kJoinThreads / ns;
        }
    }
};
__constant__ JoinDivTable c_join_div = JoinDivTable();

// join_stage for the member-stationary join: thread t stages float4 t % Q of
// members t / Q, + kJoinThreads / Q, ... (Q = STRIDE / 4 float4 per row; the
// kJoinThreads % Q threads left over idle), so the index arithmetic is fixed
// per thread instead of a division per float4.
template <std::uint32_t STRIDE>
__device__ __forceinline__ bool join_stage_reg(const JoinArgs& a, std::uint32_t v,
                                               std::uint32_t S, std::uint32_t* s_id,
                                               float* s_thr, float* s_rows,
                                               std::uint32_t* na_out, std::uint32_t* ns_out) {
    constexpr std::uint32_t Q = STRIDE / 4, MSTEP = kJoinThreads / Q;
    const std::uint32_t tid = threadIdx.x;
    const std::uint32_t cnt = a.cnt[v];
    const std::uint32_t na = cnt & 0xFFFFu, ns = cnt >> 16;
    *na_out = na;
    *ns_out = ns;
    if (na == 0) return false;  // block-uniform
    for (std::uint32_t i = tid; i < ns; i += kJoinThreads) {
        s_id[i] = a.ids[std::size_t(v) * S + i];
        s_thr[i] = a.thr[std::size_t(v) * S + i];
    }
    __syncthreads();
    const std::uint32_t qi = tid / Q, qf = tid - qi * Q;
    const float4* src = reinterpret_cast<const float4*>(a.pts);
    float4* dst = reinterpret_cast<float4*>(s_rows);
    if (qi < MSTEP)
        for (std::uint32_t m = qi; m < ns; m += MSTEP)
            dst[m * Q + qf] = src[std::size_t(s_id[m]) * Q + qf];
    __syncthreads();
    return true;
}

// A thread's walk over the (new i, member j) pairs p = i * ns + j, p = p0,
// p0 + step, ...: (i, j) starts at (p0 / ns, p0 % ns) and advances by
// (step / ns, step % ns) with one carry, so no pair costs a division. The
// quotients come from a rounded-up reciprocal, exact for arguments below 2^14
// and ns <= 192. Unordered new x new pairs are visited once (i < j).
struct PairWalk {
    std::uint32_t i, j, di, dj, ns, na;
    __device__ PairWalk(std::uint32_t p0, std::uint32_t step, std::uint32_t ns_,
                        std::uint32_t na_)
        : ns(ns_), na(na_) {
        const std::uint64_t recip = ((1ull << 24) + ns_ - 1) / ns_;
        i = static_cast<std::uint32_t>((std::uint64_t(p0) * recip) >> 24);
        j = p0 - i * ns_;
        di = static_cast<std::uint32_t>((std::uint64_t(step) * recip) >> 24);
        dj = step - di * ns_;
    }
    __device__ bool live() const { return !(j < na && j <= i); }
    __device__ void next() {
        i += di;
        j += dj;
        if (j >= ns) {
            j -= ns;
            ++i;
        }
    }
};

// Proposals of one node from its pair distances (the reference's try_pair
// test, nndescent.cpp:41-44, both ends): pair p = (i, j) with distance d
// proposes j to i when d < thr_i and i to j when d < thr_j. The distance
// pass has counted, per member, the proposals it receives (LDS atomics into
// s_cnt); one global atomic per receiving member reserves its range of the
// round's buffer, then the pairs are visited again and each key is written at
// the member's base plus an LDS cursor. Keys past a member's capacity in this
// round are dropped.
// Distances are read as s_D[p] (pitch 0) or from a tile, s_D[i * pitch + j].
__device__ __forceinline__ void join_emit(const JoinArgs& a, const std::uint32_t* s_id,
                                          const float* s_thr, const float* s_D,
                                          std::uint32_t pitch, std::uint32_t* s_cnt,
                                          std::uint32_t* s_base, std::uint32_t na,
                                          std::uint32_t ns) {
    const std::uint32_t tid = threadIdx.x;
    for (std::uint32_t x = tid; x < ns; x += kJoinThreads) {
        const std::uint32_t c = s_cnt[x];
        s_base[x] = c ? atomicAdd(&a.fill[s_id[x]], c) : a.cap;
        s_cnt[x] = 0;  // now the write cursor
    }
    __syncthreads();
    PairWalk w(tid, kJoinThreads, ns, na);
    for (std::uint32_t p = tid; p < na * ns; p += kJoinThreads, w.next()) {
        if (!w.live()) continue;
        const std::uint32_t i = w.i, j = w.j;
        const float d = s_D[pitch ? __umul24(i, pitch) + j : p];
        const bool to_i = d < s_thr[i], to_j = d < s_thr[j];
        if (!(to_i || to_j)) continue;
        const unsigned long long hb =
            static_cast<unsigned long long>(fbits(metric_finalize(d, a.cosine))) << 32;
        if (to_i) {
            const std::uint32_t pos = s_base[i] + atomicAdd(&s_cnt[i], 1u);
            if (pos < a.cap) a.cand[std::size_t(s_id[i]) * a.cap + pos] = hb | s_id[j];
        }
        if (to_j) {
            const std::uint32_t pos = s_base[j] + atomicAdd(&s_cnt[j], 1u);
            if (pos < a.cap) a.cand[std::size_t(s_id[j]) * a.cap + pos] = hb | s_id[i];
        }
    }
}

// Exact local join: every (new i, member j) pair's distance is computed in
// the reference's exact summation order (both rows from LDS), the pairs
// spread densely over the threads; each pair's distance is kept by pair index
// and its proposals counted on the spot. LDS: ids[S], thr[S], cnt[S],
// base[S], rows[S*stride], distances[A*S].
template <std::uint32_t STRIDE>
__global__ void __launch_bounds__(kJoinThreads) join_kernel(JoinArgs a) {
    extern __shared__ float lds[];
    if (*a.halt) return;
    const std::uint32_t S = 6u * a.smp;
    const std::uint32_t stride = STRIDE ? STRIDE : a.stride;
    std::uint32_t* s_id = reinterpret_cast<std::uint32_t*>(lds);
    float* s_thr = lds + S;
    std::uint32_t* s_cnt = reinterpret_cast<std::uint32_t*>(lds + 2 * S);
    std::uint32_t* s_base = s_cnt + S;
    float* s_rows = lds + 4 * S;
    float* s_Dp = s_rows + std::size_t(S) * stride;
    const std::uint32_t tid = threadIdx.x;
    for (std::uint32_t v = blockIdx.x; v < a.n; v += gridDim.x) {
        std::uint32_t na, ns;
        for (std::uint32_t x = tid; x < S; x += kJoinThreads) s_cnt[x] = 0;
        if (!join_stage(a, v, S, stride, s_id, s_thr, s_rows, &na, &ns)) continue;
        PairWalk w(tid, kJoinThreads, ns, na);
        for (std::uint32_t p = tid; p < na * ns; p += kJoinThreads, w.next()) {
            if (!w.live()) continue;
            const std::uint32_t i = w.i, j = w.j;
            const float d = l2sqr_exact(s_rows + __umul24(i, stride),
                                        s_rows + __umul24(j, stride), stride);
            s_Dp[p] = d;
            if (d < s_thr[i]) atomicAdd(&s_cnt[i], 1u);
            if (d < s_thr[j]) atomicAdd(&s_cnt[j], 1u);
        }
        __syncthreads();
        join_emit(a, s_id, s_thr, s_Dp, 0, s_cnt, s_base, na, ns);
        __syncthreads();
    }
}

// The t-th partner of member j in the member-stationary join: new member t,
// or new member j - 1 - t (mod na) in the circular order.
__device__ __forceinline__ std::uint32_t join_partner(bool circ, std::uint32_t j,
                                                      std::uint32_t na, std::uint32_t t) {
    if (!circ) return t;
    const std::uint32_t o = t + 1;
    return j >= o ? j - o : j + na - o;
}

// Largest value over the wavefront's lanes.
__device__ __forceinline__ unsigned int wave_max_u32(unsigned int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned int o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// Number of set bits of a wavefront mask below the calling lane.
__device__ __forceinline__ std::uint32_t lanes_below(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<std::uint32_t>(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<std::uint32_t>(m), 0u));
}

// Exact local join, member-stationary, for rows of at most kJoinRegStride
// floats and at most kJoinThreads members: thread t keeps member j = t % ns
// (its row, threshold and proposal count) in registers and visits the new
// rows i = t / ns, + G, ... (G = kJoinThreads / ns; the threads left over
// give the last members one more phase), so a pair reads one row from LDS.
// A visited pair that passes a threshold test is appended, with its exact
// distance, to its wavefront's LDS list (list_cap entries) at the ballot's
// prefix count over a running length that is the same in every lane still
// visiting (no atomics); after the per-member reservation each wavefront
// writes its list's keys one entry per lane. Passing pairs beyond the list's
// capacity keep one bit per visited pair and direction (Mask: 32 bits when a
// lane visits at most 2 smp <= 32 pairs), and the emit pass recomputes their
// distances (the same function on the same rows: the same bits). With no
// distance tile a node needs LDS for ids[S], thr[S], cnt[S], base[S],
// rows[S*STRIDE] and the two lists only.
// (At stride 64 the register budget of four waves per SIMD is enforced: it
// spills two loop-invariant registers, and three waves would cost more.)
template <std::uint32_t STRIDE, typename Mask>
__global__ void __launch_bounds__(kJoinThreads)
    __attribute__((amdgpu_waves_per_eu(STRIDE == 64 ? 4 : 1))) join_reg_kernel(JoinArgs a) {
    static_assert(STRIDE % 8 == 0 && STRIDE <= kJoinRegStride, "row kept in registers");
    extern __shared__ float lds[];
    if (*a.halt) return;
    const std::uint32_t S = 6u * a.smp;  // <= kJoinThreads (host-checked)
    std::uint32_t* s_id = reinterpret_cast<std::uint32_t*>(lds);
    float* s_thr = lds + S;
    std::uint32_t* s_cnt = reinterpret_cast<std::uint32_t*>(lds + 2 * S);
    std::uint32_t* s_base = s_cnt + S;
    float* s_rows = lds + 4 * S;
    const std::uint32_t tid = threadIdx.x, lane = tid & 63u;
    const std::uint32_t L = a.list_cap;
    // entry: exact distance bits, i | j << 8 | to_i << 16 | to_j << 17
    uint2* s_list = reinterpret_cast<uint2*>(s_rows + S * STRIDE) +
                    __builtin_amdgcn_readfirstlane(tid >> 6) * L;
    for (std::uint32_t v = blockIdx.x; v < a.n; v += gridDim.x) {
        std::uint32_t na, ns;
        for (std::uint32_t x = tid; x < S; x += kJoinThreads) s_cnt[x] = 0;
        if (!join_stage_reg<STRIDE>(a, v, S, s_id, s_thr, s_rows, &na, &ns)) continue;
        // G0 = 128 / ns threads per member (i0 = 0..G0-1, step G0); the
        // E = 128 - G0 ns threads left over become an extra phase (i0 = G0)
        // of the last E members, which then step by G0 + 1: every thread has
        // a member
        const std::uint32_t G0 = c_join_div.per_member[ns], E = kJoinThreads - G0 * ns;
        // t / ns by a rounded-up 16-bit reciprocal: exact for t < 128, ns < 500
        std::uint32_t i0 = (tid * c_join_div.recip16[ns]) >> 16;
        std::uint32_t j = tid - i0 * ns;
        if (i0 >= G0) {  // an extra thread
            j = ns - E + (tid - G0 * ns);
            i0 = G0;
        }
        const std::uint32_t G = j >= ns - E ? G0 + 1 : G0;
        // partners of member j, t = 0, 1, ..: every new member for an old j;
        // for a new j the new members i < j, except in a node without old
        // members (the first round: all pairs new x new), where a new j takes
        // j - 1 - t (mod na), (na - 1) / 2 of them plus, for an even na, the
        // opposite one when j >= na / 2: every unordered pair once, about
        // na / 2 per member instead of j. (With old members the lanes of a
        // wavefront would then read different rows in the same step instead
        // of one broadcast row.)
        const bool circ = j < na && na == ns;
        const std::uint32_t iend =
            circ ? (na - 1) / 2 + ((na % 2 == 0 && j >= na / 2) ? 1u : 0u) : (j < na ? j : na);
        float rj[STRIDE];
        // bit r: pair (partner i0 + r G, j) proposes and did not fit the list
        Mask to_i = 0, to_j = 0;
        // the list's running length: equal in the lanes still visiting,
        // never above it in the lanes that left
        std::uint32_t run = 0, cj = 0;
        {
            const float4* src = reinterpret_cast<const float4*>(s_rows + j * STRIDE);
#pragma unroll
            for (std::uint32_t q = 0; q < STRIDE / 4; ++q) {
                const float4 f = src[q];
                rj[4 * q] = f.x;
                rj[4 * q + 1] = f.y;
                rj[4 * q + 2] = f.z;
                rj[4 * q + 3] = f.w;
            }
            const float thr_j = s_thr[j];
            // the partner order is block-uniform (na == ns): one loop each
            auto visit = [&](std::uint32_t i, Mask bit) {
                const float d = l2sqr_join<STRIDE>(rj, s_rows + i * STRIDE);
                const bool pi = d < s_thr[i], pj = d < thr_j;
                if (pi) atomicAdd(&s_cnt[i], 1u);
                cj += pj ? 1u : 0u;
                const bool any = pi || pj;
                const unsigned long long bal = __ballot(any);
                const std::uint32_t pos = run + lanes_below(bal);
                run += static_cast<std::uint32_t>(__popcll(bal));
                if (any) {
                    if (pos < L) {
                        s_list[pos] = make_uint2(fbits(d), i | (j << 8) | (pi ? 0x10000u : 0u) |
                                                               (pj ? 0x20000u : 0u));
                    } else {
                        if (pi) to_i |= bit;
                        if (pj) to_j |= bit;
                    }
                }
            };
            Mask bit = 1;
            if (na == ns) {
                for (std::uint32_t t = i0; t < iend; t += G, bit <<= 1)
                    visit(join_partner(true, j, na, t), bit);
            } else {
                for (std::uint32_t t = i0; t < iend; t += G, bit <<= 1) visit(t, bit);
            }
            if (cj) atomicAdd(&s_cnt[j], cj);
        }
        run = wave_max_u32(run);
        __syncthreads();
        // one reservation per receiving member, then write at LDS cursors
        for (std::uint32_t x = tid; x < ns; x += kJoinThreads) {
            const std::uint32_t c = s_cnt[x];
            s_base[x] = c ? atomicAdd(&a.fill[s_id[x]], c) : a.cap;
            s_cnt[x] = 0;
        }
        __syncthreads();
        // the wavefront's listed pairs, one per lane
        const std::uint32_t len = run < L ? run : L;
        for (std::uint32_t e = lane; e < len; e += 64) {
            const uint2 en = s_list[e];
            const std::uint32_t ii = en.y & 0xFFu, jj = (en.y >> 8) & 0xFFu;
            const unsigned long long hb =
                static_cast<unsigned long long>(
                    fbits(metric_finalize(__uint_as_float(en.x), a.cosine)))
                << 32;
            if (en.y & 0x10000u) {
                const std::uint32_t pos = s_base[ii] + atomicAdd(&s_cnt[ii], 1u);
                if (pos < a.cap) a.cand[std::size_t(s_id[ii]) * a.cap + pos] = hb | s_id[jj];
            }
            if (en.y & 0x20000u) {
                const std::uint32_t pos = s_base[jj] + atomicAdd(&s_cnt[jj], 1u);
                if (pos < a.cap) a.cand[std::size_t(s_id[jj]) * a.cap + pos] = hb | s_id[ii];
            }
        }
        // the passing pairs the list did not hold: recomputed
        if (to_i | to_j) {
            const std::uint32_t id_j = s_id[j];
            std::uint32_t pj = to_j ? s_base[j] + atomicAdd(&s_cnt[j],
                                                            static_cast<std::uint32_t>(__popcll(to_j)))
                                    : 0u;
            Mask m = to_i | to_j;
            while (m) {
                const std::uint32_t r =
                    static_cast<std::uint32_t>(__ffsll(static_cast<unsigned long long>(m)) - 1);
                m &= m - 1;
                const std::uint32_t i = join_partner(circ, j, na, i0 + r * G);
                // a direction whose range starts at the capacity writes
                // nothing: skip it, and the distance when both do
                const bool wi = ((to_i >> r) & 1u) && s_base[i] < a.cap;
                const bool dj = (to_j >> r) & 1u;
                const bool wj = dj && pj < a.cap;
                if (wi || wj) {
                    const float d = l2sqr_join<STRIDE>(rj, s_rows + i * STRIDE);
                    const unsigned long long hb =
                        static_cast<unsigned long long>(fbits(metric_finalize(d, a.cosine))) << 32;
                    if (wi) {
                        const std::uint32_t pos = s_base[i] + atomicAdd(&s_cnt[i], 1u);
                        if (pos < a.cap) a.cand[std::size_t(s_id[i]) * a.cap + pos] = hb | id_j;
                    }
                    if (wj) a.cand[std::size_t(id_j) * a.cap + pj] = hb | s_id[i];
                }
                if (dj) ++pj;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Gram-filter local join (matrix cores)
// ---------------------------------------------------------------------------
//
// The rows of a node's members are centred on its first member, rounded to
// FP16 (or BF16) and contracted on the matrix cores (16x16x16 MFMA, FP32
// accumulation): D~(i,j) = N_i + N_j - 2 G_ij with N the squared norms of
// the rounded rows. A rigorous bound turns D~ into a lower bound on the
// exact distance r that the reference's summation would produce:
//   |S - T| <= delta = rho (|a_i| + |a_j|) + 2 abs     (rounding of the rows)
//   |D~ - S^2| <= eps = cG (|a_i| + |a_j|)^2           (FP32 accumulation)
//   r >= T^2 (1 - eta)                                  (the exact sum's rounding)
// with T the true distance, S the distance of the rounded rows, rho the
// format's unit roundoff (plus the centring's), abs its flush-to-zero floor per
// coordinate (times sqrt(K)), cG = 2 (K + 8) 2^-24 and eta = (stride + 24)
// 2^-24 + 2^-18; every bound is rounded outward. A pair whose lower bound
// reaches both members' thresholds cannot pass either try_pair test and gets
// +inf; every other pair gets its exact distance. The proposals are therefore
// exactly those of the exact join: the matrix cores only decide which exact
// distances are needed.

// Eight coordinates k..k+7 of LDS row r centred on row 0 (zeros past the
// stride or for rows at or beyond `rows`).
template <std::uint32_t STRIDE>
__device__ __forceinline__ void centred8(const float* s_rows, std::uint32_t r, std::uint32_t rows,
                                         std::uint32_t k, float (&out)[8]) {
    if (r < rows && k < STRIDE) {
        const float4* x = reinterpret_cast<const float4*>(s_rows + __umul24(r, STRIDE) + k);
        const float4* c = reinterpret_cast<const float4*>(s_rows + k);
        const float4 x0 = x[0], x1 = x[1], c0 = c[0], c1 = c[1];
        out[0] = x0.x - c0.x;
        out[1] = x0.y - c0.y;
        out[2] = x0.z - c0.z;
        out[3] = x0.w - c0.w;
        out[4] = x1.x - c1.x;
        out[5] = x1.y - c1.y;
        out[6] = x1.z - c1.z;
        out[7] = x1.w - c1.w;
    } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) out[q] = 0.0f;
    }
}

// Matrix-core filter join. One 256-thread group (four wavefronts) per node.
//
// 1. Staging as in the exact joins; per member, the squared norm N of its
//    centred row rounded to FP16 / BF16, an outward bound nb on the norms of
//    the centred row and of its rounding, and st = sqrt(thr / (1 - eta))
//    rounded up (the distance a pair must exceed to fail the member's
//    try_pair test, whatever the exact sum's rounding).
// 2. Gram tiles of 16 new rows x 16 members, one 16x16x32 MFMA per 32
//    coordinates, D~ = N_i + N_j - 2 G. With S the distance of the rounded
//    rows and T the true one, |S - T| <= delta = rho (nb_i + nb_j) + 2 abs
//    and |D~ - S^2| <= eps = cG (nb_i + nb_j)^2, so D~ - eps >= (delta +
//    max(st_i, st_j))^2 proves that the exact distance fails both tests.
//    Every other live pair (the exact joins' visited pairs: new x all, each
//    new x new pair once) is appended to an LDS list by ballot prefix.
// 3. The listed pairs get their exact distance, densely over the threads
//    (the reference's summation order), and their proposals are counted.
// 4. One reservation per receiving member, then every listed pair's keys are
//    written from the stored distance: nothing is computed twice.
// The proposals are exactly the exact joins'. Tiles are taken in batches
// whose pair count fits the list (one batch for all but the largest nodes).
// LDS: ids[S], thr[S], cnt[S], base[S], (N, nb, st, -)[S], rows[S*STRIDE],
// list idx u16[L], list distance f32[L], list length.
// Nodes per CU the filter join is compiled for: its LDS per node at sample
// 16 allows about this many; the register budget follows from it.
constexpr unsigned mfma_join_blocks(std::uint32_t stride) {
    return stride <= 24 ? 6u : 4u;
}

// AUDIT: the FP64 check of every live pair (a.audit set) is compiled in.
template <std::uint32_t STRIDE, bool BF16, bool AUDIT>
__global__ void __launch_bounds__(mfma_join_threads(STRIDE), mfma_join_blocks(STRIDE))
    join_mfma_kernel(JoinArgs a) {
    static_assert(STRIDE % 8 == 0, "rows padded to 8 floats");
    constexpr std::uint32_t KP = (STRIDE + 31u) & ~31u;
    constexpr std::uint32_t kMfmaThreads = mfma_join_threads(STRIDE);
    extern __shared__ float lds[];
    if (*a.halt) return;
    const std::uint32_t S = 6u * a.smp;
    std::uint32_t* s_id = reinterpret_cast<std::uint32_t*>(lds);
    float* s_thr = lds + S;
    std::uint32_t* s_cnt = reinterpret_cast<std::uint32_t*>(lds + 2 * S);
    std::uint32_t* s_base = s_cnt + S;
    float4* s_m4 = reinterpret_cast<float4*>(lds + 4 * S);  // S multiple of 2: 16-B aligned
    float* s_rows = lds + 8 * S;
    // rows staged per node: row_cap in a first pass (nodes with more members
    // are set aside), S in the pass over the set-aside nodes
    const std::uint32_t rows_alloc = a.row_cap && !a.deferred_pass ? a.row_cap : S;
    float* s_ld = s_rows + std::size_t(rows_alloc) * STRIDE;
    std::uint16_t* s_li = reinterpret_cast<std::uint16_t*>(s_ld + kMfmaList);
    std::uint32_t* s_len = reinterpret_cast<std::uint32_t*>(s_li + kMfmaList);
    const std::uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const float rho = GramConst<BF16>::rho * (1.0f + 0x1p-16f);
    const float abs2 = 2.0f * GramConst<BF16>::absc * sqrtf(float(KP)) * 1.01f * (1.0f + 0x1p-16f);
    const float abs_norm = GramConst<BF16>::absc * sqrtf(float(KP)) * 1.01f;
    const float cG = 2.0f * float(KP + 8) * 0x1p-24f * (1.0f + 0x1p-16f);
    const float eta = float(STRIDE + 24) * 0x1p-24f + 0x1p-18f;
    const float nscale = 1.0f + float(KP + 4) * 0x1p-24f;
    const float sscale = (1.0f + 0x1p-20f) / (1.0f - eta);
    unsigned long long n_pairs = 0, n_exact = 0, n_tiles = 0;
    unsigned int worst_rel = 0, worst_ratio = 0;
    const std::uint32_t nodes = a.deferred_pass ? *a.defer_count : a.n;
    for (std::uint32_t it = blockIdx.x; it < nodes; it += gridDim.x) {
        const std::uint32_t v = a.deferred_pass ? a.defer[it] : it;
        if (a.row_cap && !a.deferred_pass) {
            const std::uint32_t c = a.cnt[v];
            if ((c & 0xFFFFu) != 0 && (c >> 16) > a.row_cap) {  // block-uniform
                if (tid == 0) a.defer[atomicAdd(a.defer_count, 1u)] = v;
                continue;
            }
        }
        std::uint32_t na, ns;
        if (!join_stage<kMfmaThreads>(a, v, S, STRIDE, s_id, s_thr, s_rows, &na, &ns)) continue;
        // 1. member constants
        for (std::uint32_t j = tid; j < ns; j += kMfmaThreads) {
            float N = 0.0f;
#pragma unroll
            for (std::uint32_t k = 0; k < STRIDE; k += 8) {
                float c8[8];
                centred8<STRIDE>(s_rows, j, ns, k, c8);
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const float h = gram_round<BF16>(c8[q]);
                    N = __builtin_fmaf(h, h, N);
                }
            }
            const float nb = (sqrtf(N) * nscale + abs_norm) * (1.0f + 2.0f * GramConst<BF16>::rho);
            const float st = sqrtf(s_thr[j] * sscale) * (1.0f + 0x1p-20f);
            s_m4[j] = make_float4(N, nb, st, 0.0f);
        }
        const std::uint32_t rt = (na + 15u) / 16u, ct = (ns + 15u) / 16u, T = rt * ct;
        // pairs a tile can hold: none when every member column is new and at
        // or below every row (new x new pairs are taken with i < j)
        // t / ct by a rounded-up 16-bit reciprocal: exact for t < 2^16 / ct^2
        // (T <= 4 * 12 tiles at sample 32)
        const std::uint32_t ct_inv = (65536u + ct - 1u) / ct;
        auto tile_row = [&](std::uint32_t t) -> std::uint32_t { return (t * ct_inv) >> 16; };
        auto tile_cap = [&](std::uint32_t t) -> std::uint32_t {
            const std::uint32_t tr = tile_row(t);
            const std::uint32_t r0 = 16u * tr, c0 = 16u * (t - tr * ct);
            if (c0 + 16u <= na && c0 + 15u <= r0) return 0u;
            const std::uint32_t rr = na - r0 < 16u ? na - r0 : 16u;
            const std::uint32_t cc = ns - c0 < 16u ? ns - c0 : 16u;
            return rr * cc;
        };
        for (std::uint32_t t0 = 0; t0 < T;) {
            std::uint32_t t1 = t0, fill = 0;
            while (t1 < T && fill + tile_cap(t1) <= kMfmaList) fill += tile_cap(t1++);
            for (std::uint32_t x = tid; x < ns; x += kMfmaThreads) s_cnt[x] = 0;
            if (tid == 0) *s_len = 0;
            __syncthreads();
            // 2. filter: wavefront w takes tiles t0 + w, + 4, ...
            const std::uint32_t li = lane & 15u, kq = lane >> 4;
            for (std::uint32_t t = t0 + w; t < t1; t += kMfmaThreads / 64u) {
                if (tile_cap(t) == 0) continue;  // wave-uniform
                const std::uint32_t tr = tile_row(t);
                const std::uint32_t r0 = 16u * tr, c0 = 16u * (t - tr * ct);
                if (a.tally && lane == 0) ++n_tiles;
                gram_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (std::uint32_t k0 = 0; k0 < KP; k0 += 32) {
                    float x[8], y[8];
                    centred8<STRIDE>(s_rows, r0 + li, na, k0 + 8u * kq, x);
                    centred8<STRIDE>(s_rows, c0 + li, ns, k0 + 8u * kq, y);
                    acc = gram_mfma32<BF16>(x, y, acc);
                }
                const std::uint32_t j = c0 + li;
                const float4 mj = s_m4[j < ns ? j : 0];
                bool keep[4];
                std::uint32_t total = 0;
                std::uint64_t bal[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const std::uint32_t i = r0 + 4u * kq + static_cast<std::uint32_t>(q);
                    const bool live = i < na && j < ns && !(j < na && j <= i);
                    bool kp = false;
                    if (live) {
                        const float4 mi = s_m4[i];
                        const float dt = (mi.x + mj.x) - 2.0f * acc[q];
                        const float nbs = mi.y + mj.y;
                        const float qd = __builtin_fmaf(rho, nbs, abs2) + fmaxf(mi.z, mj.z);
                        const float lhs = dt - cG * nbs * nbs;
                        const bool reject = lhs >= qd * qd * (1.0f + 0x1p-18f);
                        kp = !reject;
                        if (a.tally) {
                            ++n_pairs;
                            n_exact += kp ? 1u : 0u;
                        }
                        if (AUDIT && a.audit) {
                            const float* ri = s_rows + __umul24(i, STRIDE);
                            const float* rj = s_rows + __umul24(j, STRIDE);
                            double t2 = 0.0;
                            for (std::uint32_t k = 0; k < STRIDE; ++k) {
                                const double e = double(ri[k]) - double(rj[k]);
                                t2 += e * e;
                            }
                            const double Tt = sqrt(t2);
                            if (Tt > 0.0) {
                                const double fast = sqrt(dt > 0.0f ? double(dt) : 0.0);
                                const double s2 = double(dt) - double(cG) * nbs * nbs;
                                const double delta = double(rho) * nbs + abs2;
                                const double tlo = (s2 > 0.0 ? sqrt(s2) : 0.0) - delta;
                                const float rel = static_cast<float>(fabs(fast - Tt) / Tt);
                                const float ratio =
                                    static_cast<float>((tlo > 0.0 ? tlo : 0.0) / Tt);
                                worst_rel = max(worst_rel, __float_as_uint(rel));
                                worst_ratio = max(worst_ratio, __float_as_uint(ratio));
                            }
                        }
                    }
                    keep[q] = kp;
                    bal[q] = __ballot(kp);
                    total += static_cast<std::uint32_t>(__popcll(bal[q]));
                }
                std::uint32_t base = 0;
                if (lane == 0 && total) base = atomicAdd(s_len, total);
                base = __shfl(base, 0, 64);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (keep[q]) {
                        const std::uint32_t pos =
                            base + static_cast<std::uint32_t>(__popcll(bal[q] & ((1ull << lane) - 1ull)));
                        const std::uint32_t i = r0 + 4u * kq + static_cast<std::uint32_t>(q);
                        s_li[pos] = static_cast<std::uint16_t>(i | (j << 6));
                    }
                    base += static_cast<std::uint32_t>(__popcll(bal[q]));
                }
            }
            __syncthreads();
            // 3. exact distances of the listed pairs, proposals counted
            const std::uint32_t L = *s_len;
            for (std::uint32_t e = tid; e < L; e += kMfmaThreads) {
                const std::uint32_t ij = s_li[e], i = ij & 63u, j = ij >> 6;
                float d;
                if constexpr (STRIDE <= kJoinRegStride) {
                    float rj[STRIDE];
                    const float4* src =
                        reinterpret_cast<const float4*>(s_rows + __umul24(j, STRIDE));
#pragma unroll
                    for (std::uint32_t q = 0; q < STRIDE / 4; ++q) {
                        const float4 f = src[q];
                        rj[4 * q] = f.x;
                        rj[4 * q + 1] = f.y;
                        rj[4 * q + 2] = f.z;
                        rj[4 * q + 3] = f.w;
                    }
                    d = l2sqr_reg<STRIDE>(rj, s_rows + __umul24(i, STRIDE));
                } else {
                    d = l2sqr_exact(s_rows + __umul24(i, STRIDE), s_rows + __umul24(j, STRIDE),
                                    STRIDE);
                }
                s_ld[e] = d;
                if (d < s_thr[i]) atomicAdd(&s_cnt[i], 1u);
                if (d < s_thr[j]) atomicAdd(&s_cnt[j], 1u);
            }
            __syncthreads();
            // 4. one reservation per receiving member, then the keys
            for (std::uint32_t x = tid; x < ns; x += kMfmaThreads) {
                const std::uint32_t c = s_cnt[x];
                s_base[x] = c ? atomicAdd(&a.fill[s_id[x]], c) : a.cap;
                s_cnt[x] = 0;
            }
            __syncthreads();
            for (std::uint32_t e = tid; e < L; e += kMfmaThreads) {
                const float d = s_ld[e];
                const std::uint32_t ij = s_li[e], i = ij & 63u, j = ij >> 6;
                const bool to_i = d < s_thr[i] && s_base[i] < a.cap;
                const bool to_j = d < s_thr[j] && s_base[j] < a.cap;
                if (!(to_i || to_j)) continue;
                const unsigned long long hb =
                    static_cast<unsigned long long>(fbits(metric_finalize(d, a.cosine))) << 32;
                if (to_i) {
                    const std::uint32_t pos = s_base[i] + atomicAdd(&s_cnt[i], 1u);
                    if (pos < a.cap) a.cand[std::size_t(s_id[i]) * a.cap + pos] = hb | s_id[j];
                }
                if (to_j) {
                    const std::uint32_t pos = s_base[j] + atomicAdd(&s_cnt[j], 1u);
                    if (pos < a.cap) a.cand[std::size_t(s_id[j]) * a.cap + pos] = hb | s_id[i];
                }
            }
            __syncthreads();
            t0 = t1;
        }
    }
    if (a.tally) {
        n_pairs = wave_sum(n_pairs);
        n_exact = wave_sum(n_exact);
        if (lane == 0) {
            atomicAdd(&a.tally[0], n_pairs);
            atomicAdd(&a.tally[1], n_exact);
            atomicAdd(&a.tally[2], n_tiles);
        }
    }
    if (a.audit) {
        worst_rel = wave_max_u32(worst_rel);
        worst_ratio = wave_max_u32(worst_ratio);
        if (lane == 0) {
            atomicMax(&a.audit[0], worst_rel);
            atomicMax(&a.audit[1], worst_ratio);
        }
    }
}

// ---------------------------------------------------------------------------
// merge
// ---------------------------------------------------------------------------

constexpr std::uint32_t kMaxCap = 512;
constexpr std::uint32_t kDefaultCap = 512;

// Ascending bitonic sort of 64*R keys held R per lane (element e = lane*R + r):
// partners closer than R are swapped inside a lane, farther ones across lanes
// by shuffles. Fully unrolled: no LDS, no barriers, no loop overhead.
template <int R>
__device__ __forceinline__ void wave_sort_regs(unsigned long long (&key)[R]) {
    const std::uint32_t lane = __lane_id();
#pragma unroll
    for (std::uint32_t size = 2; size <= 64u * R; size <<= 1) {
#pragma unroll
        for (std::uint32_t half = size >> 1; half > 0; half >>= 1) {
            if (half < static_cast<std::uint32_t>(R)) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int q = r ^ static_cast<int>(half);
                    if (q > r) {
                        const bool up = ((lane * R + r) & size) == 0;
                        const unsigned long long x = key[r], y = key[q];
                        const bool sw = (x > y) == up;
                        key[r] = sw ? y : x;
                        key[q] = sw ? x : y;
                    }
                }
            } else {
                const int lh = static_cast<int>(half / R);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const unsigned long long other = __shfl_xor(key[r], lh, 64);
                    const std::uint32_t e = lane * R + r;
                    const bool up = (e & size) == 0;
                    const bool lo = (e & half) == 0;
                    const unsigned long long mn = other < key[r] ? other : key[r];
                    const unsigned long long mx = other < key[r] ? key[r] : other;
                    key[r] = (lo == up) ? mn : mx;
                }
            }
        }
    }
}

// Sort the nv compacted survivors s_key[0, nv) (nv <= 64R), keep the first copy
// of every key and store the first k in s_top; returns the number of distinct
// survivors.
template <int R>
__device__ __forceinline__ std::uint32_t sort_unique_top(const unsigned long long* s_key,
                                                         std::uint32_t nv, std::uint32_t k,
                                                         unsigned long long* s_top) {
    const std::uint32_t lane = __lane_id();
    unsigned long long key[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const std::uint32_t e = lane * R + r;
        key[r] = e < nv ? s_key[e] : ~0ull;
    }
    wave_sort_regs<R>(key);
    const unsigned long long prev_lane = __shfl_up(key[R - 1], 1, 64);
    bool ok[R];
    std::uint32_t mine = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const unsigned long long prev = r > 0 ? key[r - 1] : prev_lane;
        const bool first = (r == 0 && lane == 0) || prev != key[r];
        ok[r] = key[r] != ~0ull && first;
        mine += ok[r] ? 1u : 0u;
    }
    const std::uint32_t incl = wave_inclusive_scan(mine);
    std::uint32_t rank = incl - mine;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (ok[r]) {
            if (rank < k) s_top[rank] = key[r];
            ++rank;
        }
    }
    return static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63));
}

// The same result as sort_unique_top<1> for nv <= kSmallMerge survivors, by
// counting instead of a 64-key sorting network (about 10 nv VALU operations
// against about 230): each lane's key is compared with every survivor,
// broadcast by readlane, once to find later copies of a key and once more,
// over first copies only, for its rank among the distinct keys.
__device__ __forceinline__ std::uint32_t small_unique_top(const unsigned long long* s_key,
                                                          std::uint32_t nv, std::uint32_t k,
                                                          unsigned long long* s_top) {
    const std::uint32_t lane = __lane_id();
    const unsigned long long key = lane < nv ? s_key[lane] : ~0ull;
    const int lo = static_cast<int>(static_cast<std::uint32_t>(key));
    const int hi = static_cast<int>(static_cast<std::uint32_t>(key >> 32));
    // survivors are never ~0 (the empty key), so lanes >= nv find no copy
    bool dup = false;
    for (std::uint32_t j = 0; j < nv; ++j) {
        const unsigned long long kj =
            (static_cast<unsigned long long>(static_cast<std::uint32_t>(
                 __builtin_amdgcn_readlane(hi, static_cast<int>(j))))
             << 32) |
            static_cast<std::uint32_t>(__builtin_amdgcn_readlane(lo, static_cast<int>(j)));
        dup = dup || (kj == key && j < lane);
    }
    const std::uint64_t dmask = __ballot(dup);
    std::uint32_t rank = 0;
    for (std::uint32_t j = 0; j < nv; ++j) {
        if ((dmask >> j) & 1ull) continue;  // j uniform: a later copy is not counted
        const unsigned long long kj =
            (static_cast<unsigned long long>(static_cast<std::uint32_t>(
                 __builtin_amdgcn_readlane(hi, static_cast<int>(j))))
             << 32) |
            static_cast<std::uint32_t>(__builtin_amdgcn_readlane(lo, static_cast<int>(j)));
        rank += kj < key ? 1u : 0u;
    }
    if (lane < nv && !dup && rank < k) s_top[rank] = key;
    return nv - static_cast<std::uint32_t>(__popcll(dmask));
}

// The same result as sort_unique_top<R> for 64 R / 2 < nv <= 64 R survivors
// (nv <= kMaxCap), without sorting them all: the keys go to registers; a
// sorted sample of 64 (spread over s_key) gives a threshold T with about 72
// keys below it; only those are sorted. When at least k distinct keys lie
// below T they hold the best k; otherwise (or when more than 128 lie below)
// all keys are sorted. The distinct count comes from an open-addressing set
// of the ids (same id implies same key) built over s_key once the best k are
// stored.
template <int R>
__device__ __forceinline__ std::uint32_t select_top_large(unsigned long long* s_key, std::uint32_t nv,
                                          std::uint32_t k, unsigned long long* s_top) {
    static_assert(64 * R <= kMaxCap, "s_key holds the survivors");
    const std::uint32_t lane = __lane_id();
    const std::uint64_t lt = (1ull << lane) - 1ull;
    unsigned long long key[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const std::uint32_t e = static_cast<std::uint32_t>(r) * 64u + lane;
        key[r] = e < nv ? s_key[e] : ~0ull;
    }
    unsigned long long smpl[1] = {s_key[(lane * nv) >> 6]};
    wave_sort_regs<1>(smpl);
    std::uint32_t q = (72u * 65u) / nv;
    q = q > 64u ? 63u : (q > 0u ? q - 1u : 0u);
    const unsigned long long T = __shfl(smpl[0], static_cast<int>(q), 64);
    __syncthreads();  // every read of s_key done
    std::uint32_t P = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const bool pass = key[r] < T;
        const std::uint64_t pm = __ballot(pass);
        if (pass) s_key[P + __popcll(pm & lt)] = key[r];
        P += __popcll(pm);
    }
    __syncthreads();
    std::uint32_t dp = 0;
    if (P <= 64) dp = sort_unique_top<1>(s_key, P, k, s_top);
    else if (P <= 128) dp = sort_unique_top<2>(s_key, P, k, s_top);
    __syncthreads();
    if (dp < k) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const std::uint32_t e = static_cast<std::uint32_t>(r) * 64u + lane;
            if (e < nv) s_key[e] = key[r];
        }
        __syncthreads();
        const std::uint32_t all = sort_unique_top<R>(s_key, nv, k, s_top);
        __syncthreads();
        return all;
    }
    // distinct count: ids into a set of H = 2 pow2ceil(nv) <= 2 kMaxCap slots
    std::uint32_t* tab = reinterpret_cast<std::uint32_t*>(s_key);
    const std::uint32_t lg = 33u - static_cast<std::uint32_t>(__clz(nv - 1));
    const std::uint32_t H = 1u << lg;
    for (std::uint32_t h = lane; h < H; h += 64) tab[h] = kNone;
    __syncthreads();
    std::uint32_t distinct = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        bool fresh = false;
        if (key[r] != ~0ull) {
            const std::uint32_t id = static_cast<std::uint32_t>(key[r]);
            for (std::uint32_t h = (id * 0x9E3779B1u) >> (32u - lg);; h = (h + 1) & (H - 1)) {
                const std::uint32_t prev = atomicCAS(&tab[h], kNone, id);
                if (prev == kNone) {
                    fresh = true;
                    break;
                }
                if (prev == id) break;
            }
        }
        distinct += __popcll(__ballot(fresh));
    }
    __syncthreads();
    return distinct;
}

// Number of entries of the ascending a[0, len) (len <= 64) below x (STRICT) or
// at most x.
template <bool STRICT>
__device__ __forceinline__ std::uint32_t count_before(const float* a, std::uint32_t len, float x) {
    std::uint32_t pos = 0;
#pragma unroll
    for (std::uint32_t step = 64; step > 0; step >>= 1) {
        if (pos + step <= len) {
            const float y = a[pos + step - 1];
            if (STRICT ? (y < x) : (y <= x)) pos += step;
        }
    }
    return pos;
}

// Rows of 64 proposal keys the merge loads before it tests them.
constexpr std::uint32_t kMergeBatch = 4;

// Set of a list's ids for the merge's listed test (at most kMaxK of kListSet
// slots used).
constexpr std::uint32_t kListSet = 2 * kMaxK;
__device__ __forceinline__ std::uint32_t list_slot(std::uint32_t id) {
    return (id * 0x9E3779B1u) >> 25;  // 7 bits: kListSet = 128
}

// One 64-thread group per node (nodes [0, n), candidate ids below id_limit),
// lists of at most kMaxK entries. The round's proposals
// cand[v*cap .. v*cap + min(fill, cap)) are first filtered by try_insert's
// per-candidate tests (knn_graph.cpp:37-55): not the node itself, not already
// listed (an LDS set of the list's ids), and (for a full list) closer than
// the current k-th. The survivors are sorted by (distance, id) in registers
// and the first copy of each key is kept (above 128 survivors only those
// below a sampled threshold are sorted, select_top_large). The best k are
// merged with the list: an entry's new position is its rank plus the number
// of entries of the other sorted list before it -- existing entries first
// among equal distances (the reference's sorted insertion) -- found by binary
// search.
__global__ void __launch_bounds__(64)
    merge_kernel(std::uint32_t n, std::uint32_t id_limit, std::uint32_t k,
                 std::uint32_t* __restrict__ ids,
                 float* __restrict__ dist, std::uint8_t* __restrict__ flags,
                 std::uint32_t* __restrict__ count, const unsigned long long* __restrict__ cand,
                 std::uint32_t* __restrict__ fill, std::uint32_t cap,
                 unsigned long long* __restrict__ partial, const std::uint32_t* __restrict__ halt) {
    __shared__ unsigned long long s_key[kMaxCap];
    if (halt && *halt) return;
    __shared__ unsigned long long s_top[kMaxK];
    __shared__ float s_ld[kMaxK], s_td[kMaxK];
    __shared__ std::uint32_t s_list[kListSet];
    const std::uint32_t lane = threadIdx.x;
    const std::uint64_t lt = (1ull << lane) - 1ull;
    unsigned long long local_upd = 0;
    for (std::uint32_t v = blockIdx.x; v < n; v += gridDim.x) {
        const std::size_t base = std::size_t(v) * k;
        const std::uint32_t c = count[v];
        const std::uint32_t my_id = lane < c ? ids[base + lane] : kNone;
        const float my_d = lane < c ? dist[base + lane] : __builtin_inff();
        const std::uint8_t my_f = lane < c ? flags[base + lane] : 0;
        const float last = (c == k) ? dist[base + k - 1] : __builtin_inff();
        std::uint32_t m = fill[v];
        m = m < cap ? m : cap;
        std::uint32_t t = 0;
        if (m > 0) {
            // the listed ids in an open-addressing set (kListSet slots, load
            // <= 1/2): about one probe per candidate
            for (std::uint32_t h = lane; h < kListSet; h += 64) s_list[h] = kNone;
            __syncthreads();
            if (lane < c) {
                std::uint32_t h = list_slot(my_id);
                while (atomicCAS(&s_list[h], kNone, my_id) != kNone) h = (h + 1) & (kListSet - 1);
            }
            __syncthreads();
            // 1. per-candidate tests, survivors compacted into s_key; the
            // keys are loaded kMergeBatch rows of 64 at a time, so a batch
            // waits for memory once
            std::uint32_t nv = 0;
            const unsigned long long* ck = cand + std::size_t(v) * cap;
            for (std::uint32_t b0 = 0; b0 < m; b0 += 64 * kMergeBatch) {
                unsigned long long kb[kMergeBatch];
#pragma unroll
                for (std::uint32_t r = 0; r < kMergeBatch; ++r) {
                    const std::uint32_t i = b0 + r * 64 + lane;
                    kb[r] = i < m ? ck[i] : ~0ull;
                }
#pragma unroll
                for (std::uint32_t r = 0; r < kMergeBatch; ++r) {
                    if (b0 + r * 64 >= m) break;  // wavefront-uniform
                    const unsigned long long key = kb[r];
                    const std::uint32_t cid = static_cast<std::uint32_t>(key);
                    const float cd = __uint_as_float(static_cast<std::uint32_t>(key >> 32));
                    bool ok = key != ~0ull && cid != v && cid < id_limit && cd < last;
                    if (ok) {
                        for (std::uint32_t h = list_slot(cid);; h = (h + 1) & (kListSet - 1)) {
                            const std::uint32_t x = s_list[h];
                            if (x == cid) {
                                ok = false;
                                break;
                            }
                            if (x == kNone) break;
                        }
                    }
                    const std::uint64_t vm = __ballot(ok);
                    if (ok) s_key[nv + __popcll(vm & lt)] = key;
                    nv += __popcll(vm);
                }
            }
            __syncthreads();
            // 2. sort, keep first copies, rank
            std::uint32_t nvalid = 0;
            if (nv > 0) {
                if (nv <= kSmallMerge) nvalid = small_unique_top(s_key, nv, k, s_top);
                else if (nv <= 64) nvalid = sort_unique_top<1>(s_key, nv, k, s_top);
                else if (nv <= 128) nvalid = sort_unique_top<2>(s_key, nv, k, s_top);
                else if (nv <= 256) nvalid = select_top_large<4>(s_key, nv, k, s_top);
                else nvalid = select_top_large<8>(s_key, nv, k, s_top);
            }
            t = nvalid < k ? nvalid : k;
            __syncthreads();  // s_top complete
            // Updates as the reference counts them (nndescent.cpp:48-49): every
            // candidate try_insert would accept against the list as it was
            // before this merge, even if a closer one of this round evicts it.
            if (lane == 0) local_upd += nvalid;
        }
        if (t > 0) {
            if (lane < c) s_ld[lane] = my_d;
            const unsigned long long tk = lane < t ? s_top[lane] : ~0ull;
            if (lane < t) s_td[lane] = __uint_as_float(static_cast<std::uint32_t>(tk >> 32));
            __syncthreads();
            if (lane < c) {
                const std::uint32_t pos_e = lane + count_before<true>(s_td, t, my_d);
                if (pos_e < k) {
                    ids[base + pos_e] = my_id;
                    dist[base + pos_e] = my_d;
                    flags[base + pos_e] = my_f;
                }
            }
            if (lane < t) {
                const float cd = __uint_as_float(static_cast<std::uint32_t>(tk >> 32));
                const std::uint32_t pos_c = lane + count_before<false>(s_ld, c, cd);
                if (pos_c < k) {
                    ids[base + pos_c] = static_cast<std::uint32_t>(tk);
                    dist[base + pos_c] = cd;
                    flags[base + pos_c] = 1;
                }
            }
        }
        if (lane == 0) {
            if (t > 0) count[v] = (c + t < k) ? c + t : k;
            fill[v] = 0;
        }
        __syncthreads();
    }
    // one partial per block; a single-block pass adds them (a device-wide
    // atomic per block contends on one address across the XCDs)
    if (lane == 0) partial[blockIdx.x] = local_upd;
}

using MergeFn = void (*)(std::uint32_t, std::uint32_t, std::uint32_t, std::uint32_t*, float*,
                         std::uint8_t*, std::uint32_t*, const unsigned long long*, std::uint32_t*,
                         std::uint32_t, unsigned long long*, const std::uint32_t*);

MergeFn pick_merge(std::uint32_t) { return merge_kernel; }

constexpr std::uint32_t kMergeBlocks = 65536;

// One block of kSumThreads adds the per-block partials (loads unrolled so
// that many are in flight per thread). With `halt`, a total at or below
// `stop` sets it (the descent has converged: the next round's kernels return
// at once), and a round that was skipped leaves the total as it was.
constexpr unsigned kSumThreads = 1024;
__global__ void __launch_bounds__(kSumThreads)
    sum_partials_kernel(const unsigned long long* __restrict__ partial, std::uint32_t m,
                        unsigned long long* __restrict__ out, std::uint32_t* __restrict__ halt,
                        double stop) {
    __shared__ unsigned long long s_w[kSumThreads / 64];
    if (halt && *halt) return;
    unsigned long long v = 0;
#pragma unroll 8
    for (std::uint32_t i = threadIdx.x; i < m; i += kSumThreads) v += partial[i];
    v = wave_sum(v);
    if ((threadIdx.x & 63u) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (unsigned w = 0; w < kSumThreads / 64; ++w) t += s_w[w];
        out[0] = t;
        if (halt && double(t) <= stop) *halt = 1u;
    }
}

// Work tallies of one join launch (profiling only): members staged (rows
// read), exact distances evaluated, proposals written.
__global__ void __launch_bounds__(256)
    join_tally_kernel(std::uint32_t n, const std::uint32_t* __restrict__ cnt,
                      const std::uint32_t* __restrict__ fill, std::uint32_t cap,
                      unsigned long long* __restrict__ out) {
    unsigned long long rows = 0, dists = 0, props = 0;
    for (std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += gridDim.x * blockDim.x) {
        const std::uint32_t na = cnt[v] & 0xFFFFu, ns = cnt[v] >> 16;
        if (na == 0) continue;
        rows += ns;
        // new x new pairs once, new x old pairs
        dists += static_cast<unsigned long long>(na) * ns - static_cast<unsigned long long>(na) * (na + 1) / 2;
        props += fill[v] < cap ? fill[v] : cap;
    }
    __shared__ unsigned long long s_t[3][4];
    rows = wave_sum(rows);
    dists = wave_sum(dists);
    props = wave_sum(props);
    if ((threadIdx.x & 63u) == 0) {
        s_t[0][threadIdx.x >> 6] = rows;
        s_t[1][threadIdx.x >> 6] = dists;
        s_t[2][threadIdx.x >> 6] = props;
    }
    __syncthreads();
    if (threadIdx.x < 3)
        atomicAdd(&out[threadIdx.x], s_t[threadIdx.x][0] + s_t[threadIdx.x][1] +
                                         s_t[threadIdx.x][2] + s_t[threadIdx.x][3]);
}

// ---------------------------------------------------------------------------
// exact brute force (nndescent.cpp:242-263), one thread per node in j order
// ---------------------------------------------------------------------------

__global__ void brute_force_kernel(const float* __restrict__ pts, std::uint32_t n,
                                   std::uint32_t stride, std::uint32_t k,
                                   std::uint32_t* __restrict__ ids, float* __restrict__ dist,
                                   std::uint8_t* __restrict__ flags,
                                   std::uint32_t* __restrict__ count) {
    for (std::uint32_t node = blockIdx.x * blockDim.x + threadIdx.x; node < n;
         node += gridDim.x * blockDim.x) {
        const std::size_t base = std::size_t(node) * k;
        std::uint32_t* I = ids + base;
        float* D = dist + base;
        std::uint8_t* F = flags + base;
        std::uint32_t cnt = 0;
        for (std::uint32_t s = 0; s < k; ++s) {
            I[s] = kNone;
            D[s] = __builtin_inff();
            F[s] = 0;
        }
        const float* v = pts + std::size_t(node) * stride;
        for (std::uint32_t j = 0; j < n; ++j) {
            if (j == node) continue;
            const float r = l2sqr_exact(v, pts + std::size_t(j) * stride, stride);
            const float t = D[k - 1];
            if (cnt == k && !(r < t * t)) continue;
            const float d = sqrt_exact(r);
            if (cnt == k && !(d < D[k - 1])) continue;
            std::uint32_t pos;
            if (cnt < k) {
                pos = cnt++;
            } else {
                pos = k - 1;
            }
            while (pos > 0 && D[pos - 1] > d) {
                D[pos] = D[pos - 1];
                I[pos] = I[pos - 1];
                --pos;
            }
            D[pos] = d;
            I[pos] = j;
        }
        count[node] = cnt;
    }
}

using JoinFn = void (*)(JoinArgs);

JoinFn pick_join(std::uint32_t stride) {
    switch (stride) {
        case 8: return join_kernel<8>;
        case 16: return join_kernel<16>;
        case 24: return join_kernel<24>;
        case 32: return join_kernel<32>;
        case 48: return join_kernel<48>;
        case 64: return join_kernel<64>;
        case 128: return join_kernel<128>;
        default: return join_kernel<0>;
    }
}

// Member-stationary exact join for the strides it is built for, else nullptr.
template <typename Mask>
JoinFn pick_join_reg_t(std::uint32_t stride) {
    switch (stride) {
        case 8: return join_reg_kernel<8, Mask>;
        case 16: return join_reg_kernel<16, Mask>;
        case 24: return join_reg_kernel<24, Mask>;
        case 32: return join_reg_kernel<32, Mask>;
        case 48: return join_reg_kernel<48, Mask>;
        case 64: return join_reg_kernel<64, Mask>;
        default: return nullptr;
    }
}

// Entries per wavefront of the member-stationary join's list of passing
// pairs. MI355X, 1M points, ms per join launch without a list / with 48
// entries: d = 8 1.31 / 1.29, 16 1.61 / 1.58, 24 2.34 / 2.26, 32 3.44 /
// 3.34, 64 9.08 / 8.78; 64 and 96 entries cost the stride-24 join a
// block per CU for nothing (2.33), 160 two (2.38). At stride 48 a list of 32
// keeps its eight blocks per CU (4.25 / 4.14).
std::uint32_t join_list_cap(std::uint32_t stride) { return stride == 48 ? 32u : 48u; }

// A lane of the member-stationary join visits at most A = 2 smp pairs: a
// 32-bit overflow mask serves samples of up to 16.
JoinFn pick_join_reg(std::uint32_t stride, std::uint32_t smp) {
    return 2u * smp <= 32u ? pick_join_reg_t<std::uint32_t>(stride)
                           : pick_join_reg_t<std::uint64_t>(stride);
}

template <bool BF16, bool AUDIT>
JoinFn pick_join_mfma_t(std::uint32_t stride) {
    switch (stride) {
        case 8: return join_mfma_kernel<8, BF16, AUDIT>;
        case 16: return join_mfma_kernel<16, BF16, AUDIT>;
        case 24: return join_mfma_kernel<24, BF16, AUDIT>;
        case 32: return join_mfma_kernel<32, BF16, AUDIT>;
        case 48: return join_mfma_kernel<48, BF16, AUDIT>;
        case 64: return join_mfma_kernel<64, BF16, AUDIT>;
        case 96: return join_mfma_kernel<96, BF16, AUDIT>;
        case 128: return join_mfma_kernel<128, BF16, AUDIT>;
        default: return nullptr;
    }
}

// Local-join distance modes (coresg_hip_build_params::gram).
enum : std::uint32_t { kGramAuto = 0, kGramExact = 1, kGramF16 = 2, kGramBF16 = 3 };

// The library's choice: the exact member-stationary join for rows of up to
// 64 floats, the FP16 matrix-core filter join for wider rows. In NN-Descent a
// third (d = 24) to almost half (d = 128) of the visited pairs pass a
// threshold test and need their exact distance anyway; at d = 24 an exact
// distance costs about as much as the filter's per-pair bound and list work
// over 16 x 16 tiles (MI355X, n = 1M: 2.37 ms exact vs 6.99 ms FP16 filter per
// join launch), at d = 128 it costs six times more (n = 2M: 64.5 ms exact vs
// 45.8 ms FP16 and 47.8 ms BF16 filter).
std::uint32_t resolve_gram(std::uint32_t mode, std::uint32_t stride) {
    if (mode == kGramAuto) return stride >= 96 ? kGramF16 : kGramExact;
    if (mode > kGramBF16) throw std::runtime_error("nndescent: gram must be 0..3");
    return mode;
}

struct DescentResult {
    std::uint32_t rounds = 0;
    std::uint64_t updates = 0;
};

DescentResult run_descent(Graph* gp, const float* pts, std::uint32_t n, std::uint32_t stride,
                          std::uint32_t k, std::uint32_t* ids, float* dist, std::uint8_t* flags,
                          std::uint32_t* count, std::uint32_t smp, std::uint32_t iters,
                          double delta, std::uint64_t seed, bool random_fill,
                          std::uint32_t cap, std::uint32_t gram, hipStream_t s,
                          bool cosine = false, bool graph_rounds = false) {
    DescentResult res;
    if (n == 0) return res;
    if (k > kMaxK) throw std::runtime_error("nndescent: k above 64 is not supported on device");
    if (smp == 0) smp = 1;
    if (smp > kMaxSample) smp = kMaxSample;
    const unsigned wave_blocks = capped_blocks(std::size_t(n) * 64, 256, 16384);
    if (random_fill && n >= 2) {
        std::unique_ptr<ScopedKernelTimer> t;
        if (gp) t.reset(new ScopedKernelTimer(*gp, "nnd_init"));
        const bool inst = gp && !gp->inst_off.empty();
        init_random_kernel<<<wave_blocks, 256, 0, s>>>(pts, n, stride, k, seed, ids, dist, flags,
                                                       count, inst ? gp->inst_of.p : nullptr,
                                                       inst ? gp->inst_off_d.p : nullptr, cosine);
        CSG_LAUNCH_CHECK();
    }
    if (iters == 0) return res;
    DescentBuffers local;
    DescentBuffers& b = gp ? gp->nnd : local;
    b.new_list.reserve(std::size_t(n) * smp);
    b.new_cnt.reserve(n);
    b.old_list.reserve(std::size_t(n) * 3 * smp);
    b.old_cnt.reserve(n);
    b.rev_new.reserve(std::size_t(n) * smp);
    b.rev_new_cnt.reserve(n);
    b.rev_old.reserve(std::size_t(n) * smp);
    b.rev_old_cnt.reserve(n);
    b.cand.reserve(std::size_t(n) * cap);
    b.fill.reserve(n);
    b.thr_node.reserve(n);
    b.updates.reserve(1);
    b.partial.reserve(kMergeBlocks);
    b.tally.reserve(4);
    fill_u32(b.fill.p, n, 0u, s);

    // reverse lists by target buckets when the bucket count allows
    const std::uint32_t nb = (n + kRevTargets - 1) >> kRevShift;
    const bool rev_buckets = nb <= kRevMaxBuckets;
    if (rev_buckets) {
        b.rev_pairs.reserve(std::size_t(n) * k);
        b.rev_hist.reserve(std::size_t(nb) * kRevBlocks);
        b.rev_tmp.reserve(scan_tmp_words(std::size_t(nb) * kRevBlocks));
        b.rev_total.reserve(1);
    }

    const std::uint32_t S = 6u * smp, A = 2u * smp;
    b.gids.reserve(std::size_t(n) * S);
    b.gthr.reserve(std::size_t(n) * S);
    b.gcnt.reserve(n);
    gram = resolve_gram(gram, stride);
    // the matrix-core filter join is built for the common strides; any
    // other stride takes the exact join (which proposes the same pairs)
    const bool audit = gp && gp->profiling && gp->audit;
    const JoinFn mfma_fn =
        gram == kGramF16    ? (audit ? pick_join_mfma_t<false, true>(stride)
                                     : pick_join_mfma_t<false, false>(stride))
        : gram == kGramBF16 ? (audit ? pick_join_mfma_t<true, true>(stride)
                                     : pick_join_mfma_t<true, false>(stride))
                            : nullptr;
    if (!mfma_fn) gram = kGramExact;
    const bool join_reg = gram == kGramExact && S <= kJoinThreads && pick_join_reg(stride, smp);
    const JoinFn join_fn = mfma_fn    ? mfma_fn
                           : join_reg ? pick_join_reg(stride, smp)
                                      : pick_join(stride);
    // the filter join's first pass stages at most row_cap rows per node (0:
    // all S); its second pass, over the nodes set aside, stages S
    const std::uint32_t row_cap = mfma_fn && mfma_row_cap(stride) < S ? mfma_row_cap(stride) : 0u;
    auto mfma_lds = [&](std::uint32_t rows) {
        return (8 * std::size_t(S) + std::size_t(rows) * stride + kMfmaList) * sizeof(float) +
               kMfmaList * sizeof(std::uint16_t) + sizeof(std::uint32_t);
    };
    const std::size_t lds_full = mfma_fn ? mfma_lds(S) : 0;
    const std::uint32_t list_cap = join_reg ? join_list_cap(stride) : 0u;
    const std::size_t lds =
        mfma_fn ? mfma_lds(row_cap ? row_cap : S)
        : join_reg
            ? (4 * std::size_t(S) + std::size_t(S) * stride) * sizeof(float) +
                  2 * std::size_t(list_cap) * sizeof(uint2)
            : (4 * std::size_t(S) + std::size_t(S) * stride + std::size_t(A) * S) * sizeof(float);
    if (mfma_fn && S > 1024u) throw std::runtime_error("nndescent: sample too large for the filter join");
    const unsigned join_threads = mfma_fn ? mfma_join_threads(stride) : kJoinThreads;
    const unsigned join_grid = static_cast<unsigned>(std::min<std::size_t>(n, 65536));
    if (std::max(lds, lds_full) > 150 * 1024)
        throw std::runtime_error("nndescent: candidate tile exceeds LDS");
    if (std::max(lds, lds_full) > 64 * 1024)
        CSG_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(join_fn),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(std::max(lds, lds_full))));
    if (row_cap) {
        b.defer.reserve(n);
        b.defer_count.reserve(1);
    }
    // Convergence (nndescent.cpp:172, 204) is decided on the device: the round
    // whose update count is at or below `stop` sets `halt`, and every kernel of
    // the rounds after it returns at once. The host enqueues round r + 1
    // before it reads round r's counts (through pinned memory and an event),
    // so the device never waits for the host between rounds.
    b.halt.reserve(1);
    fill_u32(b.halt.p, 1, 0u, s);
    b.readback.reserve(iters);
    b.round_done.create();
    JoinArgs ja{pts, n, stride, smp, b.gids.p, b.gthr.p, b.gcnt.p, b.cand.p, b.fill.p, cap,
                nullptr, nullptr, b.halt.p, cosine};
    ja.row_cap = row_cap;
    ja.list_cap = list_cap;
    ja.defer = row_cap ? b.defer.p : nullptr;
    ja.defer_count = row_cap ? b.defer_count.p : nullptr;
    const unsigned merge_blocks = static_cast<unsigned>(std::min<std::size_t>(n, kMergeBlocks));
    const bool tallies = gp && gp->profiling;
    const bool gram_audit = tallies && gram != kGramExact;
    if (gram_audit) {
        b.gram_tally.reserve(3);
        b.gram_audit.reserve(2);
        ja.tally = b.gram_tally.p;
        ja.audit = gp->audit ? b.gram_audit.p : nullptr;
    }
    const double stop = delta * double(k) * double(n);
    // Reads round r's counts once its copies have landed; true: it converged.
    auto consume = [&](std::uint32_t r, bool wait) -> bool {
        if (wait) CSG_CHECK(hipEventSynchronize(b.round_done.e[r & 1u]));
        const RoundReadback& rb = b.readback.p[r];
        if (tallies) gp->add_work("nnd_join", {rb.tally[0], rb.tally[1], rb.tally[2], n});
        if (gram_audit) {
            gp->add_work("nnd_gram", {rb.gram_tally[0], rb.gram_tally[1], 0, 0});
            // matrix-core work: 16 x 16 tiles, 32-coordinate steps per tile
            gp->add_work("nnd_mfma", {rb.gram_tally[2]});
            gp->max_work("nnd_mfma", 1, (stride + 31u) / 32u);
            gp->max_work("nnd_gram", 2, rb.gram_audit[0]);
            gp->max_work("nnd_gram", 3, rb.gram_audit[1]);
        }
        res.updates += rb.upd;
        res.rounds = r + 1;
        return double(rb.upd) <= stop;
    };
    // One round's launches and the copies of its counts into `readback`.
    auto enqueue_round = [&](std::uint32_t round) {
        {
            std::unique_ptr<ScopedKernelTimer> t;
            if (gp) t.reset(new ScopedKernelTimer(*gp, "nnd_sample"));
            if (rev_buckets) {
                const std::size_t hl = std::size_t(nb) * sizeof(std::uint32_t);
                const auto sample_fn = k <= 16 ? sample_kernel<16, true>
                                       : k <= 32 ? sample_kernel<32, true>
                                                 : sample_kernel<64, true>;
                sample_fn<<<kRevBlocks, 256, hl, s>>>(
                    n, k, ids, flags, count, smp, b.new_list.p, b.new_cnt.p, b.old_list.p,
                    b.old_cnt.p, b.rev_new.p, b.rev_new_cnt.p, b.rev_old.p, b.rev_old_cnt.p,
                    b.rev_hist.p, nb, dist, b.thr_node.p, b.halt.p, cosine);
                CSG_LAUNCH_CHECK();
                exclusive_scan_u32(b.rev_hist.p, std::size_t(nb) * kRevBlocks, b.rev_total.p,
                                   b.rev_tmp.p, s);
                const auto scatter_fn = k <= 16 ? rev_scatter_kernel<16>
                                        : k <= 32 ? rev_scatter_kernel<32>
                                                  : rev_scatter_kernel<64>;
                scatter_fn<<<kRevBlocks, 256, hl, s>>>(n, smp, b.new_list.p, b.new_cnt.p,
                                                       b.old_list.p, b.old_cnt.p, b.rev_hist.p,
                                                       nb, b.rev_pairs.p, b.halt.p);
                CSG_LAUNCH_CHECK();
                rev_bucket_kernel<<<nb, kRevThreads, 0, s>>>(
                    n, smp, b.rev_pairs.p, b.rev_hist.p, kRevBlocks, b.rev_total.p, b.rev_new.p,
                    b.rev_new_cnt.p, b.rev_old.p, b.rev_old_cnt.p, b.halt.p);
                CSG_LAUNCH_CHECK();
            } else {
                fill_u32(b.rev_new_cnt.p, n, 0u, s);
                fill_u32(b.rev_old_cnt.p, n, 0u, s);
                const auto sample_fn = k <= 16 ? sample_kernel<16, false>
                                       : k <= 32 ? sample_kernel<32, false>
                                                 : sample_kernel<64, false>;
                sample_fn<<<wave_blocks, 256, 0, s>>>(
                    n, k, ids, flags, count, smp, b.new_list.p, b.new_cnt.p, b.old_list.p,
                    b.old_cnt.p, b.rev_new.p, b.rev_new_cnt.p, b.rev_old.p, b.rev_old_cnt.p,
                    nullptr, 0, dist, b.thr_node.p, b.halt.p, cosine);
                CSG_LAUNCH_CHECK();
            }
        }
        {
            std::unique_ptr<ScopedKernelTimer> t;
            if (gp) t.reset(new ScopedKernelTimer(*gp, "nnd_gather"));
            gather_kernel<32><<<wave_blocks, 256, 0, s>>>(
                n, smp, b.new_list.p, b.new_cnt.p, b.old_list.p, b.old_cnt.p,
                b.rev_new.p, b.rev_new_cnt.p, b.rev_old.p, b.rev_old_cnt.p, b.gids.p, b.gthr.p,
                b.gcnt.p, b.thr_node.p, b.halt.p);
            CSG_LAUNCH_CHECK();
        }
        if (gram_audit) {
            fill_u64(reinterpret_cast<std::uint64_t*>(b.gram_tally.p), 3, 0ull, s);
            fill_u32(b.gram_audit.p, 2, 0u, s);
        }
        {
            std::unique_ptr<ScopedKernelTimer> t;
            if (gp) t.reset(new ScopedKernelTimer(*gp, "nnd_join"));
            if (row_cap) fill_u32(b.defer_count.p, 1, 0u, s);
            hipLaunchKernelGGL(join_fn, dim3(join_grid), dim3(join_threads), lds, s, ja);
            CSG_LAUNCH_CHECK();
            if (row_cap) {  // the nodes set aside, all S rows staged
                JoinArgs jd = ja;
                jd.deferred_pass = true;
                hipLaunchKernelGGL(join_fn, dim3(kDeferBlocks), dim3(join_threads), lds_full, s,
                                   jd);
                CSG_LAUNCH_CHECK();
            }
        }
        if (tallies) {
            fill_u64(reinterpret_cast<std::uint64_t*>(b.tally.p), 3, 0ull, s);
            join_tally_kernel<<<capped_blocks(n, 256, 2048), 256, 0, s>>>(n, b.gcnt.p, b.fill.p,
                                                                          cap, b.tally.p);
            CSG_LAUNCH_CHECK();
        }
        {
            std::unique_ptr<ScopedKernelTimer> t;
            if (gp) t.reset(new ScopedKernelTimer(*gp, "nnd_merge"));
            pick_merge(k)<<<merge_blocks, 64, 0, s>>>(n, n, k, ids, dist, flags, count, b.cand.p,
                                                     b.fill.p, cap, b.partial.p, b.halt.p);
            CSG_LAUNCH_CHECK();
        }
        sum_partials_kernel<<<1, kSumThreads, 0, s>>>(b.partial.p, merge_blocks, b.updates.p,
                                                      b.halt.p, stop);
        CSG_LAUNCH_CHECK();
        RoundReadback* rb = b.readback.p + round;
        CSG_CHECK(hipMemcpyAsync(&rb->upd, b.updates.p, 8, hipMemcpyDeviceToHost, s));
        if (tallies)
            CSG_CHECK(hipMemcpyAsync(rb->tally, b.tally.p, 24, hipMemcpyDeviceToHost, s));
        if (gram_audit) {
            CSG_CHECK(hipMemcpyAsync(rb->gram_tally, b.gram_tally.p, 24, hipMemcpyDeviceToHost, s));
            CSG_CHECK(hipMemcpyAsync(rb->gram_audit, b.gram_audit.p, 8, hipMemcpyDeviceToHost, s));
        }
    };
    const bool plain = gp && gp->plain;
    // Every round of the descent as one captured hipGraph, replayed by the
    // builds of a graph after its first (graph_rounds: a build, whose sizes
    // repeat; inserts change n each time):
    // the rounds after convergence return at once (the device's halt flag),
    // and the counts are read after the replay. Profiled builds keep the
    // eager launches (their per-kernel events), the plain form its host
    // synchronisation after every launch.
    bool replay = false;
    if (graph_rounds && gp && !plain && !tallies) {
        std::vector<std::uint64_t> key = {
            reinterpret_cast<std::uint64_t>(pts), n, stride, k,
            reinterpret_cast<std::uint64_t>(ids), reinterpret_cast<std::uint64_t>(dist),
            reinterpret_cast<std::uint64_t>(flags), reinterpret_cast<std::uint64_t>(count), smp,
            iters, cap, gram, cosine ? 1u : 0u, reinterpret_cast<std::uint64_t>(join_fn),
            lds, lds_full, row_cap, join_threads, join_grid, merge_blocks, rev_buckets ? nb : 0u,
            wave_blocks, list_cap};
        std::uint64_t stop_bits;
        std::memcpy(&stop_bits, &stop, 8);
        key.push_back(stop_bits);
        for (const void* q : {static_cast<const void*>(b.new_list.p), static_cast<const void*>(b.new_cnt.p),
                              static_cast<const void*>(b.old_list.p), static_cast<const void*>(b.old_cnt.p),
                              static_cast<const void*>(b.rev_new.p), static_cast<const void*>(b.rev_new_cnt.p),
                              static_cast<const void*>(b.rev_old.p), static_cast<const void*>(b.rev_old_cnt.p),
                              static_cast<const void*>(b.cand.p), static_cast<const void*>(b.updates.p),
                              static_cast<const void*>(b.partial.p), static_cast<const void*>(b.tally.p),
                              static_cast<const void*>(b.fill.p), static_cast<const void*>(b.gids.p),
                              static_cast<const void*>(b.gcnt.p), static_cast<const void*>(b.thr_node.p),
                              static_cast<const void*>(b.rev_pairs.p), static_cast<const void*>(b.rev_hist.p),
                              static_cast<const void*>(b.rev_tmp.p), static_cast<const void*>(b.rev_total.p),
                              static_cast<const void*>(b.gthr.p), static_cast<const void*>(b.halt.p),
                              static_cast<const void*>(b.readback.p),
                              static_cast<const void*>(ja.defer),
                              static_cast<const void*>(ja.defer_count)})
            key.push_back(reinterpret_cast<std::uint64_t>(q));
        // A graph pays for its capture only when it is replayed: the first
        // build of a key runs eagerly and remembers the key, the second
        // captures, every later one replays.
        if (!gp->descent_exec || gp->descent_key != key) {
            if (gp->descent_exec) {
                std::lock_guard<std::mutex> lk(device_api_mutex());
                CSG_CHECK(hipGraphExecDestroy(gp->descent_exec));
            }
            gp->descent_exec = nullptr;
            gp->descent_key.clear();
        }
        if (!gp->descent_exec && gp->descent_key_seen == key) {
            hipGraph_t graph = nullptr;
            {
                CaptureGuard guard(s);
                for (std::uint32_t round = 0; round < iters; ++round) enqueue_round(round);
                graph = guard.end();
            }
            hipError_t e;
            {
                std::lock_guard<std::mutex> lk(device_api_mutex());
                e = hipGraphInstantiate(&gp->descent_exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
            }
            CSG_CHECK(e);
            gp->descent_key = key;
        }
        replay = gp->descent_exec != nullptr;
        gp->descent_key_seen = std::move(key);
    }
    if (replay) {
        CSG_CHECK(hipGraphLaunch(gp->descent_exec, s));
        CSG_CHECK(hipStreamSynchronize(s));
        for (std::uint32_t round = 0; round < iters; ++round)
            if (consume(round, false)) break;
        return res;
    }
    std::uint32_t last = 0;
    bool converged = false;
    for (std::uint32_t round = 0; round < iters; ++round) {
        last = round;
        if (gp && gp->yield_hook) gp->yield_hook();
        enqueue_round(round);
        CSG_CHECK(hipEventRecord(b.round_done.e[round & 1u], s));
        if (plain) {
            // the plain form: the host reads the round's count before it
            // enqueues the next round
            if (consume(round, true)) {
                converged = true;
                break;
            }
            continue;
        }
        if (round > 0 && consume(round - 1, true)) {
            converged = true;  // this round's kernels returned at once
            break;
        }
    }
    if (!converged && !plain) consume(last, true);
    // the skipped round's copies land in `readback` too: wait for them
    CSG_CHECK(hipEventSynchronize(b.round_done.e[last & 1u]));
    return res;
}

// ---------------------------------------------------------------------------
// batch insertion: exact k-NN of a batch of new points against every point,
// and the new points proposed to the existing lists they belong in
// ---------------------------------------------------------------------------

constexpr unsigned kInsThreads = 256;  // 4 waves; lane = new point, wave = column phase
constexpr unsigned kInsRows = 64;
constexpr unsigned kInsTile = 128;     // columns staged in LDS per step

struct InsertArgs {
    const float* pts;
    std::uint32_t n0, n1, k;  // existing points [0, n0), new points [n0, n1)
    std::uint32_t* ids;
    float* dist;
    std::uint8_t* flags;
    std::uint32_t* count;
    unsigned long long* cand;  // proposal buffers of the existing points
    std::uint32_t* fill;
    std::uint32_t cap;
    // column chunks: block b scans query block b / chunks against columns
    // [c * chunk_cols, (c + 1) * chunk_cols) with c = b % chunks; with more
    // than one chunk every block writes its KP best as keys to `partial`
    // (per new point, per chunk) instead of the final list
    std::uint32_t chunks, chunk_cols;
    unsigned long long* partial;
    unsigned long long* tally;  // Gram-filter scan, profiling: {pairs, exact}
    // the scanned columns start at col_base; a new point's key lists in
    // `partial` are `slots` (>= chunks) lists apart
    std::uint32_t col_base, slots;
};

// (d, id) lexicographic order on metric distances: the reference's brute
// force keeps the lower id among equal distances (nndescent.cpp:242-263).
__device__ __forceinline__ bool key_less(float d, std::uint32_t i, float e, std::uint32_t j) {
    return d < e || (d == e && i < j);
}

// Insert (d, id) into a sorted register list (unrolled: no dynamic indexing).
template <std::uint32_t KP>
__device__ __forceinline__ void list_insert(float (&td)[KP], std::uint32_t (&ti)[KP], float d,
                                            std::uint32_t id) {
#pragma unroll
    for (int i = int(KP) - 1; i > 0; --i) {
        const bool shift = key_less(d, id, td[i - 1], ti[i - 1]);
        const bool here = !shift && key_less(d, id, td[i], ti[i]);
        const float nd = shift ? td[i - 1] : (here ? d : td[i]);
        const std::uint32_t ni = shift ? ti[i - 1] : (here ? id : ti[i]);
        td[i] = nd;
        ti[i] = ni;
    }
    if (key_less(d, id, td[0], ti[0])) {
        td[0] = d;
        ti[0] = id;
    }
}

// After a graph search (its KP keys in the last of a new point's slots,
// sorted) the final list is the best k over all slots, so a column whose key
// is not below the search's k-th key cannot enter it: that key (bd, bi) bounds
// a scan's list from the start, and its square (loosened like the list's)
// the squared prefilter thr2. Without a search slot or a k-th key: no bound.
template <std::uint32_t KP>
__device__ __forceinline__ void search_bound(const InsertArgs& a, std::uint32_t p, bool has,
                                             float& bd, std::uint32_t& bi, float& thr2) {
    bd = __builtin_inff();
    bi = kNone;
    if (!has || !a.partial || a.slots <= a.chunks) return;
    const unsigned long long bk =
        a.partial[(std::size_t(p - a.n0) * a.slots + a.slots - 1) * KP + a.k - 1];
    if (bk == ~0ull) return;
    bd = __uint_as_float(static_cast<std::uint32_t>(bk >> 32));
    bi = static_cast<std::uint32_t>(bk);
    thr2 = bd * bd * (1.0f + 0x1p-20f);
}

// One block per 64 new points. Every step stages kInsTile points in LDS; a
// lane holds its new point's row in registers and computes the exact squared
// distance (reference summation order) to every 4th staged point, keeping its
// best KP in a sorted register list. The same distance offers the new point
// to the staged point's list when that one is an existing point and the
// distance beats its current k-th. The four partial lists of a new point are
// merged at the end (waves 2,3 -> 0,1 -> 0) and its first k entries written.
// Rows of up to 24 floats are held to three waves per SIMD (168 VGPRs and 12
// bytes of spilled state per lane at d = 24) instead of the two that 215
// VGPRs allow: a 50k-point batch's scan of its own pairs at d = 24 took 9.4
// -> 7.2 ms on MI355X; wider rows spill far more.
template <std::uint32_t STRIDE, std::uint32_t KP>
__global__ void __launch_bounds__(kInsThreads)
    __attribute__((amdgpu_waves_per_eu(STRIDE <= 24 ? 3 : 1))) insert_scan_kernel(InsertArgs a) {
    extern __shared__ float lds[];
    float* s_tile = lds;                                 // kInsTile * STRIDE
    float* s_thr = lds + kInsTile * STRIDE;              // kInsTile
    float* s_ld = lds;                                   // merge phase: 2 * kInsRows * KP
    std::uint32_t* s_li = reinterpret_cast<std::uint32_t*>(lds + 2 * kInsRows * KP);
    const std::uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const std::uint32_t chunk = blockIdx.x % a.chunks;
    const std::uint32_t col0 = a.col_base + chunk * a.chunk_cols;
    const std::uint32_t col1 = min(a.n1, col0 + a.chunk_cols);
    const std::uint32_t qstep = (gridDim.x / a.chunks) * kInsRows;
    for (std::uint32_t rb = a.n0 + (blockIdx.x / a.chunks) * kInsRows; rb < a.n1; rb += qstep) {
        const std::uint32_t p = rb + lane;
        const bool has = p < a.n1;
        float row[STRIDE];
#pragma unroll
        for (std::uint32_t f = 0; f < STRIDE / 4; ++f) {
            const float4 t = has ? reinterpret_cast<const float4*>(a.pts + std::size_t(p) * STRIDE)[f]
                                 : make_float4(0.f, 0.f, 0.f, 0.f);
            row[4 * f] = t.x;
            row[4 * f + 1] = t.y;
            row[4 * f + 2] = t.z;
            row[4 * f + 3] = t.w;
        }
        float td[KP];  // metric distances, sorted by (d, id)
        std::uint32_t ti[KP];
#pragma unroll
        for (std::uint32_t i = 0; i < KP; ++i) {
            td[i] = __builtin_inff();
            ti[i] = kNone;
        }
        // squared prefilter for td[KP-1], loosened by 2^-20 so that no pair
        // whose rounded sqrt ties or beats it is skipped
        float thr2 = __builtin_inff();
        float bd;
        std::uint32_t bi;
        search_bound<KP>(a, p, has, bd, bi, thr2);
        for (std::uint32_t c0 = col0; c0 < col1; c0 += kInsTile) {
            __syncthreads();  // the previous tile is consumed
            const std::uint32_t cend = min(kInsTile, col1 - c0);
            for (std::uint32_t e = threadIdx.x; e < cend * (STRIDE / 4); e += kInsThreads) {
                const std::uint32_t c = e / (STRIDE / 4), f = e - c * (STRIDE / 4);
                reinterpret_cast<float4*>(s_tile)[e] =
                    reinterpret_cast<const float4*>(a.pts + std::size_t(c0 + c) * STRIDE)[f];
            }
            for (std::uint32_t c = threadIdx.x; c < cend; c += kInsThreads) {
                const std::uint32_t j = c0 + c;
                float t = -1.0f;  // new points take no proposals here
                if (j < a.n0) {
                    t = __builtin_inff();
                    if (a.count[j] >= a.k) {
                        const float dk = a.dist[std::size_t(j) * a.k + a.k - 1];
                        t = dk * dk * (1.0f + 0x1p-20f);  // the merge makes the exact test
                    }
                }
                s_thr[c] = t;
            }
            __syncthreads();
            for (std::uint32_t c = w; c < cend; c += 4) {
                const std::uint32_t j = c0 + c;
                const float d2 = l2sqr_reg<STRIDE>(row, s_tile + c * STRIDE);
                if (!has || j == p) continue;
                if (d2 <= thr2) {
                    const float d = sqrt_exact(d2);
                    if (key_less(d, j, td[KP - 1], ti[KP - 1]) && key_less(d, j, bd, bi)) {
                        list_insert<KP>(td, ti, d, j);
                        thr2 = fminf(thr2, td[KP - 1] * td[KP - 1] * (1.0f + 0x1p-20f));
                    }
                }
                if (d2 < s_thr[c]) {
                    const std::uint32_t pos = atomicAdd(&a.fill[j], 1u);
                    if (pos < a.cap)
                        a.cand[std::size_t(j) * a.cap + pos] =
                            (static_cast<unsigned long long>(fbits(sqrt_exact(d2))) << 32) | p;
                }
            }
        }
        // merge the four partial lists of each new point
        __syncthreads();
        if (w >= 2) {
#pragma unroll
            for (std::uint32_t i = 0; i < KP; ++i) {
                s_ld[((w - 2) * kInsRows + lane) * KP + i] = td[i];
                s_li[((w - 2) * kInsRows + lane) * KP + i] = ti[i];
            }
        }
        __syncthreads();
        if (w < 2) {
            for (std::uint32_t i = 0; i < KP; ++i) {
                const float d = s_ld[(w * kInsRows + lane) * KP + i];
                const std::uint32_t id = s_li[(w * kInsRows + lane) * KP + i];
                if (id == kNone || !key_less(d, id, td[KP - 1], ti[KP - 1])) break;
                list_insert<KP>(td, ti, d, id);
            }
        }
        __syncthreads();
        if (w == 1) {
#pragma unroll
            for (std::uint32_t i = 0; i < KP; ++i) {
                s_ld[lane * KP + i] = td[i];
                s_li[lane * KP + i] = ti[i];
            }
        }
        __syncthreads();
        if (w == 0) {
            for (std::uint32_t i = 0; i < KP; ++i) {
                const float d = s_ld[lane * KP + i];
                const std::uint32_t id = s_li[lane * KP + i];
                if (id == kNone || !key_less(d, id, td[KP - 1], ti[KP - 1])) break;
                list_insert<KP>(td, ti, d, id);
            }
            if (has && a.partial) {
                unsigned long long* out =
                    a.partial + (std::size_t(p - a.n0) * a.slots + chunk) * KP;
#pragma unroll
                for (std::uint32_t i = 0; i < KP; ++i)
                    out[i] = ti[i] == kNone
                                 ? ~0ull
                                 : (static_cast<unsigned long long>(fbits(td[i])) << 32) | ti[i];
            } else if (has) {
                std::uint32_t c = 0;
#pragma unroll
                for (std::uint32_t i = 0; i < KP; ++i) {
                    if (i < a.k && ti[i] != kNone) {
                        const std::size_t o = std::size_t(p) * a.k + i;
                        a.ids[o] = ti[i];
                        a.dist[o] = td[i];
                        a.flags[o] = 1;
                        c = i + 1;
                    }
                }
                a.count[p] = c;
            }
        }
    }
}

// Write a new point's KP best (registers td/ti) as its final list or, with
// column chunks, as its chunk's keys.
template <std::uint32_t KP>
__device__ __forceinline__ void insert_write_list(const InsertArgs& a, std::uint32_t p,
                                                  std::uint32_t chunk, const float (&td)[KP],
                                                  const std::uint32_t (&ti)[KP]) {
    if (a.partial) {
        unsigned long long* out = a.partial + (std::size_t(p - a.n0) * a.slots + chunk) * KP;
#pragma unroll
        for (std::uint32_t i = 0; i < KP; ++i)
            out[i] = ti[i] == kNone ? ~0ull
                                    : (static_cast<unsigned long long>(fbits(td[i])) << 32) | ti[i];
    } else {
        std::uint32_t c = 0;
#pragma unroll
        for (std::uint32_t i = 0; i < KP; ++i) {
            if (i < a.k && ti[i] != kNone) {
                const std::size_t o = std::size_t(p) * a.k + i;
                a.ids[o] = ti[i];
                a.dist[o] = td[i];
                a.flags[o] = 1;
                c = i + 1;
            }
        }
        a.count[p] = c;
    }
}

// Gram-filter insert scan: the exact scan's results with the matrix cores
// deciding which exact distances are needed. Per 64-column tile, the
// 64 x 64 Gram of the (centred, FP16/BF16-rounded) new points and columns runs
// on 16x16x16 MFMA (wavefront w: columns 16w..16w+15, all 64 rows), giving
// fast distances D~ = N_q + N_c - 2G in LDS. A lane (one new point) then takes
// the rigorous lower bound of the exact distance (the Gram-filter join's bound)
// for each of its columns and computes the exact distance only when the bound
// does not rule out both uses: entering the point's list (bound <= its
// running threshold) and being offered to the column's list (bound below the
// column's k-th). Rows are centred on the block's first new point.
constexpr unsigned kGramTile = 64;

template <std::uint32_t STRIDE, std::uint32_t KP, bool BF16>
__global__ void __launch_bounds__(kInsThreads) insert_scan_gram_kernel(InsertArgs a) {
    extern __shared__ float lds[];
    constexpr std::uint32_t KP16 = (STRIDE + 15u) & ~15u, KS = KP16 / 16u;
    constexpr std::uint32_t GP = kGramTile + 1;              // D~ tile pitch
    float* s_tile = lds;                                      // kGramTile * STRIDE
    float* s_thr = s_tile + kGramTile * STRIDE;               // kGramTile
    float* s_N = s_thr + kGramTile;                           // kGramTile
    float* s_nb = s_N + kGramTile;                            // kGramTile
    float* s_q = s_nb + kGramTile;                            // kInsRows * STRIDE (queries)
    float* s_Nq = s_q + kInsRows * STRIDE;                    // kInsRows
    float* s_G = s_Nq + kInsRows;                             // kInsRows * GP
    float* s_ld = lds;                                        // merge phase: 2 * kInsRows * KP
    std::uint32_t* s_li = reinterpret_cast<std::uint32_t*>(lds + 2 * kInsRows * KP);
    const std::uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const std::uint32_t li = lane & 15u, lk = 4u * (lane >> 4);
    const float rho = GramConst<BF16>::rho;
    const float abs_norm = GramConst<BF16>::absc * sqrtf(float(KP16)) * 1.01f;
    const float cG = 2.0f * float(KP16 + 8) * 0x1p-24f;
    const float eta = float(STRIDE + 24) * 0x1p-24f + 0x1p-18f;
    const float nscale = 1.0f + float(KP16 + 4) * 0x1p-24f;
    const std::uint32_t chunk = blockIdx.x % a.chunks;
    const std::uint32_t col0 = a.col_base + chunk * a.chunk_cols;
    const std::uint32_t col1 = min(a.n1, col0 + a.chunk_cols);
    const std::uint32_t qstep = (gridDim.x / a.chunks) * kInsRows;
    unsigned long long n_pairs = 0, n_exact = 0;
    for (std::uint32_t rb = a.n0 + (blockIdx.x / a.chunks) * kInsRows; rb < a.n1; rb += qstep) {
        const std::uint32_t p = rb + lane;
        const bool has = p < a.n1;
        __syncthreads();  // LDS of the previous query block is consumed
        for (std::uint32_t e = threadIdx.x; e < kInsRows * (STRIDE / 4); e += kInsThreads) {
            const std::uint32_t r = e / (STRIDE / 4), f = e - r * (STRIDE / 4);
            reinterpret_cast<float4*>(s_q)[e] =
                rb + r < a.n1 ? reinterpret_cast<const float4*>(a.pts + std::size_t(rb + r) * STRIDE)[f]
                              : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        __syncthreads();
        float row[STRIDE];
#pragma unroll
        for (std::uint32_t f = 0; f < STRIDE / 4; ++f) {
            const float4 t = reinterpret_cast<const float4*>(s_q + lane * STRIDE)[f];
            row[4 * f] = t.x;
            row[4 * f + 1] = t.y;
            row[4 * f + 2] = t.z;
            row[4 * f + 3] = t.w;
        }
        // the query's rounded centred norm (centre: the block's first row)
        float Nq = 0.0f;
#pragma unroll
        for (std::uint32_t f = 0; f < STRIDE; ++f) {
            const float h = gram_round<BF16>(row[f] - s_q[f]);
            Nq = __builtin_fmaf(h, h, Nq);
        }
        if (w == 0) s_Nq[lane] = Nq;
        const float nbq = (sqrtf(Nq) * nscale + abs_norm) * (1.0f + 2.0f * rho);
        float td[KP];
        std::uint32_t ti[KP];
#pragma unroll
        for (std::uint32_t i = 0; i < KP; ++i) {
            td[i] = __builtin_inff();
            ti[i] = kNone;
        }
        float thr2 = __builtin_inff();
        float bd;
        std::uint32_t bi;
        search_bound<KP>(a, p, has, bd, bi, thr2);
        for (std::uint32_t c0 = col0; c0 < col1; c0 += kGramTile) {
            __syncthreads();  // the previous tile is consumed
            const std::uint32_t cend = min(kGramTile, col1 - c0);
            for (std::uint32_t e = threadIdx.x; e < kGramTile * (STRIDE / 4); e += kInsThreads) {
                const std::uint32_t c = e / (STRIDE / 4), f = e - c * (STRIDE / 4);
                reinterpret_cast<float4*>(s_tile)[e] =
                    c < cend ? reinterpret_cast<const float4*>(a.pts + std::size_t(c0 + c) * STRIDE)[f]
                             : make_float4(0.f, 0.f, 0.f, 0.f);
            }
            for (std::uint32_t c = threadIdx.x; c < cend; c += kInsThreads) {
                const std::uint32_t j = c0 + c;
                float t = -1.0f;  // new points take no proposals here
                if (j < a.n0) {
                    t = __builtin_inff();
                    if (a.count[j] >= a.k) {
                        const float dk = a.dist[std::size_t(j) * a.k + a.k - 1];
                        t = dk * dk * (1.0f + 0x1p-20f);  // the merge makes the exact test
                    }
                }
                s_thr[c] = t;
            }
            __syncthreads();
            if (threadIdx.x < kGramTile) {
                const std::uint32_t c = threadIdx.x;
                float N = 0.0f;
#pragma unroll
                for (std::uint32_t f = 0; f < STRIDE; ++f) {
                    const float h = gram_round<BF16>(s_tile[c * STRIDE + f] - s_q[f]);
                    N = __builtin_fmaf(h, h, N);
                }
                s_N[c] = N;
                s_nb[c] = (sqrtf(N) * nscale + abs_norm) * (1.0f + 2.0f * rho);
            }
            // Gram strip of wavefront w: columns 16w + li, all 64 rows
            float bfrag[KS][4];
#pragma unroll
            for (std::uint32_t ks = 0; ks < KS; ++ks) {
                const std::uint32_t k0 = 16u * ks + lk;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const std::uint32_t kk = k0 + static_cast<std::uint32_t>(q);
                    bfrag[ks][q] = kk < STRIDE ? s_tile[(16u * w + li) * STRIDE + kk] - s_q[kk] : 0.0f;
                }
            }
            __syncthreads();  // s_N complete
#pragma unroll
            for (std::uint32_t rt = 0; rt < kInsRows / 16u; ++rt) {
                gram_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (std::uint32_t ks = 0; ks < KS; ++ks) {
                    float afrag[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const std::uint32_t kk = 16u * ks + lk + static_cast<std::uint32_t>(q);
                        afrag[q] = kk < STRIDE ? s_q[(16u * rt + li) * STRIDE + kk] - s_q[kk] : 0.0f;
                    }
                    acc = gram_mfma<BF16>(afrag, bfrag[ks], acc);
                }
                const std::uint32_t col = 16u * w + li;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const std::uint32_t r = 16u * rt + lk + static_cast<std::uint32_t>(q);
                    s_G[r * GP + col] = (s_Nq[r] + s_N[col]) - 2.0f * acc[q];
                }
            }
            __syncthreads();
            for (std::uint32_t c = w; c < cend; c += 4) {
                const std::uint32_t j = c0 + c;
                if (!has || j == p) continue;
                const float dt = s_G[lane * GP + c];
                const float nbs = nbq + s_nb[c];
                float tlo = 0.0f;
                const float s2 = dt - cG * nbs * nbs;
                if (s2 > 0.0f) {
                    const float slo = sqrtf(s2) * (1.0f - 0x1p-20f);
                    const float delta =
                        (rho * nbs + 2.0f * abs_norm) * (1.0f + 0x1p-16f) + slo * 0x1p-21f;
                    tlo = slo - delta;
                }
                const float thr_c = s_thr[c];
                if (tlo > 0.0f) {
                    const float rlo = tlo * tlo * (1.0f - eta);
                    if (rlo > thr2 && !(rlo < thr_c)) {
                        if (a.tally) ++n_pairs;
                        continue;
                    }
                }
                if (a.tally) {
                    ++n_pairs;
                    ++n_exact;
                }
                const float d2 = l2sqr_reg<STRIDE>(row, s_tile + c * STRIDE);
                if (d2 <= thr2) {
                    const float d = sqrt_exact(d2);
                    if (key_less(d, j, td[KP - 1], ti[KP - 1]) && key_less(d, j, bd, bi)) {
                        list_insert<KP>(td, ti, d, j);
                        thr2 = fminf(thr2, td[KP - 1] * td[KP - 1] * (1.0f + 0x1p-20f));
                    }
                }
                if (d2 < thr_c) {
                    const std::uint32_t pos = atomicAdd(&a.fill[j], 1u);
                    if (pos < a.cap)
                        a.cand[std::size_t(j) * a.cap + pos] =
                            (static_cast<unsigned long long>(fbits(sqrt_exact(d2))) << 32) | p;
                }
            }
        }
        // merge the four partial lists of each new point
        __syncthreads();
        if (w >= 2) {
#pragma unroll
            for (std::uint32_t i = 0; i < KP; ++i) {
                s_ld[((w - 2) * kInsRows + lane) * KP + i] = td[i];
                s_li[((w - 2) * kInsRows + lane) * KP + i] = ti[i];
            }
        }
        __syncthreads();
        if (w < 2) {
            for (std::uint32_t i = 0; i < KP; ++i) {
                const float d = s_ld[(w * kInsRows + lane) * KP + i];
                const std::uint32_t id = s_li[(w * kInsRows + lane) * KP + i];
                if (id == kNone || !key_less(d, id, td[KP - 1], ti[KP - 1])) break;
                list_insert<KP>(td, ti, d, id);
            }
        }
        __syncthreads();
        if (w == 1) {
#pragma unroll
            for (std::uint32_t i = 0; i < KP; ++i) {
                s_ld[lane * KP + i] = td[i];
                s_li[lane * KP + i] = ti[i];
            }
        }
        __syncthreads();
        if (w == 0) {
            for (std::uint32_t i = 0; i < KP; ++i) {
                const float d = s_ld[lane * KP + i];
                const std::uint32_t id = s_li[lane * KP + i];
                if (id == kNone || !key_less(d, id, td[KP - 1], ti[KP - 1])) break;
                list_insert<KP>(td, ti, d, id);
            }
            if (has) insert_write_list<KP>(a, p, chunk, td, ti);
        }
    }
    if (a.tally) {
        n_pairs = wave_sum(n_pairs);
        n_exact = wave_sum(n_exact);
        if (lane == 0) {
            atomicAdd(&a.tally[0], n_pairs);
            atomicAdd(&a.tally[1], n_exact);
        }
    }
}

// The k best of a new point's per-chunk lists: its chunks * KP keys
// (distance bits << 32 | id, ~0 for empty) sorted in registers; keys order as
// (distance, id), the brute force's order, and ids of different chunks are
// distinct. One wavefront per new point.
template <int R>
__global__ void __launch_bounds__(64)
    insert_merge_kernel(const unsigned long long* __restrict__ partial, std::uint32_t n0,
                        std::uint32_t n1, std::uint32_t k, std::uint32_t keys,
                        std::uint32_t* __restrict__ ids, float* __restrict__ dist,
                        std::uint8_t* __restrict__ flags, std::uint32_t* __restrict__ count) {
    __shared__ unsigned long long s_key[64 * R];
    __shared__ unsigned long long s_top[kMaxK];
    const std::uint32_t lane = threadIdx.x;
    for (std::uint32_t p = n0 + blockIdx.x; p < n1; p += gridDim.x) {
        const unsigned long long* in = partial + std::size_t(p - n0) * keys;
        for (std::uint32_t i = lane; i < 64u * R; i += 64) s_key[i] = i < keys ? in[i] : ~0ull;
        __syncthreads();
        const std::uint32_t got = sort_unique_top<R>(s_key, 64u * R, k, s_top);
        __syncthreads();
        const std::uint32_t t = got < k ? got : k;
        if (lane < t) {
            const unsigned long long key = s_top[lane];
            const std::size_t o = std::size_t(p) * k + lane;
            ids[o] = static_cast<std::uint32_t>(key);
            dist[o] = __uint_as_float(static_cast<std::uint32_t>(key >> 32));
            flags[o] = 1;
        }
        if (lane == 0) count[p] = t;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// graph-search insertion: each new point finds its neighbours among the
// existing points by a beam search over the k-NN lists (the reference's
// anchored seeding, incremental_knn.cpp:118-216) instead of a scan of every
// point; one wavefront per new point
// ---------------------------------------------------------------------------

constexpr unsigned kSearchWaves = 2;         // wavefronts per block
constexpr unsigned kSearchHash = 2048;       // visited-set slots per wavefront (LDS)
constexpr unsigned kSearchVisitCap = 1536;   // no expansion past this many visits
constexpr unsigned kSearchHintScan = 256;    // hints scanned for entry points (kHintScan)
constexpr unsigned kSearchEntries = 8;       // random entry points (search_entries)
constexpr std::uint32_t kRefineRounds = 5;   // refinement rounds (the CLI's --refine-iters)

struct SearchArgs {
    const float* pts;
    std::uint32_t n0, n1, k, stride;  // existing points [0, n0), new points [n0, n1)
    const std::uint32_t* ids;         // existing lists
    const float* dist;
    const std::uint32_t* count;
    const std::uint32_t* hints;  // the selected anti-hubs: a one-level coarse index
    std::uint32_t nhints;
    unsigned long long* cand;  // proposal buffers of the existing points
    std::uint32_t* fill;
    std::uint32_t cap;
    unsigned long long* partial;  // per new point: `slots` key lists of kp keys
    std::uint32_t slots, slot, kp;
    std::uint64_t seed;
    unsigned long long* evals;  // distance evaluations (profiling), or nullptr
};

__device__ __forceinline__ std::uint64_t xorshift64(std::uint64_t* s) {
    std::uint64_t x = *s;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    *s = x;
    return x;
}

// Wave-wide ascending bitonic sort of one 64-bit key per lane.
__device__ __forceinline__ unsigned long long wave_sort_u64(unsigned long long key,
                                                            std::uint32_t lane) {
#pragma unroll
    for (std::uint32_t k2 = 2; k2 <= 64; k2 <<= 1) {
#pragma unroll
        for (std::uint32_t j = k2 >> 1; j > 0; j >>= 1) {
            const unsigned long long o = __shfl_xor(key, static_cast<int>(j), 64);
            const bool take_min = ((lane & j) == 0) == ((lane & k2) == 0);
            key = take_min ? (o < key ? o : key) : (o > key ? o : key);
        }
    }
    return key;
}

// Number of entries of the ascending 64-entry list `l` below key (strictly, or
// at most key with `le`).
__device__ __forceinline__ std::uint32_t rank_in(const unsigned long long* l,
                                                 unsigned long long key, bool le) {
    std::uint32_t lo = 0, hi = 64;
    while (lo < hi) {
        const std::uint32_t mid = (lo + hi) >> 1;
        const bool below = le ? (l[mid] <= key) : (l[mid] < key);
        if (below)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(64 * kSearchWaves) insert_search_kernel(SearchArgs a) {
    __shared__ std::uint32_t s_hash[kSearchWaves][kSearchHash];
    __shared__ unsigned long long s_beam[kSearchWaves][64], s_new[kSearchWaves][64];
    __shared__ unsigned long long s_out[kSearchWaves][128];
    __shared__ std::uint8_t s_outx[kSearchWaves][128];
    const std::uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const std::uint32_t stride = a.stride;
    unsigned long long evals = 0;
    for (std::uint32_t q = a.n0 + blockIdx.x * kSearchWaves + w; q < a.n1;
         q += gridDim.x * kSearchWaves) {
        const float* vq = a.pts + std::size_t(q) * stride;
        for (std::uint32_t i = lane; i < kSearchHash; i += 64) s_hash[w][i] = kNone;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // the beam: lane i holds the i-th best key (metric distance bits <<
        // 32 | id) and whether it was expanded
        unsigned long long beam = ~0ull;
        bool expanded = false;
        std::uint32_t visited = 0;
        std::uint64_t rs = a.seed ^ (0x9E3779B97F4A7C15ull * (std::uint64_t(q) + 1));

        // Offers one candidate id per lane (kNone: none): first visits get
        // their exact distance, are proposed to their own list when they beat
        // its k-th (the merge makes the exact test), and join the beam.
        auto offer = [&](std::uint32_t id) {
            bool fresh = false;
            if (id < a.n0) {
                std::uint32_t h = (id * 2654435761u) >> 21;  // 11 bits: kSearchHash slots
                for (int t = 0; t < 16; ++t) {
                    const std::uint32_t old = atomicCAS(&s_hash[w][h], kNone, id);
                    if (old == kNone) {
                        fresh = true;
                        break;
                    }
                    if (old == id) break;
                    h = (h + 1) & (kSearchHash - 1);
                }
            }
            unsigned long long key = ~0ull;
            if (fresh) {
                const float r = l2sqr_exact(vq, a.pts + std::size_t(id) * stride, stride);
                const float d = sqrt_exact(r);
                key = (static_cast<unsigned long long>(fbits(d)) << 32) | id;
                ++evals;
                bool propose = a.count[id] < a.k;
                if (!propose) {
                    const float dk = a.dist[std::size_t(id) * a.k + a.k - 1];
                    propose = r < dk * dk * (1.0f + 0x1p-20f);
                }
                if (propose) {
                    const std::uint32_t pos = atomicAdd(&a.fill[id], 1u);
                    if (pos < a.cap)
                        a.cand[std::size_t(id) * a.cap + pos] =
                            (static_cast<unsigned long long>(fbits(d)) << 32) | q;
                }
            }
            visited += static_cast<std::uint32_t>(__popcll(__ballot(fresh)));
            if (!__ballot(fresh)) return;
            // merge the sorted offers with the beam, keep the 64 smallest
            const unsigned long long nk = wave_sort_u64(key, lane);
            s_beam[w][lane] = beam;
            s_new[w][lane] = nk;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            const std::uint32_t pb = lane + rank_in(s_new[w], beam, false);
            const std::uint32_t pn = lane + rank_in(s_beam[w], nk, true);
            s_out[w][pb] = beam;
            s_outx[w][pb] = expanded ? 1 : 0;
            s_out[w][pn] = nk;
            s_outx[w][pn] = 0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            beam = s_out[w][lane];
            expanded = s_outx[w][lane] != 0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
        };

        // entry points: the closest of up to kSearchHintScan hints, then
        // kSearchEntries uniform draws
        if (a.nhints) {
            const std::uint32_t scan = a.nhints < kSearchHintScan ? a.nhints : kSearchHintScan;
            const std::uint32_t hstep = a.nhints / scan > 1 ? a.nhints / scan : 1;
            const std::uint32_t off = static_cast<std::uint32_t>(xorshift64(&rs) % a.nhints);
            for (std::uint32_t t0 = 0; t0 < scan; t0 += 64) {
                const std::uint32_t t = t0 + lane;
                offer(t < scan ? a.hints[(off + std::size_t(t) * hstep) % a.nhints] : kNone);
            }
        }
        {
            std::uint32_t id = kNone;
            for (std::uint32_t e = 0; e < kSearchEntries; ++e) {
                const std::uint32_t r = static_cast<std::uint32_t>(xorshift64(&rs) % a.n0);
                if (lane == e) id = r;
            }
            offer(id);
        }
        // expand the closest unexpanded entry until every entry of the beam is
        // expanded (nothing outside it can improve it) or the visit cap
        while (visited < kSearchVisitCap) {
            const unsigned long long open = __ballot(beam != ~0ull && !expanded);
            if (!open) break;
            const std::uint32_t b = static_cast<std::uint32_t>(__ffsll(static_cast<long long>(open)) - 1);
            const std::uint32_t c = static_cast<std::uint32_t>(
                __shfl(static_cast<unsigned int>(beam & 0xFFFFFFFFull), static_cast<int>(b), 64));
            if (lane == b) expanded = true;
            const std::uint32_t deg = a.count[c];
            offer(lane < deg && lane < a.k ? a.ids[std::size_t(c) * a.k + lane] : kNone);
        }
        // the kp best as this point's key list in its search slot
        if (lane < a.kp) {
            unsigned long long* out =
                a.partial + (std::size_t(q - a.n0) * a.slots + a.slot) * a.kp;
            out[lane] = beam;
        }
    }
    if (a.evals) {
        evals = wave_sum(evals);
        if (lane == 0) atomicAdd(a.evals, evals);
    }
}

using InsertFn = void (*)(InsertArgs);

template <std::uint32_t KP>
InsertFn pick_insert_stride(std::uint32_t stride) {
    switch (stride) {
        case 8: return insert_scan_kernel<8, KP>;
        case 16: return insert_scan_kernel<16, KP>;
        case 24: return insert_scan_kernel<24, KP>;
        case 32: return insert_scan_kernel<32, KP>;
        case 40: return insert_scan_kernel<40, KP>;
        case 48: return insert_scan_kernel<48, KP>;
        case 56: return insert_scan_kernel<56, KP>;
        case 64: return insert_scan_kernel<64, KP>;
        default: return nullptr;
    }
}

InsertFn pick_insert(std::uint32_t stride, std::uint32_t k, std::uint32_t* kp) {
    if (k <= 16) {
        *kp = 16;
        return pick_insert_stride<16>(stride);
    }
    if (k <= 32) {
        *kp = 32;
        return pick_insert_stride<32>(stride);
    }
    *kp = 64;
    return pick_insert_stride<64>(stride);
}

template <std::uint32_t KP, bool BF16>
InsertFn pick_insert_gram_stride(std::uint32_t stride) {
    switch (stride) {
        case 8: return insert_scan_gram_kernel<8, KP, BF16>;
        case 16: return insert_scan_gram_kernel<16, KP, BF16>;
        case 24: return insert_scan_gram_kernel<24, KP, BF16>;
        case 32: return insert_scan_gram_kernel<32, KP, BF16>;
        case 40: return insert_scan_gram_kernel<40, KP, BF16>;
        case 48: return insert_scan_gram_kernel<48, KP, BF16>;
        case 56: return insert_scan_gram_kernel<56, KP, BF16>;
        case 64: return insert_scan_gram_kernel<64, KP, BF16>;
        default: return nullptr;
    }
}

template <bool BF16>
InsertFn pick_insert_gram(std::uint32_t stride, std::uint32_t k) {
    if (k <= 16) return pick_insert_gram_stride<16, BF16>(stride);
    if (k <= 32) return pick_insert_gram_stride<32, BF16>(stride);
    return pick_insert_gram_stride<64, BF16>(stride);
}

void check_join_capacity(std::uint32_t cap) {
    if (cap < 16 || cap > kMaxCap)
        throw std::runtime_error("nndescent: join_capacity must be in 16..512");
}

}  // namespace

void build_knn(Graph& g, const coresg_hip_build_params& p, coresg_hip_build_stats* st) {
    if (!g.points_set) throw std::runtime_error("build_knn: points not set");
    const auto t0 = std::chrono::steady_clock::now();
    const std::uint32_t n = g.n, k = g.k;
    hipStream_t s = g.stream;
    const std::size_t nk = std::size_t(n) * k;
    fill_u32(g.ids.p, nk, kNone, s);
    fill_u32(reinterpret_cast<std::uint32_t*>(g.dist.p), nk, 0x7F800000u, s);
    fill_u8(g.flags.p, nk, 0, s);
    fill_u32(g.count.p, n, 0u, s);
    DescentResult r;
    if (!g.inst_off.empty() && p.brute_force_limit)
        throw std::runtime_error("build_knn: brute_force_limit is not supported with instances");
    if (p.brute_force_limit && n <= p.brute_force_limit) {
        brute_force_kernel<<<capped_blocks(n, 64), 64, 0, s>>>(g.pts.p, n, g.stride, k, g.ids.p,
                                                               g.dist.p, g.flags.p, g.count.p);
        CSG_LAUNCH_CHECK();
    } else {
        const std::uint32_t cap = p.join_capacity ? p.join_capacity : kDefaultCap;
        check_join_capacity(cap);
        r = run_descent(&g, g.pts.p, n, g.stride, k, g.ids.p, g.dist.p, g.flags.p, g.count.p,
                        p.sample, p.iters, p.delta, p.seed, true, cap, p.gram, s, false, true);
    }
    CSG_CHECK(hipStreamSynchronize(s));
    g.knn_set = true;
    g.edges_ready = false;
    g.emit_ready = false;
    g.hub_ready = false;
    if (st) {
        st->rounds = r.rounds;
        st->updates = r.updates;
        st->knn_seconds =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
}

void nndescent_host(int device, const coresg_hip_nnd_config& cfg, const float* vectors,
                    std::uint32_t* ids, float* dists, std::uint8_t* flags, std::uint32_t* counts,
                    std::uint64_t* updates) {
    (void)device;
    if (cfg.n == 0 || cfg.k == 0) return;
    if (cfg.metric != CORESG_HIP_METRIC_L2 && cfg.metric != CORESG_HIP_METRIC_COSINE)
        throw std::runtime_error("nndescent: unsupported metric for the device path "
                                 "(L2 and cosine only)");
    if (cfg.stride % 8 != 0 || cfg.stride < cfg.dim)
        throw std::runtime_error("nndescent: stride must be round_up(dim, 8)");
    OwnedStream os;
    hipStream_t s = os.s;
    const std::size_t n = cfg.n, nk = n * cfg.k;
    DBuf<float> d_pts, d_dist;
    DBuf<std::uint32_t> d_ids, d_cnt;
    DBuf<std::uint8_t> d_flags;
    d_pts.reserve(n * cfg.stride);
    d_dist.reserve(nk);
    d_ids.reserve(nk);
    d_cnt.reserve(n);
    d_flags.reserve(nk);
    CSG_CHECK(hipMemcpyAsync(d_pts.p, vectors, n * cfg.stride * 4, hipMemcpyHostToDevice, s));
    CSG_CHECK(hipMemcpyAsync(d_ids.p, ids, nk * 4, hipMemcpyHostToDevice, s));
    CSG_CHECK(hipMemcpyAsync(d_dist.p, dists, nk * 4, hipMemcpyHostToDevice, s));
    CSG_CHECK(hipMemcpyAsync(d_flags.p, flags, nk, hipMemcpyHostToDevice, s));
    CSG_CHECK(hipMemcpyAsync(d_cnt.p, counts, n * 4, hipMemcpyHostToDevice, s));
    const DescentResult r =
        run_descent(nullptr, d_pts.p, cfg.n, cfg.stride, cfg.k, d_ids.p, d_dist.p, d_flags.p,
                    d_cnt.p, cfg.sample, cfg.iters, cfg.delta, cfg.seed, cfg.random_fill != 0,
                    kDefaultCap, kGramAuto, s, cfg.metric == CORESG_HIP_METRIC_COSINE);
    CSG_CHECK(hipMemcpyAsync(ids, d_ids.p, nk * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipMemcpyAsync(dists, d_dist.p, nk * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipMemcpyAsync(flags, d_flags.p, nk, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipMemcpyAsync(counts, d_cnt.p, n * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));
    if (updates) *updates = r.updates;
}

void insert_points(Graph& g, const float* rows, std::uint32_t count, int where, int padded,
                   const coresg_hip_build_params& p, coresg_hip_build_stats* st) {
    if (!g.knn_set) throw std::runtime_error("insert: the CORE-SG is not built");
    if (!g.inst_off.empty()) throw std::runtime_error("insert: not supported with instances");
    const auto t0 = std::chrono::steady_clock::now();
    if (count == 0) return;
    const std::uint32_t n0 = g.n, k = g.k, stride = g.stride;
    const std::uint64_t n1_64 = std::uint64_t(n0) + count;
    if (n1_64 >= 0x7FFFFFFFull) throw std::runtime_error("insert: point count exceeds 2^31");
    const std::uint32_t n1 = static_cast<std::uint32_t>(n1_64);
    std::uint32_t kp = 0;
    InsertFn fn = pick_insert(stride, k, &kp);
    // the batch scan: exact (the library's choice), or the Gram filter: the
    // same lists with exact distances for only 1-3 % of the pairs, but at the
    // scan's occupancy (2 waves per SIMD: 32-entry register lists per point)
    // the per-tile Gram staging and barriers cost more than the distances
    // (MI355X, 20k points into 500k, d = 24: 92 ms exact, 117 ms FP16 filter)
    const std::uint32_t gmode = p.gram == kGramAuto ? kGramExact : p.gram;
    if (gmode > kGramBF16) throw std::runtime_error("insert: gram must be 0..3");
    if (fn && gmode == kGramF16) fn = pick_insert_gram<false>(stride, k);
    if (fn && gmode == kGramBF16) fn = pick_insert_gram<true>(stride, k);
    const bool gram_scan = gmode != kGramExact;
    const bool brute = p.brute_force_limit && n1 <= p.brute_force_limit;
    if (!brute && !fn)
        throw std::runtime_error("insert: rows wider than 64 floats are not supported on device");
    if (k > kMaxK) throw std::runtime_error("insert: k above 64 is not supported on device");
    hipStream_t s = g.stream;

    // grow the resident arrays, keeping the existing points and lists
    g.pts.grow(std::size_t(n1) * stride, std::size_t(n0) * stride, s);
    g.ids.grow(std::size_t(n1) * k, std::size_t(n0) * k, s);
    g.dist.grow(std::size_t(n1) * k, std::size_t(n0) * k, s);
    g.flags.grow(std::size_t(n1) * k, std::size_t(n0) * k, s);
    g.count.grow(n1, n0, s);
    float* dst = g.pts.p + std::size_t(n0) * stride;
    const hipMemcpyKind kind =
        where == CORESG_HIP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (padded || g.dim == stride) {
        CSG_CHECK(hipMemcpyAsync(dst, rows, std::size_t(count) * stride * 4, kind, s));
    } else {
        fill_u32(reinterpret_cast<std::uint32_t*>(dst), std::size_t(count) * stride, 0u, s);
        CSG_CHECK(hipMemcpy2DAsync(dst, std::size_t(stride) * 4, rows, std::size_t(g.dim) * 4,
                                   std::size_t(g.dim) * 4, count, kind, s));
    }
    const std::size_t nk_new = std::size_t(count) * k, base = std::size_t(n0) * k;
    fill_u32(g.ids.p + base, nk_new, kNone, s);
    fill_u32(reinterpret_cast<std::uint32_t*>(g.dist.p) + base, nk_new, 0x7F800000u, s);
    fill_u8(g.flags.p + base, nk_new, 0, s);
    fill_u32(g.count.p + n0, count, 0u, s);
    g.n = n1;
    g.knn_set = true;
    g.edges_ready = false;
    g.emit_ready = false;
    g.hub_ready = false;

    std::uint64_t applied = 0;
    std::uint32_t refine_rounds = 0;
    if (brute) {
        // exact lists for every point (incremental_knn.cpp:1169-1180)
        brute_force_kernel<<<capped_blocks(n1, 64), 64, 0, s>>>(g.pts.p, n1, stride, k, g.ids.p,
                                                                g.dist.p, g.flags.p, g.count.p);
        CSG_LAUNCH_CHECK();
    } else {
        DescentBuffers& b = g.nnd;
        const std::uint32_t cap = kDefaultCap;
        b.cand.reserve(std::size_t(n0 ? n0 : 1) * cap);
        b.fill.reserve(n0 ? n0 : 1);
        b.partial.reserve(kMergeBlocks);
        b.updates.reserve(1);
        fill_u32(b.fill.p, n0 ? n0 : 1, 0u, s);
        // Graph search (insert_search = 1): the batch's neighbours among the
        // existing points come from a beam search over their lists, and the
        // scan below covers only the batch itself (its own points' pairs);
        // otherwise the scan covers every point.
        const bool search = p.insert_search != 0 && n0 > 0;
        const std::uint32_t col_base = search ? n0 : 0u;
        const std::uint32_t ncols = n1 - col_base;
        // Column chunks so that small batches still fill the chip: about 2048
        // blocks, at most 16 chunks (16 * KP <= 512 keys per point to merge,
        // one list fewer with the search's)
        const std::uint32_t qblocks = (count + kInsRows - 1) / kInsRows;
        std::uint32_t chunks = (2048u + qblocks - 1) / qblocks;
        chunks = std::max(1u, std::min({chunks, 16u, 512u / kp - (search ? 1u : 0u),
                                        (ncols + kInsTile - 1) / kInsTile}));
        const std::uint32_t chunk_cols =
            ((ncols + chunks - 1) / chunks + kInsTile - 1) / kInsTile * kInsTile;
        chunks = (ncols + chunk_cols - 1) / chunk_cols;
        const std::uint32_t slots = chunks + (search ? 1u : 0u);
        unsigned long long* partial = nullptr;
        if (slots > 1) {
            g.ins_partial.reserve(std::size_t(count) * slots * kp);
            partial = g.ins_partial.p;
        }
        unsigned long long* tally = nullptr;
        if (gram_scan && g.profiling) {
            b.gram_tally.reserve(3);
            fill_u64(reinterpret_cast<std::uint64_t*>(b.gram_tally.p), 3, 0ull, s);
            tally = b.gram_tally.p;
        }
        InsertArgs ia{g.pts.p, n0, n1, k, g.ids.p, g.dist.p, g.flags.p, g.count.p,
                      b.cand.p, b.fill.p, cap, chunks, chunk_cols, partial, tally,
                      col_base, slots};
        if (search) {
            // entry hints: the current anti-hub selection (over the existing
            // points), as the reference passes it (incremental_knn.cpp:135-160)
            const std::uint32_t nh = static_cast<std::uint32_t>(g.selected.size());
            g.ins_hints.reserve(nh ? nh : 1);
            if (nh)
                CSG_CHECK(hipMemcpyAsync(g.ins_hints.p, g.selected.data(), std::size_t(nh) * 4,
                                         hipMemcpyHostToDevice, s));
            unsigned long long* evals = nullptr;
            if (g.profiling) {
                b.tally.reserve(3);
                fill_u64(reinterpret_cast<std::uint64_t*>(b.tally.p), 1, 0ull, s);
                evals = b.tally.p;
            }
            SearchArgs sa{g.pts.p, n0, n1, k, stride, g.ids.p, g.dist.p, g.count.p,
                          g.ins_hints.p, nh, b.cand.p, b.fill.p, cap, partial, slots, chunks,
                          kp, p.seed ^ (std::uint64_t(n0) << 20), evals};
            const unsigned sblocks = static_cast<unsigned>(
                std::min<std::size_t>((count + kSearchWaves - 1) / kSearchWaves, 16384));
            if (g.yield_hook) g.yield_hook();
            ScopedKernelTimer t(g, "insert_search");
            insert_search_kernel<<<sblocks, 64 * kSearchWaves, 0, s>>>(sa);
            CSG_LAUNCH_CHECK();
            if (evals) {
                unsigned long long ev = 0;
                CSG_CHECK(hipMemcpyAsync(&ev, evals, 8, hipMemcpyDeviceToHost, s));
                CSG_CHECK(hipStreamSynchronize(s));
                g.add_work("insert_search", {count, ev, 0, 0});
            }
        }
        const std::size_t lds_tile =
            gram_scan ? (std::size_t(kGramTile) * stride + 3 * kGramTile + std::size_t(kInsRows) * stride +
                         kInsRows + std::size_t(kInsRows) * (kGramTile + 1)) *
                            sizeof(float)
                      : (std::size_t(kInsTile) * stride + kInsTile) * sizeof(float);
        const std::size_t lds_merge = std::size_t(4) * kInsRows * kp * sizeof(float);
        const std::size_t lds = std::max(lds_tile, lds_merge);
        if (lds > 64 * 1024)
            CSG_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          static_cast<int>(lds)));
        const unsigned blocks = static_cast<unsigned>(
            std::min<std::size_t>(qblocks, 65535 / chunks) * chunks);
        if (g.yield_hook) g.yield_hook();
        {
            ScopedKernelTimer t(g, "insert_scan");
            hipLaunchKernelGGL(fn, dim3(blocks), dim3(kInsThreads), lds, s, ia);
            CSG_LAUNCH_CHECK();
            if (partial) {
                const std::uint32_t keys = slots * kp;
                const unsigned mblocks = static_cast<unsigned>(std::min<std::size_t>(count, 65536));
                if (keys <= 64)
                    insert_merge_kernel<1><<<mblocks, 64, 0, s>>>(partial, n0, n1, k, keys, g.ids.p,
                                                                  g.dist.p, g.flags.p, g.count.p);
                else if (keys <= 128)
                    insert_merge_kernel<2><<<mblocks, 64, 0, s>>>(partial, n0, n1, k, keys, g.ids.p,
                                                                  g.dist.p, g.flags.p, g.count.p);
                else if (keys <= 256)
                    insert_merge_kernel<4><<<mblocks, 64, 0, s>>>(partial, n0, n1, k, keys, g.ids.p,
                                                                  g.dist.p, g.flags.p, g.count.p);
                else
                    insert_merge_kernel<8><<<mblocks, 64, 0, s>>>(partial, n0, n1, k, keys, g.ids.p,
                                                                  g.dist.p, g.flags.p, g.count.p);
                CSG_LAUNCH_CHECK();
            }
        }
        if (n0) {
            if (g.yield_hook) g.yield_hook();
            const unsigned merge_blocks = static_cast<unsigned>(std::min<std::size_t>(n0, kMergeBlocks));
            pick_merge(k)<<<merge_blocks, 64, 0, s>>>(n0, n1, k, g.ids.p, g.dist.p, g.flags.p,
                                                     g.count.p, b.cand.p, b.fill.p, cap,
                                                     b.partial.p, nullptr);
            CSG_LAUNCH_CHECK();
            sum_partials_kernel<<<1, kSumThreads, 0, s>>>(b.partial.p, merge_blocks, b.updates.p,
                                                          nullptr, 0.0);
            CSG_LAUNCH_CHECK();
            unsigned long long upd = 0;
            CSG_CHECK(hipMemcpyAsync(&upd, b.updates.p, 8, hipMemcpyDeviceToHost, s));
            CSG_CHECK(hipStreamSynchronize(s));
            applied = upd;
        }
        if (search) {
            // restricted refinement (the reference's refine_nnd after the
            // merge, incremental_knn.cpp:873-987): descent rounds over the
            // grown graph, which only nodes holding entries not yet joined --
            // the batch's lists and the entries it just placed -- take part in;
            // up to kRefineRounds, stopping at the usual update threshold
            const DescentResult r =
                run_descent(&g, g.pts.p, n1, stride, k, g.ids.p, g.dist.p, g.flags.p, g.count.p,
                            p.sample, kRefineRounds, p.delta, p.seed ^ 0x5851F42D4C957F2Dull,
                            false, kDefaultCap, p.gram, s);
            applied += r.updates;
            refine_rounds = r.rounds;
        }
        if (tally) {
            unsigned long long gt[2] = {0, 0};
            CSG_CHECK(hipMemcpyAsync(gt, tally, sizeof(gt), hipMemcpyDeviceToHost, s));
            CSG_CHECK(hipStreamSynchronize(s));
            g.add_work("insert_gram", {gt[0], gt[1], 0, 0});
        }
    }
    CSG_CHECK(hipStreamSynchronize(s));
    const double knn_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (g.yield_hook) g.yield_hook();
    build_bridges(g, p, st);
    if (st) {
        st->knn_seconds = knn_s;
        st->rounds = refine_rounds;
        st->updates = applied;
        st->total_seconds =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
}

}  // namespace coresg_hip
