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
//   2. join: one 128-thread group per node stages its candidate rows in LDS,
//      computes every new x (new u old) distance in the reference's exact
//      summation order, keeps for each member the best T candidates that beat
//      its current k-th distance, and appends each as a packed
//      (distance bits << 32 | id) key to the target's bounded candidate buffer
//      (atomic cursor; proposals beyond the capacity are dropped, which only
//      happens in the first rounds of a cold start).
//   3. merge: one 64-thread group per node drops duplicates, self and ids
//      already listed, ranks the buffer by counting, and merges the best into
//      the sorted list (new entries after equal distances, flagged new).
// The early stop is the reference's: updates <= delta * k * n.
#include <algorithm>
#include <chrono>
#include <memory>
#include <vector>

#include "build.hpp"
#include "scan.hpp"

namespace coresg_hip {
namespace {

constexpr std::uint32_t kMaxK = 64;        // list length handled by one wavefront
constexpr std::uint32_t kMaxSample = 32;   // reference cap (cuda_common.cuh:32)
constexpr std::uint32_t kJoinThreads = 128;

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
                       std::uint32_t* __restrict__ count) {
    const std::uint32_t lane = threadIdx.x & 63u;
    const std::uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const std::uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    for (std::uint32_t node = wave; node < n; node += waves) {
        const std::size_t base = std::size_t(node) * k;
        const std::uint32_t cnt0 = count[node];
        const std::uint32_t needed = k - cnt0;
        const std::uint32_t attempts = needed + needed / 2 + 4;
        // NodeRng(seed, node, 0): state s0, one discarded draw; draw t uses
        // state s0 + (t + 2) * phi.
        const std::uint64_t s0 = seed ^ (0x9E3779B97F4A7C15ull * (std::uint64_t(node) + 1)) ^
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
                cand = static_cast<std::uint32_t>(z >> 32) % n;
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
            for (std::uint32_t j = 0; j < 64; ++j) {
                const std::uint32_t oc = __shfl(cand, j, 64);
                ok = ok && !(j < lane && oc == cand);
            }
            const std::uint64_t mask = __ballot(ok);
            const std::uint32_t rank = accepted + __popcll(mask & ((1ull << lane) - 1ull));
            const bool take = ok && rank < needed;
            float d = __builtin_inff();
            if (take)
                d = sqrt_exact(l2sqr_exact(pts + std::size_t(node) * stride,
                                           pts + std::size_t(cand) * stride, stride));
            // move taken candidates into lanes [accepted, accepted + taken)
            const std::uint64_t tmask = __ballot(take);
            const std::uint32_t taken = __popcll(tmask);
            for (std::uint32_t r = 0; r < taken; ++r) {
                // lane holding the r-th taken candidate
                std::uint64_t mm = tmask;
                for (std::uint32_t q = 0; q < r; ++q) mm &= mm - 1ull;
                const int src = __ffsll(static_cast<unsigned long long>(mm)) - 1;
                const std::uint32_t cid = __shfl(cand, src, 64);
                const float cd = __shfl(d, src, 64);
                if (lane == accepted + r) {
                    acc_id = cid;
                    acc_d = cd;
                }
            }
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

__global__ void __launch_bounds__(256)
    sample_kernel(std::uint32_t n, std::uint32_t k, const std::uint32_t* __restrict__ ids,
                  std::uint8_t* __restrict__ flags, const std::uint32_t* __restrict__ count,
                  std::uint32_t smp, std::uint32_t* __restrict__ new_list,
                  std::uint32_t* __restrict__ new_cnt, std::uint32_t* __restrict__ old_list,
                  std::uint32_t* __restrict__ old_cnt, std::uint32_t* __restrict__ rev_new,
                  std::uint32_t* __restrict__ rev_new_cnt, std::uint32_t* __restrict__ rev_old,
                  std::uint32_t* __restrict__ rev_old_cnt) {
    const std::uint32_t lane = threadIdx.x & 63u;
    const std::uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const std::uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    const std::uint32_t max_new = smp, max_old = 3u * smp;
    for (std::uint32_t v = wave; v < n; v += waves) {
        const std::size_t base = std::size_t(v) * k;
        const std::uint32_t c = count[v];
        const bool has = lane < c;
        const std::uint32_t id = has ? ids[base + lane] : kNone;
        const bool isnew = has && flags[base + lane] != 0;
        const bool isold = has && !isnew;
        const std::uint64_t lt = (1ull << lane) - 1ull;
        const std::uint64_t mnew = __ballot(isnew), mold = __ballot(isold);
        const std::uint32_t rn = __popcll(mnew & lt), ro = __popcll(mold & lt);
        const bool take_new = isnew && rn < max_new && id < n;
        const bool take_old = isold && ro < max_old && id < n;
        if (isnew && rn < max_new) flags[base + lane] = 0;
        if (take_new) {
            new_list[std::size_t(v) * max_new + rn] = id;
            const std::uint32_t pos = atomicAdd(&rev_new_cnt[id], 1u);
            if (pos < smp) rev_new[std::size_t(id) * smp + pos] = v;
        }
        if (take_old) {
            old_list[std::size_t(v) * max_old + ro] = id;
            const std::uint32_t pos = atomicAdd(&rev_old_cnt[id], 1u);
            if (pos < smp) rev_old[std::size_t(id) * smp + pos] = v;
        }
        if (lane == 0) {
            const std::uint32_t tn = __popcll(mnew), to = __popcll(mold);
            new_cnt[v] = tn < max_new ? tn : max_new;
            old_cnt[v] = to < max_old ? to : max_old;
        }
    }
}

// ---------------------------------------------------------------------------
// local join
// ---------------------------------------------------------------------------

struct JoinArgs {
    const float* pts;
    std::uint32_t n, stride, k, smp;
    const float* dist;
    const std::uint32_t *new_list, *new_cnt, *old_list, *old_cnt;
    const std::uint32_t *rev_new, *rev_new_cnt, *rev_old, *rev_old_cnt;
    unsigned long long* cand;  // n x cap proposal buffer
    std::uint32_t* fill;       // proposals offered per node this round
    std::uint32_t cap;
    std::uint32_t keep;        // best candidates proposed per member and node
};

// LDS layout (floats/u32): ids[S], thr[S], rows[S*stride], D[A*S].
// KEEP = candidates proposed per member (the best that beat its k-th).
template <std::uint32_t KEEP>
__global__ void __launch_bounds__(kJoinThreads) join_kernel(JoinArgs a) {
    extern __shared__ float lds[];
    const std::uint32_t smp = a.smp;
    const std::uint32_t S = 6u * smp;       // A + 3*smp old + smp reverse old
    std::uint32_t* s_id = reinterpret_cast<std::uint32_t*>(lds);
    float* s_thr = lds + S;
    float* s_rows = lds + 2 * S;
    float* s_D = s_rows + std::size_t(S) * a.stride;
    __shared__ std::uint32_t s_na, s_ns;
    const std::uint32_t tid = threadIdx.x;
    const std::uint32_t stride = a.stride;

    for (std::uint32_t v = blockIdx.x; v < a.n; v += gridDim.x) {
        // candidate ids: A = new u reverse-new, then old u reverse-old
        if (tid == 0) {
            const std::uint32_t nn = a.new_cnt[v];
            std::uint32_t rn = a.rev_new_cnt[v];
            rn = rn < smp ? rn : smp;
            const std::uint32_t no = a.old_cnt[v];
            std::uint32_t ro = a.rev_old_cnt[v];
            ro = ro < smp ? ro : smp;
            s_na = nn + rn;
            s_ns = nn + rn + no + ro;
        }
        __syncthreads();
        const std::uint32_t na = s_na, ns = s_ns;
        if (na == 0) {
            __syncthreads();
            continue;
        }
        for (std::uint32_t i = tid; i < ns; i += kJoinThreads) {
            const std::uint32_t nn = a.new_cnt[v];
            std::uint32_t rn = a.rev_new_cnt[v];
            rn = rn < smp ? rn : smp;
            const std::uint32_t no = a.old_cnt[v];
            std::uint32_t id;
            if (i < nn) id = a.new_list[std::size_t(v) * smp + i];
            else if (i < nn + rn) id = a.rev_new[std::size_t(v) * smp + (i - nn)];
            else if (i < nn + rn + no) id = a.old_list[std::size_t(v) * 3 * smp + (i - nn - rn)];
            else id = a.rev_old[std::size_t(v) * smp + (i - nn - rn - no)];
            s_id[i] = id < a.n ? id : kNone;
        }
        __syncthreads();
        // keep the first occurrence of every id (A before old)
        for (std::uint32_t i = tid; i < ns; i += kJoinThreads) {
            const std::uint32_t id = s_id[i];
            bool dup = (id == kNone);
            for (std::uint32_t j = 0; j < i && !dup; ++j) dup = (s_id[j] == id);
            const float t = (id == kNone) ? 0.0f : a.dist[std::size_t(id) * a.k + a.k - 1];
            s_thr[i] = dup ? -1.0f : t * t;  // -1 marks a dropped member
        }
        __syncthreads();
        for (std::uint32_t e = tid; e < ns * stride; e += kJoinThreads) {
            const std::uint32_t i = e / stride, d = e - i * stride;
            const std::uint32_t id = s_id[i];
            s_rows[e] = (s_thr[i] >= 0.0f) ? a.pts[std::size_t(id) * stride + d] : 0.0f;
        }
        __syncthreads();
        for (std::uint32_t e = tid; e < na * ns; e += kJoinThreads) {
            const std::uint32_t r = e / ns, c = e - r * ns;
            float val = __builtin_inff();
            if (r != c && s_thr[r] >= 0.0f && s_thr[c] >= 0.0f)
                val = l2sqr_exact(s_rows + std::size_t(r) * stride, s_rows + std::size_t(c) * stride,
                                  stride);
            s_D[std::size_t(r) * S + c] = val;
        }
        __syncthreads();
        // one member per thread: best `keep` candidates that beat its k-th
        for (std::uint32_t x = tid; x < ns; x += kJoinThreads) {
            const float thr = s_thr[x];
            if (thr < 0.0f) continue;
            constexpr std::uint32_t kKeepMax = KEEP;
            float bd[kKeepMax];
            std::uint32_t bi[kKeepMax];
#pragma unroll
            for (std::uint32_t q = 0; q < kKeepMax; ++q) {
                bd[q] = __builtin_inff();
                bi[q] = kNone;
            }
            const std::uint32_t lim = x < na ? ns : na;
            for (std::uint32_t y = 0; y < lim; ++y) {
                const float r = x < na ? s_D[std::size_t(x) * S + y] : s_D[std::size_t(y) * S + x];
                if (!(r < thr) || !(r < bd[kKeepMax - 1])) continue;
                // insert (r, y) keeping bd ascending; `keep` caps the used prefix
                float cr = r;
                std::uint32_t cy = y;
#pragma unroll
                for (std::uint32_t q = 0; q < kKeepMax; ++q) {
                    if (cr < bd[q]) {
                        const float td = bd[q];
                        const std::uint32_t ti = bi[q];
                        bd[q] = cr;
                        bi[q] = cy;
                        cr = td;
                        cy = ti;
                    }
                }
            }
            const std::uint32_t target = s_id[x];
            std::uint32_t nk = 0;
#pragma unroll
            for (std::uint32_t q = 0; q < kKeepMax; ++q) nk += (bi[q] != kNone) ? 1u : 0u;
            if (nk == 0) continue;
            const std::uint32_t pos = atomicAdd(&a.fill[target], nk);
            unsigned long long* base = a.cand + std::size_t(target) * a.cap;
#pragma unroll
            for (std::uint32_t q = 0; q < kKeepMax; ++q) {
                if (q >= nk || pos + q >= a.cap) break;  // bi is filled front to back
                const std::uint32_t cand = s_id[bi[q]];
                const float d = sqrt_exact(bd[q]);
                base[pos + q] = (static_cast<unsigned long long>(fbits(d)) << 32) | cand;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// merge
// ---------------------------------------------------------------------------

constexpr std::uint32_t kMaxCap = 256;

// One 64-thread group per node. Candidates of the round are in
// cand[v*cap .. v*cap + min(fill, cap)). A candidate survives if it is not the
// node itself, not already listed, the first copy of its key, and (for a full
// list) closer than the current k-th; survivors are ranked by (distance, id)
// and the best k merged with the list by rank counting.
__global__ void __launch_bounds__(64)
    merge_kernel(std::uint32_t n, std::uint32_t k, std::uint32_t* __restrict__ ids,
                 float* __restrict__ dist, std::uint8_t* __restrict__ flags,
                 std::uint32_t* __restrict__ count, const unsigned long long* __restrict__ cand,
                 std::uint32_t* __restrict__ fill, std::uint32_t cap,
                 unsigned long long* __restrict__ updates) {
    __shared__ unsigned long long s_key[kMaxCap];
    __shared__ std::uint8_t s_ok[kMaxCap];
    __shared__ std::uint32_t s_ex[kMaxK];
    __shared__ unsigned long long s_top[kMaxK];
    const std::uint32_t lane = threadIdx.x;
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
        if (lane < c) s_ex[lane] = my_id;
        for (std::uint32_t i = lane; i < m; i += 64) s_key[i] = cand[std::size_t(v) * cap + i];
        __syncthreads();
        for (std::uint32_t i = lane; i < m; i += 64) {
            const unsigned long long key = s_key[i];
            const std::uint32_t cid = static_cast<std::uint32_t>(key);
            const float cd = __uint_as_float(static_cast<std::uint32_t>(key >> 32));
            bool ok = cid != v && cid < n && cd < last;
            for (std::uint32_t j = 0; j < c && ok; ++j) ok = s_ex[j] != cid;
            for (std::uint32_t j = 0; j < i && ok; ++j) ok = s_key[j] != key;
            s_ok[i] = ok ? 1 : 0;
        }
        __syncthreads();
        std::uint32_t nvalid = 0;
        for (std::uint32_t i = lane; i < m; i += 64) {
            if (!s_ok[i]) continue;
            const unsigned long long key = s_key[i];
            std::uint32_t rank = 0;
            for (std::uint32_t j = 0; j < m; ++j) rank += (s_ok[j] && s_key[j] < key) ? 1u : 0u;
            if (rank < k) s_top[rank] = key;
            ++nvalid;
        }
        nvalid = wave_sum(nvalid);
        __syncthreads();
        const std::uint32_t t = nvalid < k ? nvalid : k;  // candidates that can enter
        const bool has_c = lane < t;
        const unsigned long long key = has_c ? s_top[lane] : ~0ull;
        const std::uint32_t cid = static_cast<std::uint32_t>(key);
        const float cd = __uint_as_float(static_cast<std::uint32_t>(key >> 32));
        // merge: existing entries precede candidates of equal distance
        std::uint32_t pos_e = lane, pos_c = lane;
        for (std::uint32_t j = 0; j < t; ++j) {
            const float od = __shfl(cd, j, 64);
            pos_e += (od < my_d) ? 1u : 0u;
        }
        for (std::uint32_t j = 0; j < c; ++j) {
            const float od = __shfl(my_d, j, 64);
            pos_c += (od <= cd) ? 1u : 0u;
        }
        if (lane < c && pos_e < k) {
            ids[base + pos_e] = my_id;
            dist[base + pos_e] = my_d;
            flags[base + pos_e] = my_f;
        }
        const bool ins = has_c && pos_c < k;
        if (ins) {
            ids[base + pos_c] = cid;
            dist[base + pos_c] = cd;
            flags[base + pos_c] = 1;
        }
        if (lane == 0) {
            count[v] = (c + t < k) ? c + t : k;
            fill[v] = 0;
        }
        const unsigned long long inserted = __popcll(__ballot(ins));
        if (lane == 0) local_upd += inserted;
        __syncthreads();
    }
    if (lane == 0 && local_upd) atomicAdd(updates, local_upd);
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

struct DescentBuffers {
    DBuf<std::uint32_t> new_list, new_cnt, old_list, old_cnt;
    DBuf<std::uint32_t> rev_new, rev_new_cnt, rev_old, rev_old_cnt;
    DBuf<unsigned long long> cand, updates;
    DBuf<std::uint32_t> fill;
};

struct DescentResult {
    std::uint32_t rounds = 0;
    std::uint64_t updates = 0;
};

DescentResult run_descent(Graph* gp, const float* pts, std::uint32_t n, std::uint32_t stride,
                          std::uint32_t k, std::uint32_t* ids, float* dist, std::uint8_t* flags,
                          std::uint32_t* count, std::uint32_t smp, std::uint32_t iters,
                          double delta, std::uint64_t seed, bool random_fill,
                          std::uint32_t keep, std::uint32_t cap, hipStream_t s) {
    DescentResult res;
    if (n == 0) return res;
    if (k > kMaxK) throw std::runtime_error("nndescent: k above 64 is not supported on device");
    if (smp == 0) smp = 1;
    if (smp > kMaxSample) smp = kMaxSample;
    const unsigned wave_blocks = capped_blocks(std::size_t(n) * 64, 256, 16384);
    if (random_fill && n >= 2) {
        init_random_kernel<<<wave_blocks, 256, 0, s>>>(pts, n, stride, k, seed, ids, dist, flags,
                                                       count);
        CSG_LAUNCH_CHECK();
    }
    if (iters == 0) return res;
    DescentBuffers b;
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
    b.updates.reserve(1);
    fill_u32(b.fill.p, n, 0u, s);

    const std::uint32_t S = 6u * smp, A = 2u * smp;
    const std::size_t lds = (2 * std::size_t(S) + std::size_t(S) * stride + std::size_t(A) * S) *
                            sizeof(float);
    if (lds > 150 * 1024) throw std::runtime_error("nndescent: candidate tile exceeds LDS");
    const void* join_fn = keep <= 8    ? reinterpret_cast<const void*>(join_kernel<8>)
                          : keep <= 16 ? reinterpret_cast<const void*>(join_kernel<16>)
                                       : reinterpret_cast<const void*>(join_kernel<32>);
    if (lds > 64 * 1024)
        CSG_CHECK(hipFuncSetAttribute(join_fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(lds)));
    JoinArgs ja{pts, n, stride, k, smp, dist, b.new_list.p, b.new_cnt.p, b.old_list.p,
                b.old_cnt.p, b.rev_new.p, b.rev_new_cnt.p, b.rev_old.p, b.rev_old_cnt.p,
                b.cand.p, b.fill.p, cap, keep};
    const unsigned join_blocks = static_cast<unsigned>(std::min<std::size_t>(n, 65536));
    const double stop = delta * double(k) * double(n);
    for (std::uint32_t round = 0; round < iters; ++round) {
        fill_u32(b.rev_new_cnt.p, n, 0u, s);
        fill_u32(b.rev_old_cnt.p, n, 0u, s);
        fill_u64(reinterpret_cast<std::uint64_t*>(b.updates.p), 1, 0ull, s);
        {
            std::unique_ptr<ScopedKernelTimer> t;
            if (gp) t.reset(new ScopedKernelTimer(*gp, "nnd_sample"));
            sample_kernel<<<wave_blocks, 256, 0, s>>>(
                n, k, ids, flags, count, smp, b.new_list.p, b.new_cnt.p, b.old_list.p,
                b.old_cnt.p, b.rev_new.p, b.rev_new_cnt.p, b.rev_old.p, b.rev_old_cnt.p);
            CSG_LAUNCH_CHECK();
        }
        {
            std::unique_ptr<ScopedKernelTimer> t;
            if (gp) t.reset(new ScopedKernelTimer(*gp, "nnd_join"));
            if (keep <= 8)
                join_kernel<8><<<join_blocks, kJoinThreads, lds, s>>>(ja);
            else if (keep <= 16)
                join_kernel<16><<<join_blocks, kJoinThreads, lds, s>>>(ja);
            else
                join_kernel<32><<<join_blocks, kJoinThreads, lds, s>>>(ja);
            CSG_LAUNCH_CHECK();
        }
        {
            std::unique_ptr<ScopedKernelTimer> t;
            if (gp) t.reset(new ScopedKernelTimer(*gp, "nnd_merge"));
            merge_kernel<<<static_cast<unsigned>(std::min<std::size_t>(n, 65536)), 64, 0, s>>>(
                n, k, ids, dist, flags, count, b.cand.p, b.fill.p, cap, b.updates.p);
            CSG_LAUNCH_CHECK();
        }
        unsigned long long upd = 0;
        CSG_CHECK(hipMemcpyAsync(&upd, b.updates.p, 8, hipMemcpyDeviceToHost, s));
        CSG_CHECK(hipStreamSynchronize(s));
        res.updates += upd;
        res.rounds = round + 1;
        if (double(upd) <= stop) break;
    }
    return res;
}

// Proposals per member and node: the list length, within [8, 32]. Fewer leave
// the descent at a worse fixed point (pairs that become old-old are never
// compared again).
std::uint32_t default_keep(std::uint32_t k) { return k < 8 ? 8 : (k > 32 ? 32 : k); }

void check_join_params(std::uint32_t keep, std::uint32_t cap) {
    if (keep == 0 || keep > 32) throw std::runtime_error("nndescent: join_keep must be in 1..32");
    if (cap < 16 || cap > kMaxCap)
        throw std::runtime_error("nndescent: join_bucket must be in 16..256");
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
    if (p.brute_force_limit && n <= p.brute_force_limit) {
        brute_force_kernel<<<capped_blocks(n, 64), 64, 0, s>>>(g.pts.p, n, g.stride, k, g.ids.p,
                                                               g.dist.p, g.flags.p, g.count.p);
        CSG_LAUNCH_CHECK();
    } else {
        const std::uint32_t keep = p.join_keep ? p.join_keep : default_keep(k);
        const std::uint32_t cap = p.join_bucket ? p.join_bucket : kMaxCap;
        check_join_params(keep, cap);
        r = run_descent(&g, g.pts.p, n, g.stride, k, g.ids.p, g.dist.p, g.flags.p, g.count.p,
                        p.sample, p.iters, p.delta, p.seed, true, keep, cap, s);
    }
    CSG_CHECK(hipStreamSynchronize(s));
    g.knn_set = true;
    g.edges_ready = false;
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
    if (cfg.metric != CORESG_HIP_METRIC_L2)
        throw std::runtime_error("nndescent: unsupported metric for the device path (L2 only)");
    if (cfg.stride % 8 != 0 || cfg.stride < cfg.dim)
        throw std::runtime_error("nndescent: stride must be round_up(dim, 8)");
    hipStream_t s;
    CSG_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
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
                    default_keep(cfg.k), kMaxCap, s);
    CSG_CHECK(hipMemcpyAsync(ids, d_ids.p, nk * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipMemcpyAsync(dists, d_dist.p, nk * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipMemcpyAsync(flags, d_flags.p, nk, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipMemcpyAsync(counts, d_cnt.p, n * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));
    CSG_CHECK(hipStreamDestroy(s));
    if (updates) *updates = r.updates;
}

}  // namespace coresg_hip
