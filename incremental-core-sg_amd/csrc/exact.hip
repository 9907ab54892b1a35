// The exact MST of the complete mutual-reachability graph over the resident
// points: CoreSG::exact_mst (coresg.cpp:419-430), i.e. mst_prim_virtual
// (mst.cpp:172-220) whose row of vertex i is CpuBackend::mrd_block's row
// (cpu_backend.cpp:61-90): max(d(i, j), max(core[i], core[j])) with the exact
// FP32 distance.
//
// Prim is sequential in its picks, and its tie rules fix which tree comes out
// when weights tie (MRD weights tie often: every edge inside a dense region
// takes the larger core distance). They are kept exactly:
//   * best[j] improves only on a strictly smaller weight, so from[j] is the
//     earliest tree vertex reaching the minimum;
//   * the pick is the smallest j among the not-yet-tree vertices of minimum
//     best[j], i.e. the minimum of (best bits << 32 | j) over them (weights
//     are non-negative, so their bit patterns order like the floats).
//
// One cooperative launch runs every step. Block b owns vertices
// [b*chunk, (b+1)*chunk): their best/from/in-tree state is touched by no other
// block. A step is: the new tree vertex's row is read into LDS, each owned
// vertex is relaxed against it, the block's minimum key goes to a candidate
// slot, one grid barrier, and then every block reduces the same candidate
// array to the same pick. Candidate arrays alternate by step parity, so the
// next step's writes never meet a block still reading this step's. The owner
// of the pick records the step's edge. The barrier counter only grows
// (target = step * blocks), and every wait is bounded: a wait that exceeds
// the bound sets an error flag and ends its block, after which every other
// block ends at its next wait, so the launch always drains.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/coresg_hip.h"
#include "common.hpp"
#include "graph.hpp"

namespace coresg_hip {
namespace {

constexpr unsigned kThreads = 256;
constexpr std::uint32_t kMaxStride = 1024;  // row of the current vertex, kept in LDS
constexpr std::uint32_t kSpinLimit = 1u << 22;
constexpr std::uint32_t kErrBarrier = 1u;

struct Prim {
    const float* pts;
    const float* core;
    float* best;
    std::uint32_t* from;
    std::uint8_t* in_tree;
    std::uint64_t* cand;  // 2 * blocks
    std::uint32_t* out_a;  // per step: smaller endpoint, kNone when the step adds no edge
    std::uint32_t* out_b;
    float* out_w;
    std::uint32_t* bar;
    std::uint32_t* err;
    std::uint32_t n, stride, chunk;
};

// core_dist (knn_graph.hpp:157-163) of every vertex for min(mpts, k)
// (coresg.cpp:68-70): +inf for an empty list, else the entry at position
// min(mpts, count) - 1, mpts 0 reading the first.
__global__ void __launch_bounds__(kThreads)
    exact_core_kernel(const float* __restrict__ dist, const std::uint32_t* __restrict__ count,
                      std::uint32_t n, std::uint32_t k, std::uint32_t mpts,
                      float* __restrict__ core) {
    for (std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += gridDim.x * blockDim.x) {
        const std::uint32_t c = count[v];
        std::uint32_t pos = mpts == 0 ? 0 : mpts - 1;
        if (pos >= c) pos = c - 1;
        core[v] = c == 0 ? INFINITY : dist[std::size_t(v) * k + pos];
    }
}

// Bounded grid barrier; returns false (for every thread of the block) when
// the wait gave up.
__device__ __forceinline__ bool grid_wait(const Prim& p, std::uint32_t target, bool* s_ok) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(p.bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        bool ok = true;
        std::uint32_t it = 0;
        while (__hip_atomic_load(p.bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++it > kSpinLimit) {
                atomicOr(p.err, kErrBarrier);
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok;
}

__global__ void __launch_bounds__(kThreads) exact_prim_kernel(Prim p) {
    __shared__ __align__(16) float crow[kMaxStride];
    __shared__ std::uint64_t wmin[kThreads / kWave];
    __shared__ bool s_ok;
    const std::uint32_t tid = threadIdx.x, G = gridDim.x;
    const std::uint32_t j0 = blockIdx.x * p.chunk;
    const std::uint32_t j1 = j0 + p.chunk < p.n ? j0 + p.chunk : p.n;
    for (std::uint32_t j = j0 + tid; j < j1; j += kThreads) {
        p.best[j] = j == 0 ? 0.0f : INFINITY;
        p.from[j] = kNone;
        p.in_tree[j] = j == 0;
    }
    std::uint32_t current = 0;
    for (std::uint32_t step = 1; step < p.n; ++step) {
        const float* row = p.pts + std::size_t(current) * p.stride;
        for (std::uint32_t t = tid; t < p.stride; t += kThreads) crow[t] = row[t];
        const float cc = p.core[current];
        __syncthreads();
        std::uint64_t key_min = kNone64;
        for (std::uint32_t j = j0 + tid; j < j1; j += kThreads) {
            if (p.in_tree[j]) continue;
            const float d = sqrt_exact(l2sqr_exact(crow, p.pts + std::size_t(j) * p.stride, p.stride));
            const float w = fmax_ref(d, fmax_ref(cc, p.core[j]));
            float b = p.best[j];
            if (w < b) {
                b = w;
                p.best[j] = w;
                p.from[j] = current;
            }
            const std::uint64_t key = (std::uint64_t(fbits(b)) << 32) | j;
            key_min = key < key_min ? key : key_min;
        }
        key_min = wave_min_u64(key_min);
        if ((tid & (kWave - 1)) == 0) wmin[tid / kWave] = key_min;
        __syncthreads();
        std::uint64_t* cand = p.cand + std::size_t(step & 1u) * G;
        if (tid == 0) {
            std::uint64_t m = wmin[0];
#pragma unroll
            for (unsigned w = 1; w < kThreads / kWave; ++w) m = wmin[w] < m ? wmin[w] : m;
            cand[blockIdx.x] = m;
        }
        if (!grid_wait(p, step * G, &s_ok)) return;
        std::uint64_t k = kNone64;
        for (std::uint32_t i = tid; i < G; i += kThreads) {
            const std::uint64_t c =
                __hip_atomic_load(&cand[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            k = c < k ? c : k;
        }
        k = wave_min_u64(k);
        __syncthreads();  // every wave is done with wmin from the block reduction
        if ((tid & (kWave - 1)) == 0) wmin[tid / kWave] = k;
        __syncthreads();
        k = wmin[0];
#pragma unroll
        for (unsigned w = 1; w < kThreads / kWave; ++w) k = wmin[w] < k ? wmin[w] : k;
        if (k == kNone64) break;  // no vertex left outside the tree
        const std::uint32_t pick = static_cast<std::uint32_t>(k);
        if (tid == 0 && pick >= j0 && pick < j1) {
            const float pw = __uint_as_float(static_cast<std::uint32_t>(k >> 32));
            const std::uint32_t f = p.from[pick];
            p.in_tree[pick] = 1;
            if (f != kNone && pw < INFINITY) {
                p.out_a[step - 1] = f < pick ? f : pick;
                p.out_b[step - 1] = f < pick ? pick : f;
                p.out_w[step - 1] = pw;
            } else {
                p.out_a[step - 1] = kNone;
            }
        }
        current = pick;
        // The next step's first __syncthreads (after the row load) orders the
        // owner's in_tree write before any of its threads reads it.
    }
}

// The same Prim with each block's vertices resident: their rows in LDS, one
// column per vertex (consecutive lanes read consecutive words), their best /
// from / in-tree state in registers (vertex j0 + tid + r * kThreads in slot
// r). A step's only global traffic is the new tree vertex's row, one 64-bit
// candidate word per block and the polls of the other blocks' words.
//
// Candidate word: tag (step & 3) in bits 62-63, then the key best bits << 31 |
// j (weights are non-negative, so their bits fit in 31; j < 2^31). A block
// publishes its word with one relaxed device-scope store into the buffer of
// the step's parity and polls the G words of that buffer until every tag is
// the step's: the words are the barrier and the data at once, with no
// counter and no release/acquire cache maintenance (the points are read-only
// and everything else a step reads is in LDS or registers). A block can run
// at most one step ahead of the slowest -- to finish step s+1 it needs every
// block's word of step s+1 -- so the other buffer still holds step s-1's tags
// while step s is read, and two tag bits tell them apart. Zero-filled buffers
// hold tag 0, which no step polls for in a zero-filled buffer (step 4 polls
// buffer 0 only after step 2 wrote tag 2 there).
constexpr std::uint64_t kPayload = (1ull << 62) - 1;

template <std::uint32_t R>
__global__ void __launch_bounds__(kThreads) exact_prim_lds_kernel(Prim p) {
    extern __shared__ __align__(16) float lds[];
    constexpr std::uint32_t ld = R * kThreads;  // columns: the block's vertices
    float* crow = lds + std::size_t(ld) * p.stride;
    __shared__ std::uint64_t wmin[kThreads / kWave];
    const std::uint32_t tid = threadIdx.x, G = gridDim.x;
    const std::uint32_t j0 = blockIdx.x * p.chunk;
    const std::uint32_t j1 = j0 + p.chunk < p.n ? j0 + p.chunk : p.n;
    float best[R], core[R];
    std::uint32_t from[R];
    bool out[R];  // outside the tree and a real vertex
#pragma unroll
    for (std::uint32_t r = 0; r < R; ++r) {
        const std::uint32_t j = j0 + tid + r * kThreads;
        const bool real = j < j1;
        best[r] = INFINITY;
        from[r] = kNone;
        out[r] = real && j != 0;
        core[r] = real ? p.core[j] : 0.0f;
        for (std::uint32_t i = 0; i < p.stride; ++i)
            lds[std::size_t(i) * ld + tid + r * kThreads] =
                real ? p.pts[std::size_t(j) * p.stride + i] : 0.0f;
    }
    std::uint32_t current = 0;
    for (std::uint32_t step = 1; step < p.n; ++step) {
        const float* row = p.pts + std::size_t(current) * p.stride;
        for (std::uint32_t t = tid; t < p.stride; t += kThreads) crow[t] = row[t];
        const float cc = p.core[current];
        __syncthreads();
        std::uint64_t key_min = kPayload;
#pragma unroll
        for (std::uint32_t r = 0; r < R; ++r) {
            if (!out[r]) continue;
            const float* col = lds + tid + r * kThreads;
            float acc0[8], acc1[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                acc0[q] = 0.0f;
                acc1[q] = 0.0f;
            }
            std::uint32_t i = 0;
            for (; i + 16 <= p.stride; i += 16) {
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const float d0 = crow[i + q] - col[std::size_t(i + q) * ld];
                    const float d1 = crow[i + 8 + q] - col[std::size_t(i + 8 + q) * ld];
                    acc0[q] = __builtin_fmaf(d0, d0, acc0[q]);
                    acc1[q] = __builtin_fmaf(d1, d1, acc1[q]);
                }
            }
            if (i + 8 <= p.stride) {
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const float d0 = crow[i + q] - col[std::size_t(i + q) * ld];
                    acc0[q] = __builtin_fmaf(d0, d0, acc0[q]);
                }
            }
            float v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = acc0[q] + acc1[q];
            const float l0 = v[0] + v[4], l1 = v[1] + v[5], l2 = v[2] + v[6], l3 = v[3] + v[7];
            const float d = sqrt_exact((l0 + l1) + (l2 + l3));
            const float w = fmax_ref(d, fmax_ref(cc, core[r]));
            if (w < best[r]) {
                best[r] = w;
                from[r] = current;
            }
            const std::uint64_t key =
                (std::uint64_t(fbits(best[r])) << 31) | (j0 + tid + r * kThreads);
            key_min = key < key_min ? key : key_min;
        }
        key_min = wave_min_u64(key_min);
        if ((tid & (kWave - 1)) == 0) wmin[tid / kWave] = key_min;
        __syncthreads();
        std::uint64_t* cand = p.cand + std::size_t(step & 1u) * G;
        const std::uint64_t tag = std::uint64_t(step & 3u) << 62;
        if (tid == 0) {
            std::uint64_t m = wmin[0];
#pragma unroll
            for (unsigned q = 1; q < kThreads / kWave; ++q) m = wmin[q] < m ? wmin[q] : m;
            __hip_atomic_store(&cand[blockIdx.x], tag | m, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        // Poll every block's word of this step (G <= kThreads: one per thread).
        std::uint64_t seen = kPayload;
        std::uint32_t it = 0;
        bool ready = tid >= G;
        for (;;) {
            if (!ready) {
                const std::uint64_t c =
                    __hip_atomic_load(&cand[tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((c & ~kPayload) == tag) {
                    ready = true;
                    seen = c & kPayload;
                }
            }
            if (__syncthreads_and(ready)) break;
            if (++it > kSpinLimit) {
                if (tid == 0) atomicOr(p.err, kErrBarrier);
                return;  // block-uniform: every thread saw the same count
            }
            __builtin_amdgcn_s_sleep(1);
        }
        seen = wave_min_u64(seen);
        if ((tid & (kWave - 1)) == 0) wmin[tid / kWave] = seen;
        __syncthreads();
        std::uint64_t k = wmin[0];
#pragma unroll
        for (unsigned q = 1; q < kThreads / kWave; ++q) k = wmin[q] < k ? wmin[q] : k;
        if (k == kPayload) break;  // no vertex left outside the tree
        const std::uint32_t pick = static_cast<std::uint32_t>(k & 0x7FFFFFFFu);
        if (pick >= j0 && pick < j1 && (pick - j0) % kThreads == tid) {
            const std::uint32_t r = (pick - j0) / kThreads;
            const float pw = __uint_as_float(static_cast<std::uint32_t>(k >> 31));
#pragma unroll
            for (std::uint32_t q = 0; q < R; ++q) {
                if (q != r) continue;
                out[q] = false;
                const std::uint32_t f = from[q];
                if (f != kNone && pw < INFINITY) {
                    p.out_a[step - 1] = f < pick ? f : pick;
                    p.out_b[step - 1] = f < pick ? pick : f;
                    p.out_w[step - 1] = pw;
                } else {
                    p.out_a[step - 1] = kNone;
                }
            }
        }
        current = pick;
    }
}

// The same Prim as exact_prim_kernel with one launch per step instead of one
// cooperative launch for all (coresg_hip_graph_set_exact_launch(g, 1)): the
// kernel boundary is the step's barrier, so nothing waits on another block.
// Each block relaxes its owned vertices against the current tree vertex and
// publishes its minimum key; the block that finishes last (a device-scope
// counter) reduces the keys to the pick, records the step's edge, marks the
// pick in the tree and makes it the next step's current vertex. ctl[0] = the
// current vertex, ctl[1] = the finished-block counter, ctl[2] = the step;
// launches past the last step return at once, so a captured batch of
// launches can be replayed over the whole run.
__global__ void __launch_bounds__(kThreads) exact_prim_step_kernel(Prim p, std::uint32_t* ctl) {
    __shared__ __align__(16) float crow[kMaxStride];
    __shared__ std::uint64_t wmin[kThreads / kWave];
    __shared__ bool s_last;
    const std::uint32_t step = ctl[2];
    if (step >= p.n) return;  // block-uniform
    const std::uint32_t tid = threadIdx.x, G = gridDim.x;
    const std::uint32_t j0 = blockIdx.x * p.chunk;
    const std::uint32_t j1 = j0 + p.chunk < p.n ? j0 + p.chunk : p.n;
    const std::uint32_t current = ctl[0];
    std::uint64_t key_min = kNone64;
    if (current != kNone) {
        const float* row = p.pts + std::size_t(current) * p.stride;
        for (std::uint32_t t = tid; t < p.stride; t += kThreads) crow[t] = row[t];
        const float cc = p.core[current];
        __syncthreads();
        for (std::uint32_t j = j0 + tid; j < j1; j += kThreads) {
            if (p.in_tree[j]) continue;
            const float d =
                sqrt_exact(l2sqr_exact(crow, p.pts + std::size_t(j) * p.stride, p.stride));
            const float w = fmax_ref(d, fmax_ref(cc, p.core[j]));
            float b = p.best[j];
            if (w < b) {
                b = w;
                p.best[j] = w;
                __hip_atomic_store(&p.from[j], current, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            const std::uint64_t key = (std::uint64_t(fbits(b)) << 32) | j;
            key_min = key < key_min ? key : key_min;
        }
    }
    key_min = wave_min_u64(key_min);
    if ((tid & (kWave - 1)) == 0) wmin[tid / kWave] = key_min;
    __syncthreads();
    if (tid == 0) {
        std::uint64_t m = wmin[0];
#pragma unroll
        for (unsigned w = 1; w < kThreads / kWave; ++w) m = wmin[w] < m ? wmin[w] : m;
        __hip_atomic_store(&p.cand[blockIdx.x], m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // release: this block's key and from[] stores before the count
        const std::uint32_t done =
            __hip_atomic_fetch_add(&ctl[1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = done == G - 1;
    }
    __syncthreads();
    if (!s_last) return;  // block-uniform
    std::uint64_t k = kNone64;
    for (std::uint32_t i = tid; i < G; i += kThreads) {
        const std::uint64_t c =
            __hip_atomic_load(&p.cand[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        k = c < k ? c : k;
    }
    k = wave_min_u64(k);
    __syncthreads();
    if ((tid & (kWave - 1)) == 0) wmin[tid / kWave] = k;
    __syncthreads();
    if (tid == 0) {
        k = wmin[0];
#pragma unroll
        for (unsigned w = 1; w < kThreads / kWave; ++w) k = wmin[w] < k ? wmin[w] : k;
        std::uint32_t next = kNone;
        if (k != kNone64 && step >= 1) {
            const std::uint32_t pick = static_cast<std::uint32_t>(k);
            const float pw = __uint_as_float(static_cast<std::uint32_t>(k >> 32));
            const std::uint32_t f =
                __hip_atomic_load(&p.from[pick], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            p.in_tree[pick] = 1;
            if (f != kNone && pw < INFINITY) {
                p.out_a[step - 1] = f < pick ? f : pick;
                p.out_b[step - 1] = f < pick ? pick : f;
                p.out_w[step - 1] = pw;
            } else {
                p.out_a[step - 1] = kNone;
            }
            next = pick;
        } else if (step >= 1) {
            p.out_a[step - 1] = kNone;
        }
        ctl[0] = next;
        ctl[1] = 0;
        ctl[2] = step + 1;
    }
}

// Initial state of the stepped Prim: vertex 0 in the tree and current.
__global__ void __launch_bounds__(kThreads) exact_step_init_kernel(Prim p, std::uint32_t* ctl) {
    for (std::uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < p.n;
         j += gridDim.x * blockDim.x) {
        p.best[j] = j == 0 ? 0.0f : INFINITY;
        p.from[j] = kNone;
        p.in_tree[j] = j == 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl[0] = 0;  // current
        ctl[1] = 0;  // finished blocks
        ctl[2] = 1;  // step
    }
}

constexpr std::uint32_t kStepsPerGraph = 256;  // launches per captured replay

using PrimKernel = void (*)(Prim);

// Base distance of each (u, v) pair, metric_distance (distance.hpp:196-201).
__global__ void __launch_bounds__(kThreads)
    edge_base_kernel(const float* __restrict__ pts, std::uint32_t stride,
                     const std::uint32_t* __restrict__ u, const std::uint32_t* __restrict__ v,
                     std::size_t m, float* __restrict__ w) {
    for (std::size_t e = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
         e += std::size_t(gridDim.x) * blockDim.x)
        w[e] = sqrt_exact(l2sqr_exact(pts + std::size_t(u[e]) * stride,
                                      pts + std::size_t(v[e]) * stride, stride));
}

std::uint32_t find_root(std::vector<std::uint32_t>& parent, std::uint32_t x) {
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

}  // namespace

void exact_mst(Graph& g, std::uint32_t mpts, void* info_out, std::uint32_t* out_u,
               std::uint32_t* out_v, float* out_w, std::uint32_t* out_labels) {
    auto* info = static_cast<coresg_hip_mst_info*>(info_out);
    if (!g.points_set) throw std::runtime_error("exact_mst: points not set");
    if (!g.knn_set) throw std::runtime_error("exact_mst: k-NN lists not set");
    if (g.stride > kMaxStride)
        throw std::runtime_error("exact_mst: row stride above " + std::to_string(kMaxStride));
    const std::uint32_t n = g.n;
    const std::uint32_t m = mpts < g.k ? mpts : g.k;
    std::vector<std::uint32_t> hu, hv;
    std::vector<float> hw;
    if (n >= 2) {
        int cus = 0, coop = 0, per_cu = 0;
        CSG_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g.device));
        CSG_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, g.device));
        if (!coop) throw std::runtime_error("exact_mst: device lacks cooperative launch");
        CSG_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(exact_prim_kernel), kThreads, 0));
        if (per_cu < 1 || cus < 1) throw std::runtime_error("exact_mst: kernel does not fit a CU");
        const std::uint32_t want = (n + kThreads - 1) / kThreads;
        // At most one block per CU (co-resident) and per thread of a block
        // (the resident kernel polls one candidate word per thread).
        const std::uint32_t cap = std::uint32_t(cus) < kThreads ? std::uint32_t(cus) : kThreads;
        const std::uint32_t blocks = want < cap ? want : cap;
        const std::uint32_t chunk = (n + blocks - 1) / blocks;
        if (std::uint64_t(n) * blocks >= (1ull << 32))
            throw std::runtime_error("exact_mst: n too large for the step counter");
        // Resident variant when the block's rows fit in LDS and the keys' index
        // field: R vertices per thread.
        PrimKernel kern = exact_prim_kernel;
        std::size_t shmem = 0;
        const std::uint32_t per = (chunk + kThreads - 1) / kThreads;
        const std::uint32_t R = per <= 1 ? 1 : per <= 2 ? 2 : per <= 4 ? 4 : 0;
        int lds_max = 0;
        CSG_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                        g.device));
        const std::size_t need = (std::size_t(R) * kThreads + 1) * g.stride * sizeof(float);
        if (R != 0 && n < (1u << 31) && need + 4096 <= std::size_t(lds_max)) {
            kern = R == 1 ? exact_prim_lds_kernel<1>
                          : R == 2 ? exact_prim_lds_kernel<2> : exact_prim_lds_kernel<4>;
            shmem = need;
            CSG_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          static_cast<int>(shmem)));
            int fits = 0;
            CSG_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &fits, reinterpret_cast<const void*>(kern), kThreads, shmem));
            if (fits < 1) {
                kern = exact_prim_kernel;
                shmem = 0;
            }
        }

        DBuf<float> core, best, ow;
        DBuf<std::uint32_t> from, oa, ob, ctl;
        DBuf<std::uint8_t> in_tree;
        DBuf<std::uint64_t> cand;
        core.reserve(n);
        best.reserve(n);
        from.reserve(n);
        in_tree.reserve(n);
        oa.reserve(n - 1);
        ob.reserve(n - 1);
        ow.reserve(n - 1);
        cand.reserve(2 * std::size_t(blocks));
        ctl.reserve(2);
        CSG_CHECK(hipMemsetAsync(ctl.p, 0, 2 * sizeof(std::uint32_t), g.stream));
        CSG_CHECK(hipMemsetAsync(cand.p, 0, 2 * std::size_t(blocks) * sizeof(std::uint64_t),
                                 g.stream));
        CSG_CHECK(hipMemsetAsync(oa.p, 0xFF, 4 * std::size_t(n - 1), g.stream));
        exact_core_kernel<<<blocks_for(n, kThreads), kThreads, 0, g.stream>>>(
            g.dist.p, g.count.p, n, g.k, m, core.p);
        CSG_LAUNCH_CHECK();
        Prim p{g.pts.p, core.p, best.p, from.p, in_tree.p, cand.p, oa.p, ob.p, ow.p,
               ctl.p, ctl.p + 1, n, g.stride, chunk};
        void* args[] = {&p};
        if (g.exact_stepped) {
            // one launch per Prim step, kStepsPerGraph of them captured once
            // and replayed until every step has run
            DBuf<std::uint32_t> sctl;
            sctl.reserve(3);
            exact_step_init_kernel<<<blocks_for(n, kThreads), kThreads, 0, g.stream>>>(p, sctl.p);
            CSG_LAUNCH_CHECK();
            if (g.plain) {  // the plain form: every launch on its own, synchronised
                ScopedKernelTimer t(g, "exact_prim");
                for (std::uint32_t step = 1; step < n; ++step) {
                    exact_prim_step_kernel<<<blocks, kThreads, 0, g.stream>>>(p, sctl.p);
                    CSG_LAUNCH_CHECK();
                }
            }
            hipGraph_t graph = nullptr;
            if (!g.plain) {
                CaptureGuard cap(g.stream);
                for (std::uint32_t i = 0; i < kStepsPerGraph; ++i) {
                    exact_prim_step_kernel<<<blocks, kThreads, 0, g.stream>>>(p, sctl.p);
                    CSG_CHECK(hipGetLastError());
                }
                graph = cap.end();
            }
            if (graph) {
                hipGraphExec_t exec = nullptr;
                hipError_t e;
                {
                    std::lock_guard<std::mutex> lk(device_api_mutex());
                    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(graph);
                }
                CSG_CHECK(e);
                hipError_t le = hipSuccess;
                {
                    ScopedKernelTimer t(g, "exact_prim");
                    for (std::uint32_t done = 0; done + 1 < n && le == hipSuccess;
                         done += kStepsPerGraph)
                        le = hipGraphLaunch(exec, g.stream);
                }
                const hipError_t se = hipStreamSynchronize(g.stream);
                {
                    std::lock_guard<std::mutex> lk(device_api_mutex());
                    (void)hipGraphExecDestroy(exec);
                }
                CSG_CHECK(le);
                CSG_CHECK(se);
            }
        } else {
            // A cooperative launch cannot be captured; say so instead of
            // launching into a stream that a failed capture left open.
            hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
            CSG_CHECK(hipStreamIsCapturing(g.stream, &cap));
            if (cap != hipStreamCaptureStatusNone)
                throw std::runtime_error("exact_mst: the graph's stream is capturing (status " +
                                         std::to_string(int(cap)) + ")");
            ScopedKernelTimer t(g, "exact_prim");
            const hipError_t e = hipLaunchCooperativeKernel(
                reinterpret_cast<const void*>(kern), dim3(blocks), dim3(kThreads), args,
                static_cast<unsigned>(shmem), g.stream);
            if (e != hipSuccess) {
                hipStreamCaptureStatus after = hipStreamCaptureStatusNone;
                const hipError_t qe = hipStreamIsCapturing(g.stream, &after);
                throw HipError(std::string("hipLaunchCooperativeKernel(exact_prim, ") +
                                   std::to_string(blocks) + " blocks, " + std::to_string(shmem) +
                                   " B LDS) -> " + hipGetErrorName(e) + ": " +
                                   hipGetErrorString(e) + " [stream capture status " +
                                   (qe == hipSuccess ? std::to_string(int(after))
                                                     : std::string(hipGetErrorName(qe))) +
                                   "]",
                               e);
            }
        }
        std::uint32_t err = 0;
        std::vector<std::uint32_t> a(n - 1), b(n - 1);
        std::vector<float> w(n - 1);
        CSG_CHECK(hipMemcpyAsync(&err, ctl.p + 1, 4, hipMemcpyDeviceToHost, g.stream));
        CSG_CHECK(hipMemcpyAsync(a.data(), oa.p, 4 * (n - 1), hipMemcpyDeviceToHost, g.stream));
        CSG_CHECK(hipMemcpyAsync(b.data(), ob.p, 4 * (n - 1), hipMemcpyDeviceToHost, g.stream));
        CSG_CHECK(hipMemcpyAsync(w.data(), ow.p, 4 * (n - 1), hipMemcpyDeviceToHost, g.stream));
        CSG_CHECK(hipStreamSynchronize(g.stream));
        if (err) throw std::runtime_error("exact_mst: grid barrier wait gave up");
        for (std::uint32_t s = 0; s + 1 < n; ++s) {
            if (a[s] == kNone) continue;
            hu.push_back(a[s]);
            hv.push_back(b[s]);
            hw.push_back(w[s]);
        }
    }
    // Labels by first appearance over a union-find of the accepted edges and
    // the weight as the double sum in acceptance order (mst.cpp:43-59;
    // types.hpp:58-62).
    std::vector<std::uint32_t> parent(n);
    for (std::uint32_t i = 0; i < n; ++i) parent[i] = i;
    std::uint32_t comps = n;
    double weight = 0.0;
    for (std::size_t e = 0; e < hu.size(); ++e) {
        const std::uint32_t ra = find_root(parent, hu[e]), rb = find_root(parent, hv[e]);
        if (ra != rb) {
            parent[ra < rb ? rb : ra] = ra < rb ? ra : rb;
            --comps;
        }
        weight += double(hw[e]);
    }
    if (out_labels) {
        std::vector<std::uint32_t> remap(n, kNone);
        std::uint32_t next = 0;
        for (std::uint32_t i = 0; i < n; ++i) {
            const std::uint32_t r = find_root(parent, i);
            if (remap[r] == kNone) remap[r] = next++;
            out_labels[i] = remap[r];
        }
    }
    if (out_u) std::memcpy(out_u, hu.data(), hu.size() * 4);
    if (out_v) std::memcpy(out_v, hv.data(), hv.size() * 4);
    if (out_w) std::memcpy(out_w, hw.data(), hw.size() * 4);
    if (info) {
        info->mpts = mpts;
        info->num_edges = static_cast<std::uint32_t>(hu.size());
        info->num_components = comps;
        info->rounds = n >= 2 ? n - 1 : 0;
        info->weight = weight;
    }
}

// MissingEdgeMode::ExactMst bridges (coresg.cpp:157-163): the exact MST at
// k_max, each edge re-stored with its base distance.
void exact_mst_bridges(Graph& g, std::vector<std::uint32_t>* bu, std::vector<std::uint32_t>* bv,
                       std::vector<float>* bw) {
    const std::size_t cap = g.n >= 2 ? g.n - 1 : 1;
    bu->assign(cap, 0);
    bv->assign(cap, 0);
    bw->assign(cap, 0.0f);
    coresg_hip_mst_info info{};
    exact_mst(g, g.k, &info, bu->data(), bv->data(), bw->data(), nullptr);
    const std::size_t m = info.num_edges;
    bu->resize(m);
    bv->resize(m);
    bw->resize(m);
    if (m == 0) return;
    DBuf<std::uint32_t> du, dv;
    DBuf<float> dw;
    du.reserve(m);
    dv.reserve(m);
    dw.reserve(m);
    CSG_CHECK(hipMemcpyAsync(du.p, bu->data(), 4 * m, hipMemcpyHostToDevice, g.stream));
    CSG_CHECK(hipMemcpyAsync(dv.p, bv->data(), 4 * m, hipMemcpyHostToDevice, g.stream));
    edge_base_kernel<<<capped_blocks(m, kThreads), kThreads, 0, g.stream>>>(g.pts.p, g.stride, du.p,
                                                                           dv.p, m, dw.p);
    CSG_LAUNCH_CHECK();
    CSG_CHECK(hipMemcpyAsync(bw->data(), dw.p, 4 * m, hipMemcpyDeviceToHost, g.stream));
    CSG_CHECK(hipStreamSynchronize(g.stream));
}

}  // namespace coresg_hip
