// Dense mutual-reachability blocks and the bridge MST.
//
//   * mrd_block (reference ABI cuda_api.hpp:68-70; CPU semantics
//     src/cpu_backend.cpp:61-90): rows of the m x m MRD matrix with the exact
//     host distance (reference summation order, IEEE sqrt), 0 on the diagonal,
//     max(d, max(c_i, c_j)) elsewhere.
//   * mst_over_subset (src/coresg.cpp:97-141): gather the subset, core at
//     k_max, dense MRD matrix, Prim from node 0 with strict improvement and the
//     first index on ties (src/mst.cpp:172-220), edges re-stored with the base
//     distance in global ids.
//   * mrd_row_min (cuda_api.hpp:74-78): per-row minimum across components.
#include <cstring>
#include <mutex>
#include <vector>

#include "build.hpp"
#include "hubness.hpp"
#include "scan.hpp"

namespace coresg_hip {
namespace {

constexpr int kTile = 16;

// One 16x16 tile of the matrix per 256-thread block; both row blocks staged
// in LDS (stride <= kMaxLdsStride floats each), exact per-pair distance.
constexpr std::uint32_t kMaxLdsStride = 1024;

__global__ void __launch_bounds__(256)
    mrd_tile_kernel(const float* __restrict__ pts, std::uint32_t m, std::uint32_t stride,
                    const float* __restrict__ core, std::uint32_t r0, std::uint32_t rows,
                    float* __restrict__ out, bool cosine) {
    extern __shared__ float lds[];
    float* A = lds;                    // kTile rows
    float* B = lds + kTile * stride;   // kTile rows
    const int tx = threadIdx.x & (kTile - 1), ty = threadIdx.x / kTile;
    const std::uint32_t row_base = blockIdx.y * kTile, col_base = blockIdx.x * kTile;
    for (std::uint32_t e = threadIdx.x; e < kTile * stride; e += 256) {
        const std::uint32_t rr = e / stride, dd = e - rr * stride;
        const std::uint32_t gi = r0 + row_base + rr;
        A[e] = (row_base + rr < rows && gi < m) ? pts[std::size_t(gi) * stride + dd] : 0.0f;
        const std::uint32_t gj = col_base + rr;
        B[e] = (gj < m) ? pts[std::size_t(gj) * stride + dd] : 0.0f;
    }
    __syncthreads();
    const std::uint32_t lr = row_base + ty, gc = col_base + tx;
    if (lr >= rows || gc >= m) return;
    const std::uint32_t gr = r0 + lr;
    float value = 0.0f;
    if (gr != gc) {
        const float d = metric_finalize(l2sqr_exact(A + ty * stride, B + tx * stride, stride), cosine);
        value = fmax_ref(d, fmax_ref(core[gr], core[gc]));
    }
    out[std::size_t(lr) * m + gc] = value;
}

__global__ void __launch_bounds__(256)
    mrd_plain_kernel(const float* __restrict__ pts, std::uint32_t m, std::uint32_t stride,
                     const float* __restrict__ core, std::uint32_t r0, std::uint32_t rows,
                     float* __restrict__ out, bool cosine) {
    const std::size_t total = std::size_t(rows) * m;
    for (std::size_t t = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
         t += std::size_t(gridDim.x) * blockDim.x) {
        const std::uint32_t lr = static_cast<std::uint32_t>(t / m);
        const std::uint32_t gc = static_cast<std::uint32_t>(t - std::size_t(lr) * m);
        const std::uint32_t gr = r0 + lr;
        float value = 0.0f;
        if (gr != gc) {
            const float d = metric_finalize(l2sqr_exact(pts + std::size_t(gr) * stride,
                                                        pts + std::size_t(gc) * stride, stride),
                                            cosine);
            value = fmax_ref(d, fmax_ref(core[gr], core[gc]));
        }
        out[t] = value;
    }
}

}  // namespace

// Rows [r0, r0 + rows) of the m x m MRD matrix of resident operands, on s.
void launch_mrd(const float* pts, std::uint32_t m, std::uint32_t stride, const float* core,
                std::uint32_t r0, std::uint32_t rows, float* out, hipStream_t s, bool cosine) {
    if (rows == 0 || m == 0) return;
    if (stride <= kMaxLdsStride) {
        dim3 grid((m + kTile - 1) / kTile, (rows + kTile - 1) / kTile);
        const std::size_t lds = std::size_t(2) * kTile * stride * sizeof(float);
        mrd_tile_kernel<<<grid, 256, lds, s>>>(pts, m, stride, core, r0, rows, out, cosine);
    } else {
        mrd_plain_kernel<<<capped_blocks(std::size_t(rows) * m, 256), 256, 0, s>>>(
            pts, m, stride, core, r0, rows, out, cosine);
    }
    CSG_LAUNCH_CHECK();
}

namespace {

// Argmin of (value, index) across a 256-thread block; ties -> lower index.
__device__ __forceinline__ void block_argmin(float* v, std::uint32_t* i, float* sv,
                                             std::uint32_t* si) {
    for (int off = 32; off > 0; off >>= 1) {
        const float ov = __shfl_xor(*v, off, 64);
        const std::uint32_t oi = __shfl_xor(*i, off, 64);
        if (ov < *v || (ov == *v && oi < *i)) {
            *v = ov;
            *i = oi;
        }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sv[w] = *v;
        si[w] = *i;
    }
    __syncthreads();
    float bv = sv[0];
    std::uint32_t bi = si[0];
    for (int t = 1; t < 4; ++t) {
        if (sv[t] < bv || (sv[t] == bv && si[t] < bi)) {
            bv = sv[t];
            bi = si[t];
        }
    }
    __syncthreads();
    *v = bv;
    *i = bi;
}

// Prim over the dense matrix, one 256-thread block. best/from/in_tree live in
// global scratch of m entries each. Writes accepted edges (local ids,
// canonical u < v, MRD weight) in acceptance order; out_count[0] = count.
__device__ void prim_run(const float* __restrict__ mat, std::uint32_t m, float* __restrict__ best,
                         std::uint32_t* __restrict__ from, std::uint8_t* __restrict__ in_tree,
                         std::uint32_t* __restrict__ ou, std::uint32_t* __restrict__ ov,
                         float* __restrict__ ow, std::uint32_t* __restrict__ out_count) {
    __shared__ float sv[4];
    __shared__ std::uint32_t si[4];
    for (std::uint32_t j = threadIdx.x; j < m; j += 256) {
        best[j] = __builtin_inff();
        from[j] = kNone;
        in_tree[j] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        in_tree[0] = 1;
        best[0] = 0.0f;
    }
    __syncthreads();
    std::uint32_t current = 0, t = 0;
    for (std::uint32_t step = 1; step < m; ++step) {
        const float* row = mat + std::size_t(current) * m;
        float lv = __builtin_inff();
        std::uint32_t li = kNone;
        for (std::uint32_t j = threadIdx.x; j < m; j += 256) {
            if (in_tree[j]) continue;
            const float w = row[j];
            float b = best[j];
            if (w < b) {
                b = w;
                best[j] = w;
                from[j] = current;
            }
            if (b < lv || (b == lv && j < li)) {
                lv = b;
                li = j;
            }
        }
        block_argmin(&lv, &li, sv, si);
        if (li == kNone) break;
        if (threadIdx.x == 0) {
            in_tree[li] = 1;
            const std::uint32_t f = from[li];
            if (f != kNone && lv < __builtin_inff()) {
                ou[t] = f < li ? f : li;
                ov[t] = f < li ? li : f;
                ow[t] = lv;
            }
        }
        if (from[li] != kNone && lv < __builtin_inff()) ++t;
        __syncthreads();
        current = li;
    }
    if (threadIdx.x == 0) *out_count = t;
}

__global__ void __launch_bounds__(256)
    prim_kernel(const float* __restrict__ mat, std::uint32_t m, float* __restrict__ best,
                std::uint32_t* __restrict__ from, std::uint8_t* __restrict__ in_tree,
                std::uint32_t* __restrict__ ou, std::uint32_t* __restrict__ ov,
                float* __restrict__ ow, std::uint32_t* __restrict__ out_count) {
    prim_run(mat, m, best, from, in_tree, ou, ov, ow, out_count);
}

// Several subsets at once: subset i's points are rows [so, so + m) of the
// gathered arrays, its matrix sits at mat_off, its tiles at tile_off.
struct SubsetMeta {
    std::uint32_t so, m, tile_off, tiles_x;
    std::uint64_t mat_off;
};

__device__ __forceinline__ std::uint32_t subset_of_tile(const SubsetMeta* meta, std::uint32_t nsets,
                                                        std::uint32_t tile) {
    std::uint32_t lo = 0, hi = nsets;  // largest i with meta[i].tile_off <= tile
    while (hi - lo > 1) {
        const std::uint32_t mid = (lo + hi) / 2;
        if (meta[mid].tile_off <= tile) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(256)
    mrd_tiles_batched_kernel(const float* __restrict__ sub, std::uint32_t stride,
                             const float* __restrict__ core, const SubsetMeta* __restrict__ meta,
                             std::uint32_t nsets, float* __restrict__ mat) {
    extern __shared__ float lds[];
    float* A = lds;
    float* B = lds + kTile * stride;
    const SubsetMeta M = meta[subset_of_tile(meta, nsets, blockIdx.x)];
    const std::uint32_t t = blockIdx.x - M.tile_off;
    const std::uint32_t row_base = (t / M.tiles_x) * kTile, col_base = (t % M.tiles_x) * kTile;
    const float* pts = sub + std::size_t(M.so) * stride;
    const float* cr = core + M.so;
    const int tx = threadIdx.x & (kTile - 1), ty = threadIdx.x / kTile;
    for (std::uint32_t e = threadIdx.x; e < kTile * stride; e += 256) {
        const std::uint32_t rr = e / stride, dd = e - rr * stride;
        const std::uint32_t gi = row_base + rr, gj = col_base + rr;
        A[e] = gi < M.m ? pts[std::size_t(gi) * stride + dd] : 0.0f;
        B[e] = gj < M.m ? pts[std::size_t(gj) * stride + dd] : 0.0f;
    }
    __syncthreads();
    const std::uint32_t gr = row_base + ty, gc = col_base + tx;
    if (gr >= M.m || gc >= M.m) return;
    float value = 0.0f;
    if (gr != gc) {
        const float d = sqrt_exact(l2sqr_exact(A + ty * stride, B + tx * stride, stride));
        value = fmax_ref(d, fmax_ref(cr[gr], cr[gc]));
    }
    mat[M.mat_off + std::size_t(gr) * M.m + gc] = value;
}

__global__ void __launch_bounds__(256)
    prim_batched_kernel(const float* __restrict__ mat, const SubsetMeta* __restrict__ meta,
                        float* __restrict__ best, std::uint32_t* __restrict__ from,
                        std::uint8_t* __restrict__ in_tree, std::uint32_t* __restrict__ ou,
                        std::uint32_t* __restrict__ ov, float* __restrict__ ow,
                        std::uint32_t* __restrict__ counts) {
    const SubsetMeta M = meta[blockIdx.x];
    if (M.m < 2) {
        if (threadIdx.x == 0) counts[blockIdx.x] = 0;
        return;
    }
    prim_run(mat + M.mat_off, M.m, best + M.so, from + M.so, in_tree + M.so, ou + M.so,
             ov + M.so, ow + M.so, counts + blockIdx.x);
}

// Base weights and global ids of every subset's accepted edges (a block each).
__global__ void __launch_bounds__(256)
    base_weight_batched_kernel(const float* __restrict__ sub, std::uint32_t stride,
                               const std::uint32_t* __restrict__ ids,
                               const SubsetMeta* __restrict__ meta,
                               const std::uint32_t* __restrict__ counts,
                               std::uint32_t* __restrict__ ou, std::uint32_t* __restrict__ ov,
                               float* __restrict__ ow) {
    const SubsetMeta M = meta[blockIdx.x];
    const std::uint32_t c = counts[blockIdx.x];
    for (std::uint32_t t = threadIdx.x; t < c; t += blockDim.x) {
        const std::uint32_t e = M.so + t;
        const std::uint32_t a = M.so + ou[e], b = M.so + ov[e];
        ow[e] = sqrt_exact(l2sqr_exact(sub + std::size_t(a) * stride,
                                       sub + std::size_t(b) * stride, stride));
        const std::uint32_t ga = ids[a], gb = ids[b];
        ou[e] = ga < gb ? ga : gb;
        ov[e] = ga < gb ? gb : ga;
    }
}

__global__ void subset_gather_kernel(const float* __restrict__ pts, std::uint32_t stride,
                                     const std::uint32_t* __restrict__ ids, std::uint32_t m,
                                     const float* __restrict__ dist,
                                     const std::uint32_t* __restrict__ count, std::uint32_t k,
                                     float* __restrict__ sub, float* __restrict__ core) {
    for (std::size_t t = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
         t < std::size_t(m) * stride; t += std::size_t(gridDim.x) * blockDim.x) {
        const std::uint32_t i = static_cast<std::uint32_t>(t / stride);
        const std::uint32_t d = static_cast<std::uint32_t>(t - std::size_t(i) * stride);
        sub[t] = pts[std::size_t(ids[i]) * stride + d];
        if (d == 0) {
            // core_dist(ids[i], k_max) (knn_graph.hpp:157-163)
            const std::uint32_t c = count[ids[i]];
            core[i] = c ? dist[std::size_t(ids[i]) * k + (k - 1 < c - 1 ? k - 1 : c - 1)]
                        : __builtin_inff();
        }
    }
}

__global__ void base_weight_kernel(const float* __restrict__ sub, std::uint32_t stride,
                                   const std::uint32_t* __restrict__ ids,
                                   const std::uint32_t* __restrict__ count,
                                   std::uint32_t* __restrict__ ou, std::uint32_t* __restrict__ ov,
                                   float* __restrict__ ow) {
    const std::uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *count) return;
    const std::uint32_t a = ou[t], b = ov[t];
    ow[t] = sqrt_exact(l2sqr_exact(sub + std::size_t(a) * stride, sub + std::size_t(b) * stride,
                                   stride));
    const std::uint32_t ga = ids[a], gb = ids[b];
    ou[t] = ga < gb ? ga : gb;
    ov[t] = ga < gb ? gb : ga;
}

__global__ void __launch_bounds__(256)
    row_min_kernel(const float* __restrict__ pts, std::uint32_t m, std::uint32_t stride,
                   const float* __restrict__ core, const std::uint32_t* __restrict__ comp,
                   std::uint32_t r0, std::uint32_t rows, float* __restrict__ out_min,
                   std::uint32_t* __restrict__ out_arg, bool cosine) {
    __shared__ float sv[4];
    __shared__ std::uint32_t si[4];
    for (std::uint32_t lr = blockIdx.x; lr < rows; lr += gridDim.x) {
        const std::uint32_t gr = r0 + lr;
        const float* a = pts + std::size_t(gr) * stride;
        const std::uint32_t my = comp ? comp[gr] : 0u;
        float lv = __builtin_inff();
        std::uint32_t li = kNone;
        for (std::uint32_t j = threadIdx.x; j < m; j += 256) {
            if (j == gr || (comp && comp[j] == my)) continue;
            const float d =
                metric_finalize(l2sqr_exact(a, pts + std::size_t(j) * stride, stride), cosine);
            const float w = fmax_ref(d, fmax_ref(core[gr], core[j]));
            if (w < lv || (w == lv && j < li)) {
                lv = w;
                li = j;
            }
        }
        block_argmin(&lv, &li, sv, si);
        if (threadIdx.x == 0) {
            out_min[lr] = lv;
            out_arg[lr] = li;
        }
        __syncthreads();
    }
}

// Per-thread residency of the MRD operands: Prim over a virtual graph asks for
// one row at a time over the same points, so they stay on the device between
// calls. A host pointer says nothing about what it points at -- a caller that
// frees one vector and allocates the next (fresh core distances per mpts, a new
// gathered subset) often gets the same address back with other contents -- so
// the cache also keeps host copies of what it uploaded: the core distances are
// compared on every call (m floats, far less than the row's m distances) and the
// points whenever a row block starts at row 0, which is how every Prim
// (mst_prim_virtual starts from vertex 0, src/mst.cpp:186-192) and every
// row-split matrix (JobScheduler::mrd_matrix, backend.hpp:149-164) begins.
struct MrdCache {
    int device = -1;
    const float* pts_key = nullptr;
    const float* core_key = nullptr;
    std::size_t m = 0;
    std::uint32_t stride = 0;
    std::vector<float> host_pts, host_core;  // what the device copies hold
    DBuf<float> pts, core, out;
    hipStream_t stream = nullptr;
    void release() {
        pts.release();
        core.release();
        out.release();
        host_pts.clear();
        host_core.clear();
        pts_key = core_key = nullptr;
        m = 0;
        stride = 0;
        destroy_stream(stream);
        stream = nullptr;
        device = -1;
    }
};

// Caches of threads that have exited, freed by the next call on a live thread
// (a leaked, never destroyed list: threads can exit during static destruction).
struct RetiredCaches {
    std::mutex mu;
    std::vector<MrdCache*> list;
};
RetiredCaches& retired_caches() {
    static RetiredCaches* r = new RetiredCaches();
    return *r;
}

void free_retired_caches() {
    std::vector<MrdCache*> r;
    {
        std::lock_guard<std::mutex> lk(retired_caches().mu);
        r.swap(retired_caches().list);
    }
    for (MrdCache* c : r) {
        c->release();
        delete c;
    }
}

// The calling thread's cache. An exiting thread makes no HIP call: by then the
// runtime, or a tool attached to it, may have torn down its state for the
// thread (a profiler failed on exactly that call), so the cache is handed to
// the retired list; at process exit the driver reclaims what is left.
struct MrdCacheHolder {
    MrdCache* c = nullptr;
    MrdCache& get() {
        if (!c) c = new MrdCache();
        return *c;
    }
    ~MrdCacheHolder() {
        if (!c) return;
        std::lock_guard<std::mutex> lk(retired_caches().mu);
        retired_caches().list.push_back(c);
    }
};
thread_local MrdCacheHolder t_holder;

void ensure_resident(int device, const float* host_pts, std::size_t m, std::uint32_t stride,
                     const float* host_core, bool check_points) {
    free_retired_caches();
    MrdCache& t_cache = t_holder.get();
    const std::size_t np = m * stride;
    if (t_cache.device == device && t_cache.pts_key == host_pts &&
        t_cache.core_key == host_core && t_cache.m == m && t_cache.stride == stride) {
        const bool pts_same =
            !check_points || std::memcmp(t_cache.host_pts.data(), host_pts, np * 4) == 0;
        if (pts_same && std::memcmp(t_cache.host_core.data(), host_core, m * 4) == 0) return;
        if (pts_same) {  // new core distances over the same points
            t_cache.host_core.assign(host_core, host_core + m);
            CSG_CHECK(hipMemcpyAsync(t_cache.core.p, t_cache.host_core.data(), m * 4,
                                     hipMemcpyHostToDevice, t_cache.stream));
            return;
        }
    }
    t_cache.release();
    t_cache.device = device;
    create_stream(&t_cache.stream);
    t_cache.pts.reserve(np);
    t_cache.core.reserve(m);
    t_cache.host_pts.assign(host_pts, host_pts + np);
    t_cache.host_core.assign(host_core, host_core + m);
    CSG_CHECK(hipMemcpyAsync(t_cache.pts.p, t_cache.host_pts.data(), np * 4,
                             hipMemcpyHostToDevice, t_cache.stream));
    CSG_CHECK(hipMemcpyAsync(t_cache.core.p, t_cache.host_core.data(), m * 4,
                             hipMemcpyHostToDevice, t_cache.stream));
    t_cache.pts_key = host_pts;
    t_cache.core_key = host_core;
    t_cache.m = m;
    t_cache.stride = stride;
}

void check_metric(int metric) {
    if (metric != CORESG_HIP_METRIC_L2 && metric != CORESG_HIP_METRIC_COSINE)
        throw std::runtime_error("unsupported metric for the device path (L2 and cosine only)");
}

}  // namespace

void mst_over_subset(Graph& g, const std::vector<std::uint32_t>& ids,
                     std::vector<std::uint32_t>* u, std::vector<std::uint32_t>* v,
                     std::vector<float>* w) {
    const std::uint32_t m = static_cast<std::uint32_t>(ids.size());
    if (m < 2) return;
    hipStream_t s = g.stream;
    const std::uint32_t stride = g.stride;
    DBuf<std::uint32_t> d_ids, from, ou, ov, cnt;
    DBuf<float> sub, core, mat, best, ow;
    DBuf<std::uint8_t> in_tree;
    d_ids.reserve(m);
    sub.reserve(std::size_t(m) * stride);
    core.reserve(m);
    mat.reserve(std::size_t(m) * m);
    best.reserve(m);
    from.reserve(m);
    in_tree.reserve(m);
    ou.reserve(m);
    ov.reserve(m);
    ow.reserve(m);
    cnt.reserve(1);
    CSG_CHECK(hipMemcpyAsync(d_ids.p, ids.data(), std::size_t(m) * 4, hipMemcpyHostToDevice, s));
    subset_gather_kernel<<<capped_blocks(std::size_t(m) * stride, 256), 256, 0, s>>>(
        g.pts.p, stride, d_ids.p, m, g.dist.p, g.count.p, g.k, sub.p, core.p);
    CSG_LAUNCH_CHECK();
    {
        ScopedKernelTimer t(g, "bridge_mrd_matrix");
        launch_mrd(sub.p, m, stride, core.p, 0, m, mat.p, s);
    }
    {
        ScopedKernelTimer t(g, "bridge_prim");
        prim_kernel<<<1, 256, 0, s>>>(mat.p, m, best.p, from.p, in_tree.p, ou.p, ov.p, ow.p,
                                      cnt.p);
        CSG_LAUNCH_CHECK();
    }
    base_weight_kernel<<<blocks_for(m, 256), 256, 0, s>>>(sub.p, stride, d_ids.p, cnt.p, ou.p,
                                                          ov.p, ow.p);
    CSG_LAUNCH_CHECK();
    std::uint32_t t = 0;
    CSG_CHECK(hipMemcpyAsync(&t, cnt.p, 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));
    const std::size_t base = u->size();
    u->resize(base + t);
    v->resize(base + t);
    w->resize(base + t);
    if (t) {
        CSG_CHECK(hipMemcpyAsync(u->data() + base, ou.p, std::size_t(t) * 4,
                                 hipMemcpyDeviceToHost, s));
        CSG_CHECK(hipMemcpyAsync(v->data() + base, ov.p, std::size_t(t) * 4,
                                 hipMemcpyDeviceToHost, s));
        CSG_CHECK(hipMemcpyAsync(w->data() + base, ow.p, std::size_t(t) * 4,
                                 hipMemcpyDeviceToHost, s));
        CSG_CHECK(hipStreamSynchronize(s));
    }
}

void mst_over_subsets(Graph& g, const std::vector<std::vector<std::uint32_t>>& sets,
                      std::vector<std::uint32_t>* u, std::vector<std::uint32_t>* v,
                      std::vector<float>* w) {
    const std::uint32_t stride = g.stride;
    if (stride > kMaxLdsStride) {  // rows too long for the LDS tiles: one by one
        for (const auto& ids : sets) mst_over_subset(g, ids, u, v, w);
        return;
    }
    hipStream_t s = g.stream;
    const std::uint32_t nsets = static_cast<std::uint32_t>(sets.size());
    if (nsets == 0) return;
    std::vector<SubsetMeta> meta(nsets);
    std::vector<std::uint32_t> all;
    std::uint64_t mat_total = 0;
    std::uint32_t tiles = 0;
    for (std::uint32_t i = 0; i < nsets; ++i) {
        const std::uint32_t m = static_cast<std::uint32_t>(sets[i].size());
        const std::uint32_t tx = (m + kTile - 1) / kTile;
        meta[i] = SubsetMeta{static_cast<std::uint32_t>(all.size()), m, tiles, tx, mat_total};
        all.insert(all.end(), sets[i].begin(), sets[i].end());
        mat_total += std::uint64_t(m) * m;
        tiles += m >= 2 ? tx * tx : 0;
        if (m < 2) meta[i].tiles_x = 0;
    }
    const std::uint32_t M = static_cast<std::uint32_t>(all.size());
    if (M == 0) return;
    DBuf<std::uint32_t> d_ids, from, ou, ov, cnt;
    DBuf<float> sub, core, mat, best, ow;
    DBuf<std::uint8_t> in_tree;
    DBuf<SubsetMeta> d_meta;
    d_ids.reserve(M);
    sub.reserve(std::size_t(M) * stride);
    core.reserve(M);
    mat.reserve(mat_total ? mat_total : 1);
    best.reserve(M);
    from.reserve(M);
    in_tree.reserve(M);
    ou.reserve(M);
    ov.reserve(M);
    ow.reserve(M);
    cnt.reserve(nsets);
    d_meta.reserve(nsets);
    CSG_CHECK(hipMemcpyAsync(d_ids.p, all.data(), std::size_t(M) * 4, hipMemcpyHostToDevice, s));
    CSG_CHECK(hipMemcpyAsync(d_meta.p, meta.data(), sizeof(SubsetMeta) * nsets,
                             hipMemcpyHostToDevice, s));
    subset_gather_kernel<<<capped_blocks(std::size_t(M) * stride, 256), 256, 0, s>>>(
        g.pts.p, stride, d_ids.p, M, g.dist.p, g.count.p, g.k, sub.p, core.p);
    CSG_LAUNCH_CHECK();
    if (tiles) {
        ScopedKernelTimer t(g, "bridge_mrd_matrix");
        const std::size_t lds = std::size_t(2) * kTile * stride * sizeof(float);
        mrd_tiles_batched_kernel<<<tiles, 256, lds, s>>>(sub.p, stride, core.p, d_meta.p, nsets,
                                                         mat.p);
        CSG_LAUNCH_CHECK();
    }
    {
        ScopedKernelTimer t(g, "bridge_prim");
        prim_batched_kernel<<<nsets, 256, 0, s>>>(mat.p, d_meta.p, best.p, from.p, in_tree.p,
                                                  ou.p, ov.p, ow.p, cnt.p);
        CSG_LAUNCH_CHECK();
    }
    base_weight_batched_kernel<<<nsets, 256, 0, s>>>(sub.p, stride, d_ids.p, d_meta.p, cnt.p,
                                                     ou.p, ov.p, ow.p);
    CSG_LAUNCH_CHECK();
    std::vector<std::uint32_t> hc(nsets), hu(M), hv(M);
    std::vector<float> hw(M);
    CSG_CHECK(hipMemcpyAsync(hc.data(), cnt.p, std::size_t(nsets) * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipMemcpyAsync(hu.data(), ou.p, std::size_t(M) * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipMemcpyAsync(hv.data(), ov.p, std::size_t(M) * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipMemcpyAsync(hw.data(), ow.p, std::size_t(M) * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));
    for (std::uint32_t i = 0; i < nsets; ++i) {
        const std::uint32_t so = meta[i].so;
        u->insert(u->end(), hu.begin() + so, hu.begin() + so + hc[i]);
        v->insert(v->end(), hv.begin() + so, hv.begin() + so + hc[i]);
        w->insert(w->end(), hw.begin() + so, hw.begin() + so + hc[i]);
    }
}

void mrd_block_host(int device, const float* pts, std::size_t m, std::uint32_t dim,
                    std::uint32_t stride, int metric, const float* core, std::size_t r0,
                    std::size_t r1, float* out) {
    (void)dim;
    if (m == 0 || r1 <= r0) return;
    check_metric(metric);
    if (stride % 8 != 0) throw std::runtime_error("mrd_block: stride must be a multiple of 8");
    if (r1 > m) throw std::runtime_error("mrd_block: row range exceeds m");
    ensure_resident(device, pts, m, stride, core, r0 == 0);
    MrdCache& t_cache = t_holder.get();
    const std::size_t rows = r1 - r0;
    t_cache.out.reserve(rows * m);
    launch_mrd(t_cache.pts.p, static_cast<std::uint32_t>(m), stride, t_cache.core.p,
               static_cast<std::uint32_t>(r0), static_cast<std::uint32_t>(rows), t_cache.out.p,
               t_cache.stream, metric == CORESG_HIP_METRIC_COSINE);
    CSG_CHECK(hipMemcpyAsync(out, t_cache.out.p, rows * m * 4, hipMemcpyDeviceToHost,
                             t_cache.stream));
    CSG_CHECK(hipStreamSynchronize(t_cache.stream));
}

void mrd_row_min_host(int device, const float* pts, std::size_t m, std::uint32_t dim,
                      std::uint32_t stride, int metric, const float* core,
                      const std::uint32_t* component, std::size_t r0, std::size_t r1,
                      float* out_min, std::uint32_t* out_argmin) {
    (void)dim;
    (void)device;
    if (m == 0 || r1 <= r0) return;
    check_metric(metric);
    if (stride % 8 != 0) throw std::runtime_error("mrd_row_min: stride must be a multiple of 8");
    if (r1 > m) throw std::runtime_error("mrd_row_min: row range exceeds m");
    OwnedStream os;
    hipStream_t s = os.s;
    const std::size_t rows = r1 - r0;
    DBuf<float> d_pts, d_core, d_min;
    DBuf<std::uint32_t> d_comp, d_arg;
    d_pts.reserve(m * stride);
    d_core.reserve(m);
    d_min.reserve(rows);
    d_arg.reserve(rows);
    CSG_CHECK(hipMemcpyAsync(d_pts.p, pts, m * stride * 4, hipMemcpyHostToDevice, s));
    CSG_CHECK(hipMemcpyAsync(d_core.p, core, m * 4, hipMemcpyHostToDevice, s));
    if (component) {
        d_comp.reserve(m);
        CSG_CHECK(hipMemcpyAsync(d_comp.p, component, m * 4, hipMemcpyHostToDevice, s));
    }
    row_min_kernel<<<static_cast<unsigned>(rows < 65536 ? rows : 65536), 256, 0, s>>>(
        d_pts.p, static_cast<std::uint32_t>(m), stride, d_core.p, component ? d_comp.p : nullptr,
        static_cast<std::uint32_t>(r0), static_cast<std::uint32_t>(rows), d_min.p, d_arg.p,
        metric == CORESG_HIP_METRIC_COSINE);
    CSG_LAUNCH_CHECK();
    CSG_CHECK(hipMemcpyAsync(out_min, d_min.p, rows * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipMemcpyAsync(out_argmin, d_arg.p, rows * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));
}

void release_thread_cache() {
    free_retired_caches();
    if (t_holder.c) t_holder.c->release();
}

}  // namespace coresg_hip
