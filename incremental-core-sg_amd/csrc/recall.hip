// Sampled k-NN recall on device (src/eval.cpp:24-97).
//
// Queries are drawn on the host exactly as the reference draws them (partial
// Fisher-Yates over std::mt19937_64(seed)). For a chunk of queries the device
// computes every exact squared distance, finds each query's k smallest
// (distance, id) keys with a batched MSB radix select, and collects them; the
// host compares them with the stored lists and accumulates the means in query
// order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "build.hpp"
#include "scan.hpp"

namespace coresg_hip {
namespace {

constexpr unsigned kThreads = 256;

__global__ void __launch_bounds__(kThreads)
    query_dist_kernel(const float* __restrict__ pts, std::uint32_t n, std::uint32_t stride,
                      const std::uint32_t* __restrict__ qids, float* __restrict__ out) {
    const std::uint32_t q = blockIdx.y;
    const std::uint32_t qi = qids[q];
    const float* a = pts + std::size_t(qi) * stride;
    for (std::uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += gridDim.x * blockDim.x) {
        out[std::size_t(q) * n + j] =
            (j == qi) ? __builtin_inff() : l2sqr_exact(a, pts + std::size_t(j) * stride, stride);
    }
}

struct QState {
    unsigned long long pref, mask;
    unsigned int rank, pad;
};

__device__ __forceinline__ unsigned long long qkey(const float* d, std::uint32_t n,
                                                   std::uint32_t q, std::uint32_t j) {
    return (static_cast<unsigned long long>(fbits(d[std::size_t(q) * n + j])) << 32) | j;
}

__global__ void __launch_bounds__(kThreads)
    qhist_kernel(const float* __restrict__ d, std::uint32_t n, const QState* __restrict__ st,
                 int digit, std::uint32_t* __restrict__ hist) {
    __shared__ std::uint32_t h[256];
    const std::uint32_t q = blockIdx.y;
    h[threadIdx.x] = 0;
    __syncthreads();
    const QState s = st[q];
    const int shift = 56 - 8 * digit;
    for (std::uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += gridDim.x * blockDim.x) {
        const unsigned long long key = qkey(d, n, q, j);
        if ((key & s.mask) != s.pref) continue;
        atomicAdd(&h[(key >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[q * 256 + threadIdx.x], h[threadIdx.x]);
}

__global__ void __launch_bounds__(256)
    qpick_kernel(QState* __restrict__ st, int digit, std::uint32_t* __restrict__ hist) {
    __shared__ std::uint32_t c[256];
    const std::uint32_t q = blockIdx.x;
    c[threadIdx.x] = hist[q * 256 + threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        QState s = st[q];
        std::uint32_t cum = 0, b = 0;
        for (b = 0; b < 256; ++b) {
            if (s.rank < cum + c[b]) break;
            cum += c[b];
        }
        if (b == 256) b = 255;
        s.rank -= cum;
        const int shift = 56 - 8 * digit;
        s.pref |= static_cast<unsigned long long>(b) << shift;
        s.mask |= 0xFFull << shift;
        st[q] = s;
    }
    __syncthreads();
    hist[q * 256 + threadIdx.x] = 0;
}

__global__ void __launch_bounds__(kThreads)
    qcollect_kernel(const float* __restrict__ d, std::uint32_t n, const QState* __restrict__ st,
                    std::uint32_t kk, std::uint32_t* __restrict__ out,
                    std::uint32_t* __restrict__ cnt) {
    const std::uint32_t q = blockIdx.y;
    const unsigned long long t = st[q].pref;
    for (std::uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += gridDim.x * blockDim.x) {
        if (qkey(d, n, q, j) <= t) {
            const std::uint32_t pos = atomicAdd(&cnt[q], 1u);
            if (pos < kk) out[std::size_t(q) * kk + pos] = j;
        }
    }
}

}  // namespace

void sampled_recall(Graph& g, std::size_t queries, std::uint64_t seed, double* recall,
                    double* ratio) {
    const std::uint32_t n = g.n, k = g.k;
    *recall = 0.0;
    *ratio = 0.0;
    if (n == 0 || k == 0 || queries == 0) return;
    if (!g.knn_set) throw std::runtime_error("sampled_recall: no k-NN lists");
    std::vector<std::uint32_t> qs(n);
    for (std::uint32_t i = 0; i < n; ++i) qs[i] = i;
    std::mt19937_64 rng(seed);
    if (queries < n) {
        for (std::size_t i = 0; i < queries; ++i) std::swap(qs[i], qs[i + rng() % (n - i)]);
        qs.resize(queries);
    }
    const std::uint32_t nq = static_cast<std::uint32_t>(qs.size());
    const std::uint32_t kk = k < n - 1 ? k : n - 1;  // |truth| (eval.cpp:55-68)
    hipStream_t s = g.stream;
    const std::uint32_t chunk = static_cast<std::uint32_t>(
        std::max<std::size_t>(1, std::min<std::size_t>(64, (std::size_t(1) << 28) / n)));
    DBuf<float> d;
    DBuf<std::uint32_t> d_q, hist, truth, tcnt;
    DBuf<QState> st;
    d.reserve(std::size_t(chunk) * n);
    d_q.reserve(chunk);
    hist.reserve(std::size_t(chunk) * 256);
    truth.reserve(std::size_t(chunk) * (kk ? kk : 1));
    tcnt.reserve(chunk);
    st.reserve(chunk);
    std::vector<std::uint32_t> h_truth(std::size_t(chunk) * (kk ? kk : 1));
    std::vector<std::uint32_t> row_ids(k);
    std::vector<float> row_d(k);
    double sum_hit = 0.0, sum_true = 0.0, sum_approx = 0.0;
    for (std::uint32_t q0 = 0; q0 < nq; q0 += chunk) {
        const std::uint32_t qc = std::min(chunk, nq - q0);
        CSG_CHECK(hipMemcpyAsync(d_q.p, qs.data() + q0, std::size_t(qc) * 4,
                                 hipMemcpyHostToDevice, s));
        const unsigned bx = std::min<unsigned>(blocks_for(n, kThreads), 1024);
        query_dist_kernel<<<dim3(bx, qc), kThreads, 0, s>>>(g.pts.p, n, g.stride, d_q.p, d.p);
        CSG_LAUNCH_CHECK();
        std::vector<QState> init(qc, QState{0, 0, kk ? kk - 1 : 0, 0});
        CSG_CHECK(hipMemcpyAsync(st.p, init.data(), qc * sizeof(QState), hipMemcpyHostToDevice,
                                 s));
        fill_u32(hist.p, std::size_t(qc) * 256, 0u, s);
        fill_u32(tcnt.p, qc, 0u, s);
        if (kk > 0) {
            for (int dg = 0; dg < 8; ++dg) {
                qhist_kernel<<<dim3(bx, qc), kThreads, 0, s>>>(d.p, n, st.p, dg, hist.p);
                CSG_LAUNCH_CHECK();
                qpick_kernel<<<qc, 256, 0, s>>>(st.p, dg, hist.p);
                CSG_LAUNCH_CHECK();
            }
            qcollect_kernel<<<dim3(bx, qc), kThreads, 0, s>>>(d.p, n, st.p, kk, truth.p, tcnt.p);
            CSG_LAUNCH_CHECK();
        }
        std::vector<QState> fin(qc);
        CSG_CHECK(hipMemcpyAsync(fin.data(), st.p, qc * sizeof(QState), hipMemcpyDeviceToHost,
                                 s));
        CSG_CHECK(hipMemcpyAsync(h_truth.data(), truth.p, std::size_t(qc) * kk * 4,
                                 hipMemcpyDeviceToHost, s));
        CSG_CHECK(hipStreamSynchronize(s));
        for (std::uint32_t t = 0; t < qc; ++t) {
            const std::uint32_t q = qs[q0 + t];
            std::uint32_t cnt = 0;
            CSG_CHECK(hipMemcpyAsync(row_ids.data(), g.ids.p + std::size_t(q) * k,
                                     std::size_t(k) * 4, hipMemcpyDeviceToHost, s));
            CSG_CHECK(hipMemcpyAsync(row_d.data(), g.dist.p + std::size_t(q) * k,
                                     std::size_t(k) * 4, hipMemcpyDeviceToHost, s));
            CSG_CHECK(hipMemcpyAsync(&cnt, g.count.p + q, 4, hipMemcpyDeviceToHost, s));
            CSG_CHECK(hipStreamSynchronize(s));
            const std::uint32_t ca = cnt < k ? cnt : k;
            std::size_t h = 0;
            for (std::uint32_t j = 0; j < ca; ++j)
                for (std::uint32_t z = 0; z < kk; ++z)
                    if (h_truth[std::size_t(t) * kk + z] == row_ids[j]) {
                        ++h;
                        break;
                    }
            sum_hit += kk ? double(h) / double(kk) : 0.0;
            const std::uint32_t kth_bits = static_cast<std::uint32_t>(fin[t].pref >> 32);
            float kth;
            std::memcpy(&kth, &kth_bits, 4);
            sum_true += kk ? double(std::sqrt(kth)) : 0.0;
            sum_approx += ca ? double(row_d[ca - 1]) : 0.0;
        }
    }
    *recall = sum_hit / double(nq);
    *ratio = sum_true > 0.0 ? sum_approx / sum_true : 1.0;
}

}  // namespace coresg_hip
