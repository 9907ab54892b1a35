// Reduce-then-scan exclusive prefix sum over u32 (device-wide, recursive).
#include "scan.hpp"

namespace coresg_hip {
namespace {

constexpr unsigned kScanThreads = 256;
constexpr unsigned kScanItems = 16;
constexpr std::size_t kScanTile = std::size_t(kScanThreads) * kScanItems;

__global__ void __launch_bounds__(kScanThreads)
    scan_reduce_kernel(const std::uint32_t* __restrict__ data, std::size_t n,
                       std::uint32_t* __restrict__ block_sums) {
    const std::size_t base = std::size_t(blockIdx.x) * kScanTile;
    std::uint32_t s = 0;
    for (unsigned i = threadIdx.x; i < kScanTile; i += kScanThreads) {
        const std::size_t g = base + i;
        if (g < n) s += data[g];
    }
    s = wave_sum(s);
    __shared__ std::uint32_t ws[kScanThreads / 64];
    if (__lane_id() == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        std::uint32_t t = 0;
        for (unsigned w = 0; w < kScanThreads / 64; ++w) t += ws[w];
        block_sums[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(kScanThreads)
    scan_apply_kernel(std::uint32_t* __restrict__ data, std::size_t n,
                      const std::uint32_t* __restrict__ block_offsets,
                      std::uint32_t* __restrict__ total) {
    const std::size_t base = std::size_t(blockIdx.x) * kScanTile +
                             std::size_t(threadIdx.x) * kScanItems;
    std::uint32_t v[kScanItems];
    std::uint32_t local = 0;
#pragma unroll
    for (unsigned i = 0; i < kScanItems; ++i) {
        const std::size_t g = base + i;
        v[i] = g < n ? data[g] : 0u;
        local += v[i];
    }
    std::uint32_t btot = 0;
    std::uint32_t run = block_exclusive_scan_256(local, &btot) + block_offsets[blockIdx.x];
#pragma unroll
    for (unsigned i = 0; i < kScanItems; ++i) {
        const std::size_t g = base + i;
        if (g < n) data[g] = run;
        run += v[i];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) *total = run;
}

__global__ void scan_one_kernel(std::uint32_t* __restrict__ data, std::size_t n,
                                std::uint32_t* __restrict__ total) {
    // n <= 1: the exclusive scan of one element is 0 and the total is the element.
    if (threadIdx.x == 0) {
        const std::uint32_t t = n ? data[0] : 0u;
        if (n) data[0] = 0u;
        if (total) *total = t;
    }
}

template <typename T>
__global__ void fill_kernel(T* __restrict__ p, std::size_t n, T value) {
    for (std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += std::size_t(gridDim.x) * blockDim.x)
        p[i] = value;
}

}  // namespace

std::size_t scan_tmp_words(std::size_t n) {
    std::size_t words = 0;
    while (n > 1) {
        const std::size_t nb = (n + kScanTile - 1) / kScanTile;
        words += nb + 1;
        n = nb;
    }
    return words + 1;
}

void exclusive_scan_u32(std::uint32_t* data, std::size_t n, std::uint32_t* total,
                        std::uint32_t* tmp, hipStream_t stream) {
    if (n <= 1) {
        scan_one_kernel<<<1, 64, 0, stream>>>(data, n, total);
        CSG_LAUNCH_CHECK();
        return;
    }
    const std::size_t nb = (n + kScanTile - 1) / kScanTile;
    std::uint32_t* sums = tmp;
    scan_reduce_kernel<<<static_cast<unsigned>(nb), kScanThreads, 0, stream>>>(data, n, sums);
    CSG_LAUNCH_CHECK();
    exclusive_scan_u32(sums, nb, nullptr, tmp + nb + 1, stream);
    scan_apply_kernel<<<static_cast<unsigned>(nb), kScanThreads, 0, stream>>>(data, n, sums,
                                                                            total);
    CSG_LAUNCH_CHECK();
}

void fill_u32(std::uint32_t* p, std::size_t n, std::uint32_t value, hipStream_t stream) {
    if (n == 0) return;
    fill_kernel<std::uint32_t><<<capped_blocks(n, 256), 256, 0, stream>>>(p, n, value);
    CSG_LAUNCH_CHECK();
}

void fill_u64(std::uint64_t* p, std::size_t n, std::uint64_t value, hipStream_t stream) {
    if (n == 0) return;
    fill_kernel<std::uint64_t><<<capped_blocks(n, 256), 256, 0, stream>>>(p, n, value);
    CSG_LAUNCH_CHECK();
}

void fill_u8(std::uint8_t* p, std::size_t n, std::uint8_t value, hipStream_t stream) {
    if (n == 0) return;
    fill_kernel<std::uint8_t><<<capped_blocks(n, 256), 256, 0, stream>>>(p, n, value);
    CSG_LAUNCH_CHECK();
}

}  // namespace coresg_hip
