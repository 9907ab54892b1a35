// Device-wide exclusive prefix sums (reduce-then-scan), used for CSR offsets,
// ordered stream compaction and dense relabelling.
#pragma once

#include "common.hpp"

namespace coresg_hip {

// In-place exclusive scan of `n` u32 values; if `total` is non-null the grand
// total is written there (device pointer). `tmp` must hold scan_tmp_words(n)
// u32 words. Everything runs on `stream`.
std::size_t scan_tmp_words(std::size_t n);
void exclusive_scan_u32(std::uint32_t* data, std::size_t n, std::uint32_t* total,
                        std::uint32_t* tmp, hipStream_t stream);

// Stream-ordered fills (kernels rather than hipMemsetAsync, so that code a
// caller captures into a HIP graph records no memset nodes).
void fill_u32(std::uint32_t* p, std::size_t n, std::uint32_t value, hipStream_t stream);
void fill_u64(std::uint64_t* p, std::size_t n, std::uint64_t value, hipStream_t stream);
void fill_u8(std::uint8_t* p, std::size_t n, std::uint8_t value, hipStream_t stream);

// Wave-level inclusive prefix sum of one u32 per lane.
__device__ __forceinline__ std::uint32_t wave_inclusive_scan(std::uint32_t v) {
    const int lane = __lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        std::uint32_t o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    return v;
}

// Block-level exclusive scan for blockDim.x == 256 (4 waves). Returns the
// exclusive prefix of `v`; `*block_total` receives the block's total.
__device__ __forceinline__ std::uint32_t block_exclusive_scan_256(std::uint32_t v,
                                                                  std::uint32_t* block_total) {
    __shared__ std::uint32_t wsum[4];
    const int lane = __lane_id();
    const int wid = threadIdx.x >> 6;
    const std::uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    std::uint32_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const std::uint32_t s = wsum[w];
        if (w < wid) base += s;
        tot += s;
    }
    __syncthreads();
    *block_total = tot;
    return base + inc - v;
}

}  // namespace coresg_hip
