// Shared device and host helpers for the gfx950 CORE-SG library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace coresg_hip {

constexpr std::uint32_t kNone = 0xFFFFFFFFu;
constexpr std::uint64_t kNone64 = 0xFFFFFFFFFFFFFFFFull;
constexpr int kWave = 64;

// Thread-local message of the most recent failure on the calling host thread
// (the reference keeps the same convention, cuda/cuda_runtime_api.cu:15-23).
void set_last_error(const std::string& msg);
const char* get_last_error();

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(const std::string& m, hipError_t c) : std::runtime_error(m), code(c) {}
};

#define CSG_CHECK(expr)                                                                   \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess) {                                                           \
            throw ::coresg_hip::HipError(std::string(#expr) + " -> " +                    \
                                             hipGetErrorName(_e) + ": " +                  \
                                             hipGetErrorString(_e),                       \
                                         _e);                                             \
        }                                                                                 \
    } while (0)

// The plain form (one launch per step with a host synchronisation after each,
// the shape of the reference's CUDA driver loop, nndescent_kernels.cu:407-418):
// while an entry point serves a graph in that mode, this thread's launch checks
// also wait for the graph's stream.
inline thread_local hipStream_t t_plain_stream = nullptr;

#define CSG_LAUNCH_CHECK()                                                 \
    do {                                                                  \
        CSG_CHECK(hipGetLastError());                                     \
        if (::coresg_hip::t_plain_stream)                                 \
            CSG_CHECK(hipStreamSynchronize(::coresg_hip::t_plain_stream)); \
    } while (0)

inline unsigned blocks_for(std::size_t total, unsigned threads) {
    std::size_t b = (total + threads - 1) / threads;
    return static_cast<unsigned>(b ? b : 1);
}

// Grid-stride loop bound helper: at most `cap` blocks (enough to fill 256 CUs
// several times over), never zero.
inline unsigned capped_blocks(std::size_t total, unsigned threads, unsigned cap = 8192) {
    unsigned b = blocks_for(total, threads);
    return b > cap ? cap : b;
}

// ---------------------------------------------------------------------------
// Exact squared L2 in the reference's summation order.
//
// The reference's host kernel (include/coresg/distance.hpp:127-147, hsum256 at
// :116-125) keeps two 8-lane accumulators over 16-float chunks, adds one
// 8-float step into the first, then reduces v = acc0 + acc1 as
// ((v0+v4)+(v1+v5)) + ((v2+v6)+(v3+v7)). Reproducing that order with explicit
// fused multiply-adds makes every distance bit-identical to the host's, which
// is what keeps k-NN lists, core distances and MST weights exact. Rows are
// padded to a multiple of 8 floats, so `stride % 8 == 0`.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float l2sqr_exact(const float* __restrict__ a,
                                             const float* __restrict__ b, std::uint32_t stride) {
    float acc0[8], acc1[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        acc0[j] = 0.0f;
        acc1[j] = 0.0f;
    }
    std::uint32_t i = 0;
    for (; i + 16 <= stride; i += 16) {
        const float4* pa = reinterpret_cast<const float4*>(a + i);
        const float4* pb = reinterpret_cast<const float4*>(b + i);
        float4 a0 = pa[0], a1 = pa[1], a2 = pa[2], a3 = pa[3];
        float4 b0 = pb[0], b1 = pb[1], b2 = pb[2], b3 = pb[3];
        float d[16] = {a0.x - b0.x, a0.y - b0.y, a0.z - b0.z, a0.w - b0.w,
                       a1.x - b1.x, a1.y - b1.y, a1.z - b1.z, a1.w - b1.w,
                       a2.x - b2.x, a2.y - b2.y, a2.z - b2.z, a2.w - b2.w,
                       a3.x - b3.x, a3.y - b3.y, a3.z - b3.z, a3.w - b3.w};
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            acc0[j] = __builtin_fmaf(d[j], d[j], acc0[j]);
            acc1[j] = __builtin_fmaf(d[8 + j], d[8 + j], acc1[j]);
        }
    }
    if (i + 8 <= stride) {
        const float4* pa = reinterpret_cast<const float4*>(a + i);
        const float4* pb = reinterpret_cast<const float4*>(b + i);
        float4 a0 = pa[0], a1 = pa[1];
        float4 b0 = pb[0], b1 = pb[1];
        float d[8] = {a0.x - b0.x, a0.y - b0.y, a0.z - b0.z, a0.w - b0.w,
                      a1.x - b1.x, a1.y - b1.y, a1.z - b1.z, a1.w - b1.w};
#pragma unroll
        for (int j = 0; j < 8; ++j) acc0[j] = __builtin_fmaf(d[j], d[j], acc0[j]);
    }
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = acc0[j] + acc1[j];
    const float l0 = v[0] + v[4], l1 = v[1] + v[5], l2 = v[2] + v[6], l3 = v[3] + v[7];
    return (l0 + l1) + (l2 + l3);
}

// IEEE-754 correctly rounded square root (the host's std::sqrt). hipcc's
// default sqrtf on gfx950 is correctly rounded unless fast-math flags are
// given, which this library never uses.
__device__ __forceinline__ float sqrt_exact(float x) { return __builtin_sqrtf(x); }

// MetricOp<L2> / MetricOp<Cosine> (distance.hpp:196-218): both rank by the
// squared Euclidean distance; L2 stores sqrt(r), cosine (rows L2-normalised on
// load, so the cosine distance is exactly r / 2) stores 0.5f * r. to_rank maps
// a stored value back: v * v and 2.0f * v.
__device__ __forceinline__ float metric_finalize(float r, bool cosine) {
    return cosine ? 0.5f * r : sqrt_exact(r);
}
__device__ __forceinline__ float metric_to_rank(float v, bool cosine) {
    return cosine ? 2.0f * v : v * v;
}

// std::max(a, b) for floats: b if a < b, else a (no NaNs occur here).
__device__ __forceinline__ float fmax_ref(float a, float b) { return a < b ? b : a; }

__device__ __forceinline__ std::uint32_t fbits(float x) { return __float_as_uint(x); }

// Lane helpers for a 64-wide wavefront.
__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ std::uint64_t wave_min_u64(std::uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        std::uint64_t o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

}  // namespace coresg_hip
