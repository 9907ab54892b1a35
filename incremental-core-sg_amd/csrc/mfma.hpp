// Matrix-core helpers shared by the filter joins, the insert scan and the
// layout probe: FP16 / BF16 rounding, the formats' error constants and the
// MFMA wrappers (v_mfma_f32_16x16x16{f16,bf16} and gfx950's
// v_mfma_f32_16x16x32_{f16,bf16}).
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace coresg_hip {

typedef _Float16 gram_f16x4 __attribute__((ext_vector_type(4)));
typedef short gram_i16x4 __attribute__((ext_vector_type(4)));
typedef float gram_f32x4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ std::uint32_t bf16_bits(float x) {
    std::uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}

template <bool BF16>
__device__ __forceinline__ float gram_round(float x) {
    if constexpr (BF16) return __uint_as_float(bf16_bits(x) << 16);
    else return static_cast<float>(static_cast<_Float16>(x));
}

template <bool BF16>
struct GramConst {
    // unit roundoff of the format plus the centring's fp32 rounding
    static constexpr float rho = BF16 ? (0x1p-8f + 0x1p-23f) : (0x1p-11f + 0x1p-23f);
    // per-coordinate absolute floor (a flushed subnormal)
    static constexpr float absc = BF16 ? 0x1p-126f : 0x1p-14f;
};

template <bool BF16>
__device__ __forceinline__ gram_f32x4 gram_mfma(const float (&x)[4], const float (&y)[4],
                                                gram_f32x4 acc) {
    if constexpr (BF16) {
        gram_i16x4 av, bv;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            av[q] = static_cast<short>(bf16_bits(x[q]));
            bv[q] = static_cast<short>(bf16_bits(y[q]));
        }
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    } else {
        gram_f16x4 av, bv;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            av[q] = static_cast<_Float16>(x[q]);
            bv[q] = static_cast<_Float16>(y[q]);
        }
        return __builtin_amdgcn_mfma_f32_16x16x16f16(av, bv, acc, 0, 0, 0);
    }
}

// gfx950 matrix-core operands: eight FP16 / BF16 values per lane.
typedef _Float16 gram_f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 gram_bf16x8 __attribute__((ext_vector_type(8)));

// One v_mfma_f32_16x16x32_{f16,bf16}: lane (li = lane % 16, kq = lane / 16)
// supplies row li of A and column li of B at coordinates 8 kq .. 8 kq + 7
// and receives C[4 kq + q][li], q = 0..3 (layout checked on the device by
// coresg_hip_probe_mfma).
template <bool BF16>
__device__ __forceinline__ gram_f32x4 gram_mfma32(const float (&x)[8], const float (&y)[8],
                                                  gram_f32x4 acc) {
    if constexpr (BF16) {
        gram_bf16x8 av, bv;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            av[q] = static_cast<__bf16>(x[q]);
            bv[q] = static_cast<__bf16>(y[q]);
        }
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    } else {
        gram_f16x8 av, bv;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            av[q] = static_cast<_Float16>(x[q]);
            bv[q] = static_cast<_Float16>(y[q]);
        }
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bv, acc, 0, 0, 0);
    }
}

}  // namespace coresg_hip
