// Bandwidth probes of known algorithmic bytes, for calibrating hardware byte
// counters against the access patterns of the path: a random gather of whole
// point rows (the local join's and the insert search's pattern) and a
// streaming copy.
#include <string>

#include "../../include/coresg_hip.h"
#include "graph.hpp"
#include "handle.hpp"
#include "mfma.hpp"
#include "scan.hpp"

namespace coresg_hip {
namespace {

// Thread t reads row (t * 2654435761 + seed) mod rows, whole (stride floats,
// float4 loads), and writes the sum of its elements.
__global__ void __launch_bounds__(256)
    gather_probe_kernel(const float* __restrict__ pts, std::uint32_t rows, std::uint32_t stride,
                        std::uint64_t gathers, std::uint32_t seed, float* __restrict__ out) {
    for (std::uint64_t t = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < gathers;
         t += std::uint64_t(gridDim.x) * blockDim.x) {
        const std::uint32_t r =
            static_cast<std::uint32_t>((t * 2654435761ull + seed) % rows);
        const float4* p = reinterpret_cast<const float4*>(pts + std::size_t(r) * stride);
        float s = 0.0f;
        for (std::uint32_t q = 0; q < stride / 4; ++q) {
            const float4 v = p[q];
            s += (v.x + v.y) + (v.z + v.w);
        }
        out[t] = s;
    }
}

// One wavefront: C = A B^T through gram_mfma32 (A, B: 16 x 32, row-major),
// with the lane layout the filter join assumes.
template <bool BF16>
__global__ void __launch_bounds__(64)
    mfma_probe_kernel(const float* __restrict__ a, const float* __restrict__ b,
                      float* __restrict__ c) {
    const unsigned lane = threadIdx.x, li = lane & 15u, kq = lane >> 4;
    float x[8], y[8];
    for (int q = 0; q < 8; ++q) {
        x[q] = a[li * 32 + 8 * kq + q];
        y[q] = b[li * 32 + 8 * kq + q];
    }
    gram_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = gram_mfma32<BF16>(x, y, acc);
    for (int q = 0; q < 4; ++q) c[(4 * kq + q) * 16 + li] = acc[q];
}

}  // namespace
}  // namespace coresg_hip

extern "C" {

int coresg_hip_probe_gather(int device, size_t rows, uint32_t stride, size_t gathers,
                            double* ms) {
    return coresg_hip::guarded([&] {
        using namespace coresg_hip;
        if (rows == 0 || rows >= (1ull << 32) || stride == 0 || stride % 4 != 0 || gathers == 0)
            throw std::runtime_error("probe_gather: rows, stride (multiple of 4), gathers");
        CSG_CHECK(hipSetDevice(device));
        OwnedStream os;
        DBuf<float> pts, out;
        pts.reserve(rows * stride);
        out.reserve(gathers);
        fill_u32(reinterpret_cast<std::uint32_t*>(pts.p), rows * stride, 0x3F800000u, os.s);
        hipEvent_t e0, e1;
        CSG_CHECK(hipEventCreate(&e0));
        CSG_CHECK(hipEventCreate(&e1));
        CSG_CHECK(hipEventRecord(e0, os.s));
        gather_probe_kernel<<<capped_blocks(gathers, 256, 65536), 256, 0, os.s>>>(
            pts.p, static_cast<std::uint32_t>(rows), stride, gathers, 12345u, out.p);
        CSG_LAUNCH_CHECK();
        CSG_CHECK(hipEventRecord(e1, os.s));
        CSG_CHECK(hipEventSynchronize(e1));
        float t = 0.0f;
        CSG_CHECK(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (ms) *ms = t;
    });
}

int coresg_hip_probe_mfma(int device, int bf16, const float* a, const float* b, float* c) {
    return coresg_hip::guarded([&] {
        using namespace coresg_hip;
        if (!a || !b || !c) throw std::runtime_error("probe_mfma: null operand");
        CSG_CHECK(hipSetDevice(device));
        OwnedStream os;
        DBuf<float> da, db, dc;
        da.reserve(16 * 32);
        db.reserve(16 * 32);
        dc.reserve(16 * 16);
        CSG_CHECK(hipMemcpyAsync(da.p, a, 16 * 32 * sizeof(float), hipMemcpyHostToDevice, os.s));
        CSG_CHECK(hipMemcpyAsync(db.p, b, 16 * 32 * sizeof(float), hipMemcpyHostToDevice, os.s));
        if (bf16)
            mfma_probe_kernel<true><<<1, 64, 0, os.s>>>(da.p, db.p, dc.p);
        else
            mfma_probe_kernel<false><<<1, 64, 0, os.s>>>(da.p, db.p, dc.p);
        CSG_LAUNCH_CHECK();
        CSG_CHECK(hipMemcpyAsync(c, dc.p, 16 * 16 * sizeof(float), hipMemcpyDeviceToHost, os.s));
        CSG_CHECK(hipStreamSynchronize(os.s));
    });
}

}  // extern "C"
