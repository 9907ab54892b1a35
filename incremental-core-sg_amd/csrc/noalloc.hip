// Allocation-free variant of the MRD block entry point.
//
// coresg_hip_mrd_block stages the caller's host operands through device
// buffers it allocates (and caches per thread). Callers that may not allocate
// device memory -- code under stream capture, tight serving loops, harnesses
// that forbid hipMalloc -- get the same rows through this entry point, which
// stages everything through one fixed array in the code object's own device
// memory instead:
//
//   scratch = [ points m*stride | core m | rows (r1-r0)*m ]   (floats, each
//             part starting on a 64-float boundary)
//
// The array is kNoAllocFloats floats (32 MiB) per GPU and lives until the
// process ends. Every call on one GPU uses the same array, so calls are
// serialised per device by a host lock here; the array keeps the last call's
// contents, and nothing in it needs to start at zero (the kernel writes every
// element of the requested rows).
#include <array>
#include <mutex>
#include <string>

#include "../../include/coresg_hip.h"
#include "build.hpp"
#include "handle.hpp"

namespace coresg_hip {
namespace {

constexpr std::size_t kNoAllocFloats = std::size_t(8) << 20;  // 32 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) float scratch[kNoAllocFloats];

// The device address of `scratch` on each GPU (its host-side name is only a
// stand-in), looked up once per device with that device current.
struct ScratchTable {
    std::mutex mu;
    std::array<float*, kMaxDevices> ptr{};
    std::array<std::mutex, kMaxDevices> call;
};

ScratchTable& table() {
    static ScratchTable t;
    return t;
}

float* scratch_on(int device) {
    ScratchTable& t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    if (!t.ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw HipError(std::string("hipGetSymbolAddress(scratch) -> ") + hipGetErrorName(e), e);
        t.ptr[device] = static_cast<float*>(p);
    }
    return t.ptr[device];
}

constexpr std::size_t round64(std::size_t x) { return (x + 63) & ~std::size_t(63); }

}  // namespace

std::size_t mrd_block_noalloc_capacity() { return kNoAllocFloats; }

void mrd_block_noalloc_host(int device, const float* pts, std::size_t m, std::uint32_t stride,
                            int metric, const float* core, std::size_t r0, std::size_t r1,
                            float* out) {
    if (device < 0 || device >= kMaxDevices)
        throw std::runtime_error("mrd_block_noalloc: device id out of range");
    if (m == 0 || r1 <= r0) return;
    if (metric != CORESG_HIP_METRIC_L2 && metric != CORESG_HIP_METRIC_COSINE)
        throw std::runtime_error("mrd_block_noalloc: unsupported metric (L2 and cosine only)");
    if (stride % 8 != 0) throw std::runtime_error("mrd_block_noalloc: stride must be a multiple of 8");
    if (r1 > m) throw std::runtime_error("mrd_block_noalloc: row range exceeds m");
    if (m > 0xFFFFFFFFull) throw std::runtime_error("mrd_block_noalloc: m above 2^32 - 1");
    const std::size_t rows = r1 - r0;
    const std::size_t o_core = round64(m * stride), o_out = o_core + round64(m);
    const std::size_t need = o_out + rows * m;
    if (m * stride / stride != m || need < o_out || need > kNoAllocFloats)
        throw std::runtime_error("mrd_block_noalloc: points, core and rows need " +
                                 std::to_string(need) + " floats, above the fixed scratch of " +
                                 std::to_string(kNoAllocFloats) + " floats (32 MiB)");
    CSG_CHECK(hipSetDevice(device));
    float* base = scratch_on(device);
    std::lock_guard<std::mutex> lk(table().call[device]);
    // the null stream: creating a stream would allocate
    hipStream_t s = nullptr;
    CSG_CHECK(hipMemcpyAsync(base, pts, m * stride * 4, hipMemcpyHostToDevice, s));
    CSG_CHECK(hipMemcpyAsync(base + o_core, core, m * 4, hipMemcpyHostToDevice, s));
    launch_mrd(base, static_cast<std::uint32_t>(m), stride, base + o_core,
               static_cast<std::uint32_t>(r0), static_cast<std::uint32_t>(rows), base + o_out, s,
               metric == CORESG_HIP_METRIC_COSINE);
    CSG_CHECK(hipMemcpyAsync(out, base + o_out, rows * m * 4, hipMemcpyDeviceToHost, s));
    CSG_CHECK(hipStreamSynchronize(s));
}

}  // namespace coresg_hip

extern "C" {

int coresg_hip_mrd_block_noalloc(int device, const float* host_pts, size_t m, uint32_t dim,
                                 uint32_t stride, int metric, const float* host_core, size_t r0,
                                 size_t r1, float* host_out) {
    (void)dim;
    return coresg_hip::guarded([&] {
        coresg_hip::mrd_block_noalloc_host(device, host_pts, m, stride, metric, host_core, r0, r1,
                                           host_out);
    });
}

size_t coresg_hip_mrd_block_noalloc_capacity(void) {
    return coresg_hip::mrd_block_noalloc_capacity();
}

}  // extern "C"
