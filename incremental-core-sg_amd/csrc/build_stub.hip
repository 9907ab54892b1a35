// Temporary: build entry points not yet implemented on device.
#include "build.hpp"

namespace coresg_hip {

void build_knn(Graph&, const coresg_hip_build_params&, coresg_hip_build_stats*) {
    throw std::runtime_error("build_knn: not implemented");
}
void build_bridges(Graph&, const coresg_hip_build_params&, coresg_hip_build_stats*) {
    throw std::runtime_error("build_bridges: not implemented");
}
void sampled_recall(Graph&, std::size_t, std::uint64_t, double*, double*) {
    throw std::runtime_error("sampled_recall: not implemented");
}
void nndescent_host(int, const coresg_hip_nnd_config&, const float*, std::uint32_t*, float*,
                    std::uint8_t*, std::uint32_t*, std::uint64_t*) {
    throw std::runtime_error("nndescent: not implemented");
}
void mrd_block_host(int, const float*, std::size_t, std::uint32_t, std::uint32_t, int,
                    const float*, std::size_t, std::size_t, float*) {
    throw std::runtime_error("mrd_block: not implemented");
}
void mrd_row_min_host(int, const float*, std::size_t, std::uint32_t, std::uint32_t, int,
                      const float*, const std::uint32_t*, std::size_t, std::size_t, float*,
                      std::uint32_t*) {
    throw std::runtime_error("mrd_row_min: not implemented");
}
void release_thread_cache() {}

}  // namespace coresg_hip
