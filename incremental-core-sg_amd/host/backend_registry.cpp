// Backend factory with the HIP branch: the change a maintainer makes to the
// reference's src/backend.cpp (make_backend, :52-65) to register HipBackend.
//
// This translation unit stands in for src/backend.cpp when the reference's
// own host code is linked against libcoresg_hip.so (the drop-in integration
// build): the same four functions with the same behaviour, plus the HIP branch
// in make_backend -- a GPU is preferred when one is visible and prefer_gpu is
// set, otherwise the CPU backend.
#include <sstream>

#include "coresg/backend.hpp"
#include "coresg_hip.h"

namespace coresg {

std::unique_ptr<Backend> make_hip_backend(const std::vector<int>& device_ids, bool oom_fallback);

std::string describe(const DeviceInfo& d) {
    std::ostringstream oss;
    if (d.kind == DeviceKind::Cpu) {
        oss << "cpu (" << d.multiprocessors << " threads)";
    } else {
        oss << "hip:" << d.id << " " << d.name << " gfx" << d.cc_major << d.cc_minor << " "
            << (d.total_memory >> 20) << " MiB";
    }
    return oss.str();
}

// No CUDA runtime in this build.
bool cuda_available() { return false; }
std::vector<DeviceInfo> enumerate_cuda_devices() { return {}; }

std::unique_ptr<Backend> make_backend(const BackendOptions& opts, ThreadPool* cpu_pool) {
    const int count = coresg_hip_device_count();
    if (opts.prefer_gpu && count > 0) {
        std::vector<int> ids = opts.device_ids;
        if (ids.empty()) {
            for (int i = 0; i < count; ++i) ids.push_back(i);
        }
        return make_hip_backend(ids, opts.oom_fallback);
    }
    return make_cpu_backend_slots(opts.cpu_threads, opts.cpu_slots, cpu_pool);
}

}  // namespace coresg
