// Packed MRD edge key shared by the edge-list code and the Boruvka sweep.
#pragma once

#include "common.hpp"

namespace coresg_hip {

__device__ __forceinline__ std::uint64_t mrd_key(float w, float cu, float cv, std::uint32_t e) {
    // mrd = max(w, max(core[u], core[v])) exactly as std::max composes it.
    const float wr = fmax_ref(w, fmax_ref(cu, cv));
    return (std::uint64_t(fbits(wr)) << 32) | e;
}

// Set on an adjacency entry seen from the edge's second endpoint, so the MRD
// key composes max(core[u], core[v]) in the edge's own operand order.
constexpr std::uint32_t kSecondEnd = 0x80000000u;

}  // namespace coresg_hip
