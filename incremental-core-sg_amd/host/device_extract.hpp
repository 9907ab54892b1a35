// DeviceExtract: a capability a Backend can offer beside the reference's three
// virtuals (include/coresg/backend.hpp:72-92). The reference keeps the whole
// extraction half of CoreSG on the host (coresg.cpp:391-430); a backend that
// also derives from DeviceExtract runs it on its device, and CoreSG's
// mst / all_msts / exact_mst route to it when
// dynamic_cast<DeviceExtract*>(backend()) succeeds. The CPU backend does not
// derive from it, so the host path stays as it is.
#pragma once

#include <cstdint>
#include <vector>

#include "coresg/coresg.hpp"

namespace coresg {

class DeviceExtract {
public:
    virtual ~DeviceExtract() = default;
    // CoreSG::all_msts (coresg.cpp:412-417): one MstResult per value, the
    // same edges in the same (Kruskal) order, labels, component count and
    // weight as the host. mpts 0 reads as 1; mpts above k_max throws
    // (coresg.cpp:391-395).
    virtual std::vector<MstResult> extract_msts(const CoreSG& sg,
                                                const std::vector<std::uint32_t>& mpts) = 0;
    // CoreSG::exact_mst (coresg.cpp:419-430): Prim's edges in acceptance order.
    virtual MstResult exact_mst(const CoreSG& sg, std::uint32_t mpts) = 0;
};

// The routing CoreSG::all_msts / CoreSG::exact_mst gain (the same lines a
// maintainer adds inside those member functions).
std::vector<MstResult> all_msts_routed(const CoreSG& sg, const std::vector<std::uint32_t>& mpts);
MstResult exact_mst_routed(const CoreSG& sg, std::uint32_t mpts);

}  // namespace coresg
