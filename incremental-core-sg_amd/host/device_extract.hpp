// DeviceExtract: a capability a Backend can offer beside the reference's three
// virtuals (include/coresg/backend.hpp:72-92). The reference keeps the whole
// extraction half of CoreSG on the host (coresg.cpp:391-430), and the hubness
// half of its build (coresg.cpp:278-300); a backend that also derives from
// DeviceExtract runs them on its device, and CoreSG's mst / all_msts /
// exact_mst (and the hubness step of build) route to it when
// dynamic_cast<DeviceExtract*>(backend()) succeeds. The CPU backend does not
// derive from it, so the host path stays as it is.
#pragma once

#include <cstdint>
#include <vector>

#include "coresg/coresg.hpp"
#include "coresg/hubness.hpp"

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

    // The hubness half of CoreSG::build (coresg.cpp:278-300) for a k-NN graph:
    // components (ConnectivityTracker::rebuild, labels by first appearance),
    // reverse degrees and scores (HubnessIndex::build, compute_score) and the
    // anti-hub selection with the per-component quota (rescore_all) under
    // `params` with `budget_dim` in the d*ln|E| budget. The same values the
    // host classes hold after that sequence.
    struct Hubness {
        std::vector<std::uint32_t> reverse_degree;
        std::vector<double> score;
        std::vector<idx_t> component;
        idx_t components = 0;
        std::vector<idx_t> selected;
    };
    virtual Hubness hubness(const KnnGraph& knn, std::uint32_t budget_dim,
                            const HubnessParams& params) = 0;
};

// The routing CoreSG::all_msts / CoreSG::exact_mst gain (the same lines a
// maintainer adds inside those member functions).
std::vector<MstResult> all_msts_routed(const CoreSG& sg, const std::vector<std::uint32_t>& mpts);
MstResult exact_mst_routed(const CoreSG& sg, std::uint32_t mpts);

}  // namespace coresg
