// Test program: the hubness half of the reference's CoreSG::build through the
// DeviceExtract capability. A .coresg is loaded into the reference's own
// CoreSG (host backend), whose load() derives components, reverse degrees,
// scores and the anti-hub selection (io.cpp:316-328); HipBackend::hubness
// computes the same from the same k-NN lists on the GPU. Every value must be
// identical.
//
//   hubness_check FILE.coresg [MIN_PER_COMPONENT]    exit 0: identical
//
// Compiled in place against the reference's sources by oracle/Makefile
// (target hip-hubness-check); runs with a GPU.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "coresg/coresg.hpp"
#include "device_extract.hpp"

using namespace coresg;

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: hubness_check FILE.coresg [MIN_PER_COMPONENT]\n");
        return 2;
    }
    CoreSGParams hp;
    hp.threads = 1;
    hp.verbose = false;
    hp.backend.prefer_gpu = false;
    if (argc > 2) hp.hub.min_per_component = std::strtoul(argv[2], nullptr, 10);
    CoreSG host(hp);
    host.load(argv[1]);
    CoreSGParams dp = hp;
    dp.backend.prefer_gpu = true;
    std::unique_ptr<Backend> be = make_backend(dp.backend, nullptr);
    auto* dev = dynamic_cast<DeviceExtract*>(be.get());
    if (!dev) {
        std::fprintf(stderr, "backend '%s' offers no DeviceExtract\n", be->name().c_str());
        return 2;
    }
    const KnnGraph& knn = host.knn();
    const DeviceExtract::Hubness h = dev->hubness(knn, host.data().dim(), host.hubness().params());
    const idx_t n = knn.size();
    std::size_t bad_rdeg = 0, bad_score = 0, bad_label = 0;
    for (idx_t v = 0; v < n; ++v) {
        if (h.reverse_degree[v] != host.hubness().reverse_degree(v)) ++bad_rdeg;
        const double a = h.score[v], b = host.hubness().reverse_score(v);
        if (std::memcmp(&a, &b, sizeof(double)) != 0) ++bad_score;
        if (h.component[v] != host.connectivity().labels()[v]) ++bad_label;
    }
    const bool comps = h.components == host.connectivity().components();
    const bool sel = h.selected == host.hubness().selected();
    const bool ok = !bad_rdeg && !bad_score && !bad_label && comps && sel;
    std::printf("hubness_check %s: n=%u rdeg %zu bad, score %zu bad, labels %zu bad, "
                "components %u/%u, selected %zu/%zu %s -> %s (backend %s)\n",
                argv[1], unsigned(n), bad_rdeg, bad_score, bad_label, unsigned(h.components),
                unsigned(host.connectivity().components()), h.selected.size(),
                host.hubness().selected().size(), sel ? "equal" : "DIFFER",
                ok ? "identical" : "DIFFERENT", be->name().c_str());
    return ok ? 0 : 1;
}
