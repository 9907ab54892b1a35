// Test program: the reference's own CoreSG, loaded from a .coresg file twice --
// once with the host backend, once with HipBackend -- must give the same
// all_msts and exact_mst results whether the host computes them or the
// DeviceExtract capability routes them to the GPU.
//
//   extract_check [--exact] FILE.coresg MPTS...    exit 0: every result equal
//
// all_msts is always compared; exact_mst only with --exact.
//
// Compiled in place against the reference's sources by oracle/Makefile
// (target hip-extract-check); it runs on a machine with a GPU.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "coresg/coresg.hpp"
#include "device_extract.hpp"

using namespace coresg;

namespace {

bool same(const MstResult& a, const MstResult& b, const std::string& what) {
    bool ok = a.edges.size() == b.edges.size() && a.component == b.component &&
              a.num_components == b.num_components &&
              std::memcmp(&a.weight, &b.weight, sizeof(double)) == 0;
    for (std::size_t i = 0; ok && i < a.edges.size(); ++i)
        ok = a.edges[i].u == b.edges[i].u && a.edges[i].v == b.edges[i].v &&
             std::memcmp(&a.edges[i].w, &b.edges[i].w, sizeof(float)) == 0;
    if (!ok)
        std::fprintf(stderr, "MISMATCH %s: edges %zu vs %zu, components %u vs %u, weight %.9f vs %.9f\n",
                     what.c_str(), a.edges.size(), b.edges.size(), a.num_components,
                     b.num_components, a.weight, b.weight);
    return ok;
}

CoreSGParams params(bool gpu) {
    CoreSGParams p;
    p.threads = 1;
    p.verbose = false;
    p.backend.prefer_gpu = gpu;
    return p;
}

}  // namespace

int main(int argc, char** argv) {
    const bool with_exact = argc > 1 && std::strcmp(argv[1], "--exact") == 0;
    if (with_exact) {
        --argc;
        ++argv;
    }
    if (argc < 3) {
        std::fprintf(stderr, "usage: extract_check [--exact] FILE.coresg MPTS...\n");
        return 2;
    }
    std::vector<std::uint32_t> mpts;
    for (int i = 2; i < argc; ++i) mpts.push_back(std::strtoul(argv[i], nullptr, 10));
    CoreSG host(params(false));
    host.load(argv[1]);
    CoreSG dev(params(true));
    dev.load(argv[1]);
    if (!dynamic_cast<DeviceExtract*>(dev.backend())) {
        std::fprintf(stderr, "backend '%s' offers no DeviceExtract\n", dev.backend()->name().c_str());
        return 2;
    }
    const std::vector<MstResult> a = host.all_msts(mpts);
    const std::vector<MstResult> b = all_msts_routed(dev, mpts);
    bool ok = a.size() == b.size();
    for (std::size_t i = 0; ok && i < a.size(); ++i)
        ok = same(a[i], b[i], "mst(" + std::to_string(mpts[i]) + ")");
    // again, in reverse order: served by the device copy kept from the first call
    std::vector<std::uint32_t> rev(mpts.rbegin(), mpts.rend());
    const std::vector<MstResult> c = all_msts_routed(dev, rev);
    for (std::size_t i = 0; ok && i < a.size(); ++i)
        ok = same(a[i], c[a.size() - 1 - i], "mst(" + std::to_string(mpts[i]) + ") again");
    for (std::size_t i = 0; with_exact && ok && i < mpts.size(); ++i)
        ok = same(host.exact_mst(mpts[i]), exact_mst_routed(dev, mpts[i]),
                  "exact_mst(" + std::to_string(mpts[i]) + ")");
    std::printf("extract_check %s: %zu msts + %zu exact msts %s (backend %s)\n", argv[1],
                mpts.size(), with_exact ? mpts.size() : std::size_t(0), ok ? "equal" : "DIFFER",
                dev.backend()->name().c_str());
    return ok ? 0 : 1;
}
