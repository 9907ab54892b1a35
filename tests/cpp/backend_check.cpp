// Test program: the reference's own CoreSG::build through HipBackend, for the
// plugin semantics the reference's CUDA backend defines (cuda_backend.cpp:15-133):
//
//   backend_check small      n < 512: the k-NN job runs on the host backend
//                            (cuda_backend.cpp:62-67); lists as good as a host
//                            build's
//   backend_check oom        a device budget too small for the job: the job is
//                            recorded out of memory and finished on the host
//                            (cuda_backend.cpp:88-101); lists as good as a host
//                            build's
//   backend_check oom-stop   the same with oom_fallback off: recorded, not run
//   backend_check cosine     the cosine metric on the device: every stored
//                            distance is the reference's metric_distance of its
//                            pair, recall as good as a host build's
//
// Exit 0 when every check holds. Compiled in place against the reference's
// sources by oracle/Makefile (target hip-backend-check); runs with a GPU.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "coresg/coresg.hpp"
#include "coresg/eval.hpp"
#include "coresg_hip.h"

using namespace coresg;

namespace {

int failures = 0;

void check(bool ok, const std::string& what) {
    std::printf("  [%s] %s\n", ok ? "ok" : "FAIL", what.c_str());
    if (!ok) ++failures;
}

std::vector<float> mixture(std::size_t n, std::uint32_t d, std::uint64_t seed) {
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> g(0.0f, 1.0f);
    std::vector<std::vector<float>> centres(8, std::vector<float>(d));
    for (auto& c : centres)
        for (float& x : c) x = 6.0f * g(rng);
    std::vector<float> out(n * d);
    for (std::size_t i = 0; i < n; ++i) {
        const auto& c = centres[rng() % centres.size()];
        for (std::uint32_t j = 0; j < d; ++j) out[i * d + j] = c[j] + g(rng);
    }
    return out;
}

CoreSGParams params(bool gpu, std::uint32_t k, MetricKind metric, bool oom_fallback = true) {
    CoreSGParams p;
    p.threads = 1;
    p.verbose = false;
    p.k_max = k;
    p.metric = metric;
    p.backend.prefer_gpu = gpu;
    p.backend.oom_fallback = oom_fallback;
    p.incr.instances = 1;
    p.incr.instance_nnd.sample = 16;
    p.incr.instance_nnd.iters = 10;
    return p;
}

double recall_of(const CoreSG& sg, std::uint32_t k) {
    const KnnGraph exact = brute_force_graph(sg.data(), k, nullptr);
    return knn_recall(sg.knn(), exact, k).recall;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "";
    const std::uint32_t d = 12, k = 12;
    if (mode == "small") {
        const std::size_t n = 400;
        const std::vector<float> x = mixture(n, d, 1);
        CoreSG host(params(false, k, MetricKind::L2)), dev(params(true, k, MetricKind::L2));
        host.build(x.data(), n, d);
        const CoreSGStats st = dev.build(x.data(), n, d);
        check(dev.backend()->name() == "hip", "backend is hip");
        check(st.knn_stats.jobs_on_device == 0 && st.knn_stats.jobs_on_host == 1,
              "the 400-point job ran on the host");
        // the fallback is the host backend on all host threads (as the
        // reference's CUDA backend builds it), whose descent schedule differs
        // from a one-thread build: compared by recall
        const double rd = recall_of(dev, k), rh = recall_of(host, k);
        std::printf("  recall: fallback %.4f, host %.4f\n", rd, rh);
        check(rd >= rh - 0.01, "the host fallback's lists are as good as a host build's");
    } else if (mode == "oom" || mode == "oom-stop") {
        const bool fallback = mode == "oom";
        const std::size_t n = 6000;
        const std::vector<float> x = mixture(n, d, 2);
        if (!coresg_hip_set_device_budget(0, std::size_t(1) << 20)) {
            std::printf("set_device_budget failed: %s\n", coresg_hip_last_error());
            return 2;
        }
        CoreSG dev(params(true, k, MetricKind::L2, fallback));
        const CoreSGStats st = dev.build(x.data(), n, d);
        coresg_hip_set_device_budget(0, 0);
        check(st.knn_stats.jobs_out_of_memory == 1, "the job is recorded out of memory");
        check(st.knn_stats.jobs_on_device == 0, "no job ran on the device");
        if (fallback) {
            CoreSG host(params(false, k, MetricKind::L2));
            host.build(x.data(), n, d);
            const double rd = recall_of(dev, k), rh = recall_of(host, k);
            std::printf("  recall: fallback %.4f, host %.4f\n", rd, rh);
            check(rd >= rh - 0.01, "the host fallback's lists are as good as a host build's");
        } else {
            std::size_t filled = 0;
            for (std::uint32_t c : dev.knn().count_storage()) filled += c;
            check(filled == 0, "with oom_fallback off the job is not run");
        }
        // the budget is lifted: the same job now runs on the device
        CoreSG again(params(true, k, MetricKind::L2, fallback));
        const CoreSGStats st2 = again.build(x.data(), n, d);
        check(st2.knn_stats.jobs_on_device == 1 && st2.knn_stats.jobs_out_of_memory == 0,
              "without the budget the job runs on the device");
    } else if (mode == "cosine") {
        const std::size_t n = 8000;
        const std::vector<float> x = mixture(n, d, 3);
        CoreSG host(params(false, k, MetricKind::Cosine)), dev(params(true, k, MetricKind::Cosine));
        host.build(x.data(), n, d);
        const CoreSGStats st = dev.build(x.data(), n, d);
        check(st.knn_stats.jobs_on_device == 1, "the cosine job ran on the device");
        const KnnGraph& g = dev.knn();
        std::size_t bad = 0, entries = 0;
        for (idx_t v = 0; v < g.size(); ++v) {
            for (std::uint32_t j = 0; j < g.degree(v); ++j) {
                const idx_t q = g.ids(v)[j];
                const float want = metric_distance(MetricKind::Cosine, dev.data().row(v),
                                                   dev.data().row(q), dev.data().stride());
                if (std::memcmp(&want, &g.dists(v)[j], 4) != 0) ++bad;
                if (j > 0 && g.dists(v)[j - 1] > g.dists(v)[j]) ++bad;
                ++entries;
            }
        }
        check(entries > 0 && bad == 0, "every stored cosine distance is the host's, in order (" +
                                           std::to_string(entries) + " entries)");
        const double rd = recall_of(dev, k), rh = recall_of(host, k);
        std::printf("  recall: device %.4f, host %.4f\n", rd, rh);
        check(rd >= rh - 0.01, "device cosine recall as good as the host's");
        // MRD rows for cosine through the backend equal the host backend's
        std::vector<float> core(200);
        for (std::size_t i = 0; i < core.size(); ++i) core[i] = 0.01f * float(i % 7);
        std::vector<float> a(core.size() * core.size()), b(a.size());
        dev.backend()->mrd_block(0, dev.data().data(), core.size(), d, dev.data().stride(),
                                 MetricKind::Cosine, core.data(), 0, core.size(), a.data());
        host.backend()->mrd_block(0, host.data().data(), core.size(), d, host.data().stride(),
                                  MetricKind::Cosine, core.data(), 0, core.size(), b.data());
        check(std::memcmp(a.data(), b.data(), a.size() * 4) == 0,
              "cosine mrd_block rows equal the host backend's");
    } else {
        std::fprintf(stderr, "usage: backend_check small|oom|oom-stop|cosine\n");
        return 2;
    }
    std::printf("backend_check %s: %s\n", mode.c_str(), failures ? "FAILED" : "passed");
    return failures ? 1 : 0;
}
