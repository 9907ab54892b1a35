// HipBackend: the reference's Backend plugin (include/coresg/backend.hpp:72-92)
// served by libcoresg_hip.so through its C ABI (include/coresg_hip.h).
//
// A drop-in beside cpu_backend.cpp: compiled against the reference's own
// headers, it implements the same three virtuals with the same semantics as
// the reference's CUDA backend (src/cuda_backend.cpp:15-133):
//   * run_knn_job: Dynamic NN-Descent on the slot's GPU, graph refined in place
//     in metric units; metrics the device does not cover and instances below
//     512 points run on the host backend; a device out-of-memory is recorded
//     in the job and, unless oom_fallback is off, finished on the host; any
//     other device failure throws.
//   * mrd_block: rows [r0, r1) of the m x m mutual-reachability matrix,
//     bit-identical to the host backend's (same distance summation order).
//   * trim: releases the calling thread's resident MRD operands.
// It also offers DeviceExtract (device_extract.hpp): CoreSG's extraction half
// (all_msts, exact_mst) on the device, over the CORE-SG the caller holds; the
// device copy of that CORE-SG (points, lists, bridges, edge list) stays
// resident across calls while the CoreSG is unchanged.
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "coresg/backend.hpp"
#include "coresg_hip.h"
#include "device_extract.hpp"

namespace coresg {
namespace {

// Metrics the device kernels cover (cuda_backend.cpp:15-21): L2 and cosine
// (rows normalised by the host Dataset, so the cosine distance is half the
// squared Euclidean one); L1 and inner product run on the host.
bool metric_to_hip(MetricKind m, int* code) {
    switch (m) {
        case MetricKind::L2: *code = CORESG_HIP_METRIC_L2; return true;
        case MetricKind::Cosine: *code = CORESG_HIP_METRIC_COSINE; return true;
        default: return false;
    }
}

void hip_check(int ok, const char* what) {
    if (!ok) throw std::runtime_error(std::string(what) + ": " + coresg_hip_last_error());
}

// A device-resident copy of a host CORE-SG: points, k-NN lists and bridges.
struct DeviceGraph {
    coresg_hip_graph* g = nullptr;
    DeviceGraph(int device, const CoreSG& sg) {
        const Dataset& d = sg.data();
        const KnnGraph& knn = sg.knn();
        const auto n = static_cast<std::uint32_t>(knn.size());
        g = coresg_hip_graph_create(device, n, d.dim(), knn.k());
        hip_check(g != nullptr, "coresg_hip_graph_create");
        hip_check(coresg_hip_graph_set_points(g, d.data(), CORESG_HIP_HOST, 1), "set_points");
        hip_check(coresg_hip_graph_set_knn(g, knn.ids_storage().data(),
                                           knn.dist_storage().data(),
                                           knn.flag_storage().data(),
                                           knn.count_storage().data(), CORESG_HIP_HOST),
                  "set_knn");
        const std::vector<Edge>& b = sg.missing_edges();
        std::vector<std::uint32_t> u(b.size()), v(b.size());
        std::vector<float> w(b.size());
        for (std::size_t i = 0; i < b.size(); ++i) {
            u[i] = b[i].u;
            v[i] = b[i].v;
            w[i] = b[i].w;
        }
        hip_check(coresg_hip_graph_set_bridges(g, u.data(), v.data(), w.data(), b.size(),
                                               sg.repair_edge_count()),
                  "set_bridges");
    }
    ~DeviceGraph() {
        if (g) coresg_hip_graph_destroy(g);
    }
    DeviceGraph(const DeviceGraph&) = delete;
    DeviceGraph& operator=(const DeviceGraph&) = delete;
};

class HipBackend : public Backend, public DeviceExtract {
public:
    HipBackend(const std::vector<int>& device_ids, bool oom_fallback)
        : ids_(device_ids), oom_fallback_(oom_fallback) {
        if (ids_.empty()) throw std::runtime_error("hip backend: no device ids");
        for (int id : ids_) {
            coresg_hip_device_props p;
            if (!coresg_hip_get_device_props(id, &p)) {
                throw std::runtime_error(std::string("hip device ") + std::to_string(id) +
                                         " unavailable: " + coresg_hip_last_error());
            }
            DeviceInfo d;
            // The reference's DeviceKind names its only GPU kind `Cuda`; GPU
            // jobs report it so the reference's job accounting
            // (incremental_knn.cpp:1039-1040) counts them as device jobs.
            d.kind = DeviceKind::Cuda;
            d.id = p.id;
            d.name = std::string(p.name) + " (" + p.arch + ")";
            d.total_memory = p.total_mem;
            d.free_memory = p.free_mem;
            d.multiprocessors = p.cu_count;
            d.cc_major = p.major;
            d.cc_minor = p.minor;
            devices_.push_back(d);
        }
        fallback_ = make_cpu_backend_slots(0, static_cast<unsigned>(ids_.size()), nullptr);
    }

    std::string name() const override { return "hip"; }
    const std::vector<DeviceInfo>& devices() const override { return devices_; }
    unsigned parallelism() const override { return static_cast<unsigned>(ids_.size()); }

    void run_knn_job(unsigned slot, KnnJob& job) override {
        int code = 0;
        if (!metric_to_hip(job.metric, &code) || job.local_n < 512) {
            fallback_->run_knn_job(slot, job);
            return;
        }
        coresg_hip_nnd_config cfg;
        cfg.n = static_cast<std::uint32_t>(job.local_n);
        cfg.dim = job.dim;
        cfg.stride = job.stride;
        cfg.k = job.graph.k();
        cfg.sample = job.params.sample;
        cfg.iters = job.params.iters;
        // The host backend seeds init_random with params.seed
        // (cpu_backend.cpp:56); using the same seed keeps the device's
        // initial graph bit-identical to the host's.
        cfg.seed = job.params.seed;
        cfg.metric = code;
        cfg.delta = static_cast<float>(job.params.delta);
        cfg.random_fill = 1;

        job.graph.grow(static_cast<idx_t>(job.local_n));
        std::uint64_t updates = 0;
        const int device = device_for(slot);
        const int ok = coresg_hip_nndescent(
            device, &cfg, job.vectors.data(), job.graph.ids_storage().data(),
            job.graph.dist_storage().data(), job.graph.flag_storage().data(),
            job.graph.count_storage().data(), &updates);
        if (!ok) {
            const std::string err = coresg_hip_last_error();
            if (err.find("out of memory") != std::string::npos ||
                err.find("hipErrorOutOfMemory") != std::string::npos) {
                job.out_of_memory = true;
                job.error = err;
                if (!oom_fallback_) return;
                fallback_->run_knn_job(slot, job);
                return;
            }
            throw std::runtime_error(std::string("hip nndescent failed on device ") +
                                     std::to_string(device) + ": " + err);
        }
        job.ran_on = DeviceKind::Cuda;
        job.ran_on_device = device;
        job.stats.rounds = job.params.iters;
        job.stats.updates = static_cast<std::size_t>(updates);
    }

    void mrd_block(unsigned slot, const float* pts, std::size_t m, std::uint32_t dim,
                   std::uint32_t stride, MetricKind metric, const float* core, std::size_t r0,
                   std::size_t r1, float* out) override {
        int code = 0;
        if (!metric_to_hip(metric, &code)) {
            fallback_->mrd_block(slot, pts, m, dim, stride, metric, core, r0, r1, out);
            return;
        }
        if (!coresg_hip_mrd_block(device_for(slot), pts, m, dim, stride, code, core, r0, r1,
                                  out)) {
            throw std::runtime_error(std::string("hip mrd_block failed: ") +
                                     coresg_hip_last_error());
        }
    }

    void trim() override {
        coresg_hip_release_thread_cache();
        std::lock_guard<std::mutex> lk(resident_mu_);
        resident_.reset();
    }

    std::vector<MstResult> extract_msts(const CoreSG& sg,
                                        const std::vector<std::uint32_t>& mpts) override {
        const std::size_t n = sg.knn().size();
        if (n < 2 || mpts.empty()) return sg.all_msts(mpts);
        std::lock_guard<std::mutex> lk(resident_mu_);
        DeviceGraph& dg = resident(sg);
        const std::size_t q = mpts.size(), n1 = n - 1;
        std::vector<coresg_hip_mst_info> info(q);
        std::vector<std::uint32_t> u(q * n1), v(q * n1), lab(q * n);
        std::vector<float> w(q * n1);
        hip_check(coresg_hip_extract(dg.g, mpts.data(), static_cast<std::uint32_t>(q), info.data(),
                                     u.data(), v.data(), w.data(), lab.data()),
                  "coresg_hip_extract");
        std::vector<MstResult> out(q);
        for (std::size_t i = 0; i < q; ++i) {
            MstResult& r = out[i];
            r.edges.reserve(info[i].num_edges);
            for (std::size_t e = 0; e < info[i].num_edges; ++e)
                r.edges.emplace_back(u[i * n1 + e], v[i * n1 + e], w[i * n1 + e]);
            // Kruskal's acceptance order is the edge_less order of the tree.
            std::sort(r.edges.begin(), r.edges.end(), EdgeLess());
            r.component.assign(lab.begin() + i * n, lab.begin() + (i + 1) * n);
            r.num_components = info[i].num_components;
            r.weight = info[i].weight;
        }
        return out;
    }

    // CoreSG::exact_mst (coresg.cpp:419-430) on the resident device copy,
    // with one launch per Prim step: in a process driving the reference's
    // CoreSG through this adapter, the one-launch cooperative form once
    // returned hipErrorIllegalState after an extract_msts call, with no cause
    // shown, so the adapter does not use cooperative launches.
    MstResult exact_mst(const CoreSG& sg, std::uint32_t mpts) override {
        int code = 0;
        const std::size_t n = sg.knn().size();
        if (!metric_to_hip(sg.data().metric(), &code) || code != CORESG_HIP_METRIC_L2 || n < 2) {
            // A fresh start for the resident MRD operands: each call hands
            // mrd_block a new core-distance vector.
            coresg_hip_release_thread_cache();
            return sg.exact_mst(mpts);
        }
        std::lock_guard<std::mutex> lk(resident_mu_);
        DeviceGraph& dg = resident(sg);
        hip_check(coresg_hip_graph_set_exact_launch(dg.g, 1), "set_exact_launch");
        coresg_hip_mst_info info{};
        std::vector<std::uint32_t> u(n - 1), v(n - 1), lab(n);
        std::vector<float> w(n - 1);
        hip_check(coresg_hip_exact_mst(dg.g, mpts, &info, u.data(), v.data(), w.data(),
                                       lab.data()),
                  "coresg_hip_exact_mst");
        MstResult r;
        r.edges.reserve(info.num_edges);
        for (std::size_t e = 0; e < info.num_edges; ++e) r.edges.emplace_back(u[e], v[e], w[e]);
        r.component = std::move(lab);
        r.num_components = info.num_components;
        r.weight = info.weight;
        return r;
    }

    Hubness hubness(const KnnGraph& knn, std::uint32_t budget_dim,
                    const HubnessParams& params) override {
        Hubness out;
        const auto n = static_cast<std::uint32_t>(knn.size());
        if (n == 0) return out;
        // the device scores with the arithmetic mean (the default rule)
        if (params.normalization != ScoreNormalization::ArithmeticMean)
            throw std::runtime_error("hip hubness: only the arithmetic-mean score is on the device");
        coresg_hip_graph* g = coresg_hip_graph_create(device_for(0), n, budget_dim ? budget_dim : 1,
                                                      knn.k());
        hip_check(g != nullptr, "coresg_hip_graph_create");
        std::unique_ptr<coresg_hip_graph, void (*)(coresg_hip_graph*)> guard(
            g, coresg_hip_graph_destroy);
        hip_check(coresg_hip_graph_set_knn(g, knn.ids_storage().data(), knn.dist_storage().data(),
                                           knn.flag_storage().data(),
                                           knn.count_storage().data(), CORESG_HIP_HOST),
                  "set_knn");
        coresg_hip_build_params p;
        coresg_hip_default_build_params(&p);
        p.antihub_multiplier = params.antihub_multiplier;
        p.max_pair_ops = params.max_pair_ops;
        p.min_selected = params.min_selected;
        p.max_selected = params.max_selected;
        p.min_per_component = params.min_per_component;
        p.score_k = params.k_for_score;
        hip_check(coresg_hip_graph_derive_hubness(g, &p), "derive_hubness");
        std::vector<std::int32_t> rd(n);
        out.score.resize(n);
        out.component.resize(n);
        std::uint32_t nc = 0;
        hip_check(coresg_hip_graph_get_hubness(g, rd.data(), out.score.data(),
                                               out.component.data(), &nc),
                  "get_hubness");
        out.reverse_degree.assign(rd.begin(), rd.end());
        out.components = nc;
        out.selected.resize(coresg_hip_graph_get_selected(g, nullptr, 0));
        coresg_hip_graph_get_selected(g, out.selected.data(), out.selected.size());
        return out;
    }

private:
    int device_for(unsigned slot) const { return ids_[slot < ids_.size() ? slot : 0]; }

    // What identifies the CoreSG a resident device copy was made from: the
    // object, its sizes and storage, and a digest of the bridges and of a
    // stride of list entries (an insert changes n; a rebuilt bridge set
    // changes the digest).
    struct ResidentKey {
        const CoreSG* sg = nullptr;
        std::size_t n = 0, k = 0, dim = 0, bridges = 0, repair = 0;
        const void *pts = nullptr, *ids = nullptr, *dist = nullptr;
        std::uint64_t digest = 0;
        bool operator==(const ResidentKey& o) const {
            return sg == o.sg && n == o.n && k == o.k && dim == o.dim && bridges == o.bridges &&
                   repair == o.repair && pts == o.pts && ids == o.ids && dist == o.dist &&
                   digest == o.digest;
        }
    };

    static std::uint64_t mix(std::uint64_t h, std::uint64_t x) {
        h ^= x + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
        return h;
    }

    static ResidentKey key_of(const CoreSG& sg) {
        ResidentKey key;
        const KnnGraph& knn = sg.knn();
        key.sg = &sg;
        key.n = knn.size();
        key.k = knn.k();
        key.dim = sg.data().dim();
        key.bridges = sg.missing_edges().size();
        key.repair = sg.repair_edge_count();
        key.pts = sg.data().data();
        key.ids = knn.ids_storage().data();
        key.dist = knn.dist_storage().data();
        std::uint64_t h = 0;
        for (const Edge& e : sg.missing_edges()) {
            std::uint32_t wb;
            std::memcpy(&wb, &e.w, 4);
            h = mix(h, (std::uint64_t(e.u) << 32) | e.v);
            h = mix(h, wb);
        }
        const std::vector<idx_t>& ids = knn.ids_storage();
        const std::vector<dist_t>& dist = knn.dist_storage();
        for (std::size_t i = 0; i < ids.size(); i += 4099) {
            std::uint32_t db;
            std::memcpy(&db, &dist[i], 4);
            h = mix(h, (std::uint64_t(ids[i]) << 32) | db);
        }
        key.digest = h;
        return key;
    }

    // The device copy of `sg`, made on first use and kept across calls (its
    // edge list stays prepared) until the CoreSG changes or trim() is called.
    DeviceGraph& resident(const CoreSG& sg) {
        const ResidentKey key = key_of(sg);
        if (!resident_ || !(resident_key_ == key)) {
            resident_.reset();
            resident_.reset(new DeviceGraph(device_for(0), sg));
            resident_key_ = key;
        }
        return *resident_;
    }

    std::vector<int> ids_;
    bool oom_fallback_ = true;
    std::vector<DeviceInfo> devices_;
    std::unique_ptr<Backend> fallback_;
    std::mutex resident_mu_;
    std::unique_ptr<DeviceGraph> resident_;
    ResidentKey resident_key_;
};

}  // namespace

std::vector<MstResult> all_msts_routed(const CoreSG& sg, const std::vector<std::uint32_t>& mpts) {
    if (auto* dev = dynamic_cast<DeviceExtract*>(sg.backend())) return dev->extract_msts(sg, mpts);
    return sg.all_msts(mpts);
}

MstResult exact_mst_routed(const CoreSG& sg, std::uint32_t mpts) {
    if (auto* dev = dynamic_cast<DeviceExtract*>(sg.backend())) return dev->exact_mst(sg, mpts);
    return sg.exact_mst(mpts);
}

std::unique_ptr<Backend> make_hip_backend(const std::vector<int>& device_ids,
                                          bool oom_fallback) {
    return std::unique_ptr<Backend>(new HipBackend(device_ids, oom_fallback));
}

}  // namespace coresg
