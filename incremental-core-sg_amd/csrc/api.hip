// C ABI of the library: error convention, device queries, the resident graph.
#include <chrono>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "../../include/coresg_hip.h"
#include "build.hpp"
#include "graph.hpp"
#include "handle.hpp"
#include "scan.hpp"

namespace coresg_hip {

namespace {
thread_local std::string t_last_error;
}

void set_last_error(const std::string& msg) { t_last_error = msg; }
const char* get_last_error() { return t_last_error.c_str(); }

Graph::~Graph() {
    resolve_timers();
    if (stream) {
        (void)hipSetDevice(device);
        (void)hipStreamSynchronize(stream);
    }
    if (sweep_exec || descent_exec) {
        std::lock_guard<std::mutex> lk(device_api_mutex());
        if (sweep_exec) (void)hipGraphExecDestroy(sweep_exec);
        if (descent_exec) (void)hipGraphExecDestroy(descent_exec);
    }
    destroy_stream(stream);
}

void Graph::ensure_scan_tmp(std::size_t count) { scan_tmp.reserve(scan_tmp_words(count)); }

void Graph::timer_start(const char* name, hipEvent_t* start) {
    hipEvent_t e;
    CSG_CHECK(hipEventCreate(&e));
    CSG_CHECK(hipEventRecord(e, stream));
    *start = e;
    (void)name;
}

void Graph::timer_stop(const char* name, hipEvent_t start) {
    hipEvent_t e;
    CSG_CHECK(hipEventCreate(&e));
    CSG_CHECK(hipEventRecord(e, stream));
    KernelTimer& t = timers[name];
    t.pending.push_back(start);
    t.pending.push_back(e);
}

void Graph::max_work(const char* name, int index, std::uint64_t value) {
    KernelTimer& t = timers[name];
    if (index >= 0 && index < 4 && value > t.work[index]) t.work[index] = value;
}

void Graph::add_work(const char* name, std::initializer_list<std::uint64_t> w) {
    KernelTimer& t = timers[name];
    int i = 0;
    for (std::uint64_t x : w) {
        if (i < 4) t.work[i] += x;
        ++i;
    }
}

void Graph::resolve_timers() {
    for (auto& kv : timers) {
        KernelTimer& t = kv.second;
        for (std::size_t i = 0; i + 1 < t.pending.size(); i += 2) {
            float ms = 0.0f;
            if (hipEventSynchronize(t.pending[i + 1]) == hipSuccess &&
                hipEventElapsedTime(&ms, t.pending[i], t.pending[i + 1]) == hipSuccess) {
                t.total_ms += ms;
                ++t.launches;
            }
            (void)hipEventDestroy(t.pending[i]);
            (void)hipEventDestroy(t.pending[i + 1]);
        }
        t.pending.clear();
    }
}

}  // namespace coresg_hip

using namespace coresg_hip;

namespace {

// The graph's device current; in the plain form, launch checks of this entry
// point also wait for the graph's stream (reset when the entry point returns).
// Per-device budgets of the reference-shaped calls (coresg_hip_set_device_budget).
constexpr int kMaxBudgetDevices = 64;
std::mutex g_budget_mu;
std::size_t g_budget[kMaxBudgetDevices] = {};

// The calling thread's allocations are charged to `device`'s budget while in
// scope.
struct BudgetScope {
    AllocBudget b;
    explicit BudgetScope(int device) {
        std::lock_guard<std::mutex> lk(g_budget_mu);
        b.limit = (device >= 0 && device < kMaxBudgetDevices) ? g_budget[device] : 0;
        t_alloc_budget = b.limit ? &b : nullptr;
    }
    ~BudgetScope() { t_alloc_budget = nullptr; }
};

void bind(const Graph& g) {
    CSG_CHECK(hipSetDevice(g.device));
    t_plain_stream = g.plain ? g.stream : nullptr;
}

}  // namespace

extern "C" {

int coresg_hip_abi_version(void) { return CORESG_HIP_ABI_VERSION; }

const char* coresg_hip_last_error(void) { return get_last_error(); }

int coresg_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int coresg_hip_get_device_props(int device, coresg_hip_device_props* out) {
    return guarded([&] {
        hipDeviceProp_t p;
        CSG_CHECK(hipGetDeviceProperties(&p, device));
        std::memset(out, 0, sizeof(*out));
        out->id = device;
        std::strncpy(out->name, p.name, sizeof(out->name) - 1);
        std::strncpy(out->arch, p.gcnArchName, sizeof(out->arch) - 1);
        out->total_mem = p.totalGlobalMem;
        out->cu_count = p.multiProcessorCount;
        out->major = p.major;
        out->minor = p.minor;
        CSG_CHECK(hipSetDevice(device));
        std::size_t fr = 0, tot = 0;
        CSG_CHECK(hipMemGetInfo(&fr, &tot));
        out->free_mem = fr;
    });
}

int coresg_hip_device_memory(int device, size_t* free_bytes, size_t* total_bytes) {
    return guarded([&] {
        CSG_CHECK(hipSetDevice(device));
        CSG_CHECK(hipMemGetInfo(free_bytes, total_bytes));
    });
}

int coresg_hip_enable_peer_access(const int* devices, size_t count) {
    // One GPU per process: nothing to enable between a device and itself.
    return guarded([&] {
        if (count > 1) throw std::runtime_error("peer access: one device per process");
        (void)devices;
    });
}

void coresg_hip_default_build_params(coresg_hip_build_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->sample = 16;
    p->iters = 10;
    p->delta = 0.001;
    p->seed = 42;
    p->brute_force_limit = 0;
    p->antihub_multiplier = 1.0;
    p->max_pair_ops = 2e9;
    p->min_selected = 2;
    p->max_selected = 0;
    p->min_per_component = 2;
    p->ensure_connected = 1;
    p->reps_per_component = 4;
    p->max_reps = 4096;
    p->join_capacity = 0;
    p->gram = 0;
}

coresg_hip_graph* coresg_hip_graph_create(int device, uint32_t n, uint32_t dim, uint32_t k_max) {
    coresg_hip_graph* h = nullptr;
    const int ok = guarded([&] {
        if (k_max == 0) throw std::runtime_error("k_max must be positive");
        if (dim == 0) throw std::runtime_error("dim must be positive");
        h = new coresg_hip_graph();
        Graph& g = h->g;
        g.device = device;
        g.n = n;
        g.dim = dim;
        g.stride = (dim + 7u) & ~7u;
        g.k = k_max;
        CSG_CHECK(hipSetDevice(device));
        create_stream(&g.stream);
        g.pts.reserve(std::size_t(n) * g.stride);
        g.ids.reserve(std::size_t(n) * k_max);
        g.dist.reserve(std::size_t(n) * k_max);
        g.flags.reserve(std::size_t(n) * k_max);
        g.count.reserve(n);
    });
    if (!ok) {
        delete h;
        return nullptr;
    }
    return h;
}

void coresg_hip_graph_destroy(coresg_hip_graph* h) {
    if (!h) return;
    (void)hipSetDevice(h->g.device);
    delete h;
}

int coresg_hip_graph_set_points(coresg_hip_graph* h, const float* rows, int where, int padded) {
    return guarded([&] {
        Graph& g = h->g;
        bind(g);
        const hipMemcpyKind kind =
            where == CORESG_HIP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        if (padded || g.dim == g.stride) {
            CSG_CHECK(hipMemcpyAsync(g.pts.p, rows, std::size_t(g.n) * g.stride * 4, kind,
                                     g.stream));
        } else {
            // zero padding, then a strided 2-D copy of the logical columns
            fill_u32(reinterpret_cast<std::uint32_t*>(g.pts.p), std::size_t(g.n) * g.stride, 0u,
                     g.stream);
            CSG_CHECK(hipMemcpy2DAsync(g.pts.p, std::size_t(g.stride) * 4, rows,
                                       std::size_t(g.dim) * 4, std::size_t(g.dim) * 4, g.n, kind,
                                       g.stream));
        }
        CSG_CHECK(hipStreamSynchronize(g.stream));
        g.points_set = true;
        g.edges_ready = false;
        g.emit_ready = false;
    });
}

int coresg_hip_graph_set_knn(coresg_hip_graph* h, const uint32_t* ids, const float* dist,
                             const uint8_t* flags, const uint32_t* count, int where) {
    return guarded([&] {
        Graph& g = h->g;
        bind(g);
        const hipMemcpyKind kind =
            where == CORESG_HIP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        const std::size_t nk = std::size_t(g.n) * g.k;
        CSG_CHECK(hipMemcpyAsync(g.ids.p, ids, nk * 4, kind, g.stream));
        CSG_CHECK(hipMemcpyAsync(g.dist.p, dist, nk * 4, kind, g.stream));
        if (flags) {
            CSG_CHECK(hipMemcpyAsync(g.flags.p, flags, nk, kind, g.stream));
        } else {
            fill_u8(g.flags.p, nk, 0, g.stream);
        }
        CSG_CHECK(hipMemcpyAsync(g.count.p, count, std::size_t(g.n) * 4, kind, g.stream));
        CSG_CHECK(hipStreamSynchronize(g.stream));
        g.knn_set = true;
        g.edges_ready = false;
        g.emit_ready = false;
        g.hub_ready = false;
    });
}

int coresg_hip_graph_get_knn(coresg_hip_graph* h, uint32_t* ids, float* dist, uint8_t* flags,
                             uint32_t* count) {
    return guarded([&] {
        Graph& g = h->g;
        bind(g);
        const std::size_t nk = std::size_t(g.n) * g.k;
        if (ids) CSG_CHECK(hipMemcpyAsync(ids, g.ids.p, nk * 4, hipMemcpyDeviceToHost, g.stream));
        if (dist)
            CSG_CHECK(hipMemcpyAsync(dist, g.dist.p, nk * 4, hipMemcpyDeviceToHost, g.stream));
        if (flags)
            CSG_CHECK(hipMemcpyAsync(flags, g.flags.p, nk, hipMemcpyDeviceToHost, g.stream));
        if (count)
            CSG_CHECK(hipMemcpyAsync(count, g.count.p, std::size_t(g.n) * 4,
                                     hipMemcpyDeviceToHost, g.stream));
        CSG_CHECK(hipStreamSynchronize(g.stream));
    });
}

int coresg_hip_graph_set_bridges(coresg_hip_graph* h, const uint32_t* u, const uint32_t* v,
                                 const float* w, size_t m, size_t repair_count) {
    return guarded([&] {
        Graph& g = h->g;
        g.bu.assign(u, u + m);
        g.bv.assign(v, v + m);
        g.bw.assign(w, w + m);
        g.repair_count = repair_count;
        g.edges_ready = false;
    });
}

size_t coresg_hip_graph_get_bridges(coresg_hip_graph* h, uint32_t* u, uint32_t* v, float* w,
                                    size_t cap, size_t* repair_count) {
    Graph& g = h->g;
    const std::size_t m = g.bu.size();
    for (std::size_t i = 0; i < m && i < cap; ++i) {
        if (u) u[i] = g.bu[i];
        if (v) v[i] = g.bv[i];
        if (w) w[i] = g.bw[i];
    }
    if (repair_count) *repair_count = g.repair_count;
    return m;
}

int coresg_hip_graph_device_arrays(coresg_hip_graph* h, float** pts, uint32_t** ids,
                                   float** dist, uint8_t** flags, uint32_t** count) {
    Graph& g = h->g;
    if (pts) *pts = g.pts.p;
    if (ids) *ids = g.ids.p;
    if (dist) *dist = g.dist.p;
    if (flags) *flags = g.flags.p;
    if (count) *count = g.count.p;
    return 1;
}

size_t coresg_hip_graph_device_bytes(const coresg_hip_graph* h) {
    const Graph& g = h->g;
    return g.pts.bytes() + g.ids.bytes() + g.dist.bytes() + g.flags.bytes() + g.count.bytes() +
           g.eu.bytes() + g.ev.bytes() + g.ew.bytes() + g.eoff.bytes() + g.core.bytes() +
           g.comp.bytes() + g.parent.bytes() + g.newroot.bytes() + g.minv.bytes() +
           g.labels_tmp.bytes() + g.best.bytes() + g.mst_keys.bytes() + g.active.bytes() +
           g.sweep_vkey.bytes() + g.sweep_vfar.bytes() + g.adj_off.bytes() + g.adj.bytes() +
           g.mst_flag.bytes() + g.scan_tmp.bytes() + g.out_pos.bytes() + g.slot_part.bytes() + g.slot_out.bytes() +
           g.slot_mpts.bytes() + g.label_buf.bytes() + g.out_u.bytes() + g.out_v.bytes() +
           g.out_w.bytes() +
           g.rdeg.bytes() + g.score.bytes() + g.conn.bytes();
}

int coresg_hip_prepare_edges(coresg_hip_graph* h, uint64_t* num_edges, uint64_t* knn_edges) {
    return guarded([&] {
        Graph& g = h->g;
        bind(g);
        prepare_edges(g);
        if (num_edges) *num_edges = g.num_edges;
        if (knn_edges) *knn_edges = g.knn_edges;
    });
}

int coresg_hip_edges(coresg_hip_graph* h, uint32_t mpts, uint32_t* u, uint32_t* v, float* w) {
    return guarded([&] {
        bind(h->g);
        reweighted_edges(h->g, mpts, u, v, w);
    });
}

int coresg_hip_extract(coresg_hip_graph* h, const uint32_t* mpts, uint32_t count,
                       coresg_hip_mst_info* info, uint32_t* out_u, uint32_t* out_v, float* out_w,
                       uint32_t* out_labels) {
    return guarded([&] {
        bind(h->g);
        extract_msts(h->g, mpts, count, info, out_u, out_v, out_w, out_labels);
    });
}

int coresg_hip_exact_mst(coresg_hip_graph* h, uint32_t mpts, coresg_hip_mst_info* info,
                         uint32_t* out_u, uint32_t* out_v, float* out_w, uint32_t* out_labels) {
    return guarded([&] {
        bind(h->g);
        exact_mst(h->g, mpts, info, out_u, out_v, out_w, out_labels);
    });
}

int coresg_hip_graph_derive_hubness(coresg_hip_graph* h, const coresg_hip_build_params* p) {
    return guarded([&] {
        bind(h->g);
        derive_hubness(h->g, *p);
    });
}

int coresg_hip_autotune_sweep(coresg_hip_graph* h, const uint32_t* mpts, uint32_t count,
                              coresg_hip_tune_result* out) {
    return guarded([&] {
        bind(h->g);
        autotune_sweep(h->g, mpts, count, out);
    });
}

size_t coresg_hip_tune_table(char* buf, size_t cap) {
    std::string t;
    guarded_host([&] { t = tune_table_text(); });
    if (buf && cap) {
        const std::size_t m = t.size() < cap - 1 ? t.size() : cap - 1;
        std::memcpy(buf, t.data(), m);
        buf[m] = '\0';
    }
    return t.size();
}

int coresg_hip_tune_table_load(const char* text) {
    return guarded_host([&] { tune_table_load(text); });
}

int coresg_hip_graph_set_plain(coresg_hip_graph* h, int on) {
    return guarded([&] { h->g.plain = on != 0; });
}

int coresg_hip_graph_set_exact_launch(coresg_hip_graph* h, int mode) {
    return guarded_host([&] {
        if (mode != 0 && mode != 1) throw std::runtime_error("exact launch mode must be 0 or 1");
        h->g.exact_stepped = mode == 1;
    });
}

int coresg_hip_graph_set_profiling(coresg_hip_graph* h, int on) {
    return guarded([&] {
        h->g.resolve_timers();
        h->g.profiling = on != 0;
    });
}

int coresg_hip_graph_kernel_times(coresg_hip_graph* h, const char* kernel, double* avg_ms,
                                  uint64_t* launches) {
    return guarded([&] {
        Graph& g = h->g;
        bind(g);
        g.resolve_timers();
        auto it = g.timers.find(kernel);
        if (it == g.timers.end()) {
            if (avg_ms) *avg_ms = 0.0;
            if (launches) *launches = 0;
            return;
        }
        if (avg_ms)
            *avg_ms = it->second.launches ? it->second.total_ms / double(it->second.launches)
                                          : 0.0;
        if (launches) *launches = it->second.launches;
    });
}

int coresg_hip_graph_kernel_work(coresg_hip_graph* h, const char* kernel, uint64_t* work4) {
    return guarded([&] {
        auto it = h->g.timers.find(kernel);
        for (int i = 0; i < 4; ++i)
            work4[i] = it == h->g.timers.end() ? 0 : it->second.work[i];
    });
}

int coresg_hip_graph_set_instances(coresg_hip_graph* h, const uint32_t* offsets,
                                   uint32_t count) {
    return guarded([&] {
        if (!h) throw std::runtime_error("set_instances: null graph");
        Graph& g = h->g;
        if (count == 0) {
            g.inst_off.clear();
        } else {
            if (!offsets) throw std::runtime_error("set_instances: null offsets");
            if (offsets[0] != 0 || offsets[count] != g.n)
                throw std::runtime_error("set_instances: offsets must run from 0 to the point count");
            for (uint32_t i = 0; i < count; ++i)
                if (offsets[i + 1] <= offsets[i])
                    throw std::runtime_error("set_instances: every instance needs a point");
            g.inst_off.assign(offsets, offsets + count + 1);
            std::vector<uint32_t> of(g.n);
            for (uint32_t i = 0; i < count; ++i)
                for (uint32_t v = offsets[i]; v < offsets[i + 1]; ++v) of[v] = i;
            g.inst_of.reserve(g.n);
            g.inst_off_d.reserve(std::size_t(count) + 1);
            CSG_CHECK(hipMemcpyAsync(g.inst_of.p, of.data(), std::size_t(g.n) * 4,
                                     hipMemcpyHostToDevice, g.stream));
            CSG_CHECK(hipMemcpyAsync(g.inst_off_d.p, offsets, (std::size_t(count) + 1) * 4,
                                     hipMemcpyHostToDevice, g.stream));
            CSG_CHECK(hipStreamSynchronize(g.stream));
        }
        g.edges_ready = false;
        g.emit_ready = false;
        g.hub_ready = false;
    });
}

int coresg_hip_extract_instances(coresg_hip_graph* h, const uint32_t* mpts, uint32_t count,
                                 coresg_hip_mst_info* info) {
    return guarded([&] {
        if (!h || !mpts || !info) throw std::runtime_error("extract_instances: null argument");
        extract_instances(h->g, mpts, count, info);
    });
}

int coresg_hip_graph_set_audit(coresg_hip_graph* h, int on) {
    return guarded([&] {
        if (!h) throw std::runtime_error("set_audit: null graph");
        h->g.audit = on != 0;
    });
}

int coresg_hip_graph_reset_timers(coresg_hip_graph* h) {
    return guarded([&] {
        h->g.resolve_timers();
        h->g.timers.clear();
    });
}

int coresg_hip_insert(coresg_hip_graph* h, const float* rows, uint32_t count, int where,
                      int padded, const coresg_hip_build_params* p, coresg_hip_build_stats* st) {
    return guarded([&] {
        bind(h->g);
        if (st) std::memset(st, 0, sizeof(*st));
        insert_points(h->g, rows, count, where, padded, *p, st);
    });
}

uint32_t coresg_hip_graph_size(const coresg_hip_graph* h) { return h ? h->g.n : 0; }

int coresg_hip_build_knn(coresg_hip_graph* h, const coresg_hip_build_params* p,
                         coresg_hip_build_stats* st) {
    return guarded([&] {
        bind(h->g);
        build_knn(h->g, *p, st);
    });
}

int coresg_hip_build_bridges(coresg_hip_graph* h, const coresg_hip_build_params* p,
                             coresg_hip_build_stats* st) {
    return guarded([&] {
        bind(h->g);
        build_bridges(h->g, *p, st);
    });
}

int coresg_hip_build(coresg_hip_graph* h, const coresg_hip_build_params* p,
                     coresg_hip_build_stats* st) {
    return guarded([&] {
        bind(h->g);
        const auto t0 = std::chrono::steady_clock::now();
        build_knn(h->g, *p, st);
        build_bridges(h->g, *p, st);
        if (st) {
            st->total_seconds =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
    });
}

size_t coresg_hip_graph_get_selected(coresg_hip_graph* h, uint32_t* out, size_t cap) {
    const Graph& g = h->g;
    for (std::size_t i = 0; i < g.selected.size() && i < cap; ++i) out[i] = g.selected[i];
    return g.selected.size();
}

int coresg_hip_graph_get_hubness(coresg_hip_graph* h, int32_t* rdeg, double* score,
                                 uint32_t* conn_labels, uint32_t* num_components) {
    return guarded([&] {
        Graph& g = h->g;
        bind(g);
        if (!g.hub_ready) throw std::runtime_error("hubness not computed (call build_bridges)");
        const std::size_t n = g.n;
        if (rdeg) CSG_CHECK(hipMemcpyAsync(rdeg, g.rdeg.p, n * 4, hipMemcpyDeviceToHost, g.stream));
        if (score)
            CSG_CHECK(hipMemcpyAsync(score, g.score.p, n * 8, hipMemcpyDeviceToHost, g.stream));
        if (conn_labels)
            CSG_CHECK(
                hipMemcpyAsync(conn_labels, g.conn.p, n * 4, hipMemcpyDeviceToHost, g.stream));
        CSG_CHECK(hipStreamSynchronize(g.stream));
        if (num_components) *num_components = g.num_components;
    });
}

int coresg_hip_sampled_recall(coresg_hip_graph* h, size_t queries, uint64_t seed,
                              double* recall, double* distance_ratio) {
    return guarded([&] {
        bind(h->g);
        sampled_recall(h->g, queries, seed, recall, distance_ratio);
    });
}

int coresg_hip_set_device_budget(int device, size_t bytes) {
    return guarded([&] {
        if (device < 0 || device >= kMaxBudgetDevices)
            throw std::runtime_error("set_device_budget: device id out of range");
        std::lock_guard<std::mutex> lk(g_budget_mu);
        g_budget[device] = bytes;
    });
}

int coresg_hip_nndescent(int device, const coresg_hip_nnd_config* cfg, const float* host_vectors,
                         uint32_t* host_ids, float* host_dists, uint8_t* host_flags,
                         uint32_t* host_counts, uint64_t* out_updates) {
    return guarded([&] {
        CSG_CHECK(hipSetDevice(device));
        BudgetScope budget(device);
        nndescent_host(device, *cfg, host_vectors, host_ids, host_dists, host_flags, host_counts,
                       out_updates);
    });
}

int coresg_hip_mrd_block(int device, const float* host_pts, size_t m, uint32_t dim,
                         uint32_t stride, int metric, const float* host_core, size_t r0,
                         size_t r1, float* host_out) {
    return guarded([&] {
        CSG_CHECK(hipSetDevice(device));
        mrd_block_host(device, host_pts, m, dim, stride, metric, host_core, r0, r1, host_out);
    });
}

int coresg_hip_mrd_row_min(int device, const float* host_pts, size_t m, uint32_t dim,
                           uint32_t stride, int metric, const float* host_core,
                           const uint32_t* host_component, size_t r0, size_t r1,
                           float* host_min, uint32_t* host_argmin) {
    return guarded([&] {
        CSG_CHECK(hipSetDevice(device));
        mrd_row_min_host(device, host_pts, m, dim, stride, metric, host_core, host_component, r0,
                         r1, host_min, host_argmin);
    });
}

void coresg_hip_release_thread_cache(void) { release_thread_cache(); }

}  // extern "C"
