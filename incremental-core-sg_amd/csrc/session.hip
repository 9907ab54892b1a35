// Serving session: one device-resident CORE-SG served by continuous batching.
//
// Replaces the reference's one-shot `coresg_cli update` / `extract` round
// trips (coresg_cli.cpp:507-570), which reload the CORE-SG from disk for every
// command, with a long-lived graph in HBM and two device workers:
//
//   * the insert worker applies CoreSG::insert (coresg.cpp:325-389) to the
//     working graph, one batch at a time in submission order, lays out the new
//     version's edge list and adjacency, and publishes them as an
//     extraction snapshot (the edge arrays move into the snapshot; only the
//     k-NN distances and counts are copied);
//   * the extract worker serves CoreSG::mst for extract requests
//     (coresg.cpp:408-417) on the latest published snapshot, on its own
//     stream, so sweeps run on the device while the next version is built.
//     Every extract waiting when the worker becomes free is admitted into the
//     next fused sweep (up to kBatchSlots mpts values): requests join as
//     slots free up. Each slot of a sweep is computed independently, so a
//     result is the same alone and in a mix; it reports the version it read.
//
// Inserts are admitted against the session's HBM budget by an analytic model
// of the device footprint of the next version (the reference plans device
// memory per point the same way, resource.cpp:72-84, here extended with the
// edge list, the adjacency, the sweep state and the snapshot). An insert that
// does not fit waits while older snapshots that extracts still hold are
// released, and fails only when it cannot fit even then.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/coresg_hip.h"
#include "build.hpp"
#include "graph.hpp"
#include "handle.hpp"

using namespace coresg_hip;

namespace {

constexpr std::uint32_t kBatchSlots = 32;  // mpts values per fused sweep
constexpr double kYieldMs = 15.0;           // longest pause of an insert for extracts

using Clock = std::chrono::steady_clock;

struct Request {
    std::uint64_t id = 0;
    int kind = 0;
    // insert
    std::vector<float> rows;
    std::uint32_t count = 0;
    // extract
    std::vector<std::uint32_t> mpts;
    bool want_edges = false, want_labels = false;
    std::vector<coresg_hip_mst_info> info;
    std::vector<std::uint32_t> u, v, labels;
    std::vector<float> w;
    // outcome
    int state = CORESG_HIP_REQ_QUEUED;
    std::string error;
    coresg_hip_build_stats insert_stats{};
    std::uint32_t graph_n = 0;
    std::uint64_t graph_version = 0;
    std::uint32_t batch_requests = 0;
    Clock::time_point t_submit, t_start, t_done;
};

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Device bytes of a served version with n points and E edges: the working
// graph (points, k-NN lists, hubness state, ownership flags, the insert scan's
// per-chunk lists, the edge list and adjacency while they are laid out) and
// its extraction snapshot (distances and counts, edge list, adjacency, a
// 32-slot sweep's per-vertex state and the labels/edge outputs), and the
// descent's buffers for the refinement rounds (proposal buffer of 512 keys,
// sampled and reverse lists of smp entries each, gathered members and
// thresholds, reverse-list pairs). Per point / per edge, in bytes; an upper
// bound of what the arrays take.
std::size_t version_bytes(std::size_t n, std::size_t E, std::uint32_t stride, std::uint32_t k,
                          std::uint32_t smp) {
    const std::size_t per_point =
        4 * stride                    // points
        + 9 * k + 4                   // ids, dist, flags, count
        + 4 + 8 + 4                   // rdeg, score, component labels
        + k                           // ownership flags
        + 8 * std::size_t(k) * 16     // insert scan: per-chunk best lists
        + 3 * 4                       // edge offsets, adjacency offsets, cursors
        + (4 * k + 4)                 // snapshot: distances, counts
        + 2 * 4                       // snapshot: offsets
        + 32 * (4 + 4 + 4 + 4 + 8 + 8 + 1)  // sweep: core, comp, parent, newroot, best, key, active
        + (8 + 4)                     // one-slot sweeps: each vertex's best edge (key, far end)
        + 4 * 4                       // labels and scan scratch
        + 8 * 512 + 4                 // descent: proposal buffer, fill counts
        + 4 * 6 * std::size_t(smp)    // new, old (3 smp), reverse new and old lists
        + 8 * 6 * std::size_t(smp)    // gathered member ids and thresholds (6 smp)
        + 8 * std::size_t(k)          // reverse-list (target, source) pairs
        + 7 * 4;                      // list counts, join threshold
    const std::size_t per_edge = 12    // (u, v, w) of the sorted list
                                 + 16  // scattered records while laying it out
                                 + 24  // adjacency, both directions
                                 + 1 + 4 + 12;  // MST flags, positions, outputs
    return n * per_point + E * per_edge;
}

}  // namespace

struct coresg_hip_session;

// The arrays one graph version needs for extraction, taken from the working
// graph when the version is published.
struct Snapshot {
    coresg_hip_session* owner = nullptr;
    coresg_hip_graph* graph = nullptr;  // extraction-only graph (owned)
    std::uint64_t version = 0;
    std::uint32_t n = 0;
    std::size_t bytes = 0;
    ~Snapshot();
};

struct coresg_hip_session {
    coresg_hip_graph* graph = nullptr;  // the working graph, owned
    coresg_hip_build_params params{};
    std::size_t hbm_budget = 0;
    std::mutex mu;
    std::condition_variable cv_insert, cv_extract, cv_done, cv_mem;
    std::deque<std::shared_ptr<Request>> inserts, extracts;
    std::unordered_map<std::uint64_t, std::shared_ptr<Request>> requests;
    std::shared_ptr<Snapshot> published;
    std::uint64_t next_id = 1, version = 0;
    std::size_t snapshots_alive = 0, snapshot_bytes_alive = 0;
    bool stop = false, busy_insert = false, busy_extract = false;
    // extract sweeps on the device now: the insert worker pauses between its
    // phases while one runs (at most kYieldMs per pause), so an extract
    // shares the device only with the insert work already queued
    std::atomic<int> extracts_running{0};
    std::thread insert_worker, extract_worker;

    void run_inserts();
    void run_extracts();
    void insert_one(Request& r);
    void extract_batch(std::vector<std::shared_ptr<Request>>& batch, const Snapshot& snap);
    std::shared_ptr<Snapshot> make_snapshot(std::uint64_t ver);
};

Snapshot::~Snapshot() {
    coresg_hip_graph_destroy(graph);
    if (owner) {
        {
            std::lock_guard<std::mutex> lk(owner->mu);
            --owner->snapshots_alive;
            owner->snapshot_bytes_alive -= bytes;
        }
        owner->cv_mem.notify_all();
    }
}

// Moves the working graph's freshly laid out edge list and adjacency into a
// new extraction-only graph, with copies of the k-NN distances and counts.
std::shared_ptr<Snapshot> coresg_hip_session::make_snapshot(std::uint64_t ver) {
    Graph& w = graph->g;
    if (w.yield_hook) w.yield_hook();
    prepare_edges(w);
    std::unique_ptr<coresg_hip_graph> h(new coresg_hip_graph());
    Graph& g = h->g;
    g.device = w.device;
    g.n = w.n;
    g.dim = w.dim;
    g.stride = w.stride;
    g.k = w.k;
    create_stream(&g.stream);
    const std::size_t nk = std::size_t(w.n) * w.k;
    g.dist.reserve(nk);
    g.count.reserve(w.n);
    CSG_CHECK(hipMemcpyAsync(g.dist.p, w.dist.p, nk * 4, hipMemcpyDeviceToDevice, w.stream));
    CSG_CHECK(hipMemcpyAsync(g.count.p, w.count.p, std::size_t(w.n) * 4, hipMemcpyDeviceToDevice,
                             w.stream));
    CSG_CHECK(hipStreamSynchronize(w.stream));
    g.eu.adopt(w.eu);
    g.ev.adopt(w.ev);
    g.ew.adopt(w.ew);
    g.eoff.adopt(w.eoff);
    g.adj_off.adopt(w.adj_off);
    g.adj.adopt(w.adj);
    g.num_edges = w.num_edges;
    g.knn_edges = w.knn_edges;
    g.knn_set = true;
    g.edges_ready = true;
    w.edges_ready = false;  // its edge arrays now belong to the snapshot
    auto s = std::make_shared<Snapshot>();
    s->version = ver;
    s->n = g.n;
    s->bytes = coresg_hip_graph_device_bytes(h.get());
    s->graph = h.release();
    {
        std::lock_guard<std::mutex> lk(mu);
        s->owner = this;
        ++snapshots_alive;
        snapshot_bytes_alive += s->bytes;
    }
    return s;
}

void coresg_hip_session::insert_one(Request& r) {
    Graph& g = graph->g;
    if (hbm_budget) {
        // the next version, with the edge count per point of the current one
        const double epp = double(published ? published->graph->g.num_edges : 0) /
                           double(g.n ? g.n : 1);
        const std::size_t n1 = std::size_t(g.n) + r.count;
        const std::size_t need = version_bytes(n1, std::size_t(epp * double(n1)) + 1, g.stride,
                                               g.k, params.sample ? params.sample : 16u);
        std::unique_lock<std::mutex> lk(mu);
        // older snapshots that in-flight extracts still hold free up as those
        // finish; the current one stays until the new version replaces it
        auto others = [&] {
            return snapshot_bytes_alive - (published ? published->bytes : 0);
        };
        while (need + others() > hbm_budget && snapshots_alive > 1 && !stop)
            cv_mem.wait(lk);
        if (need + others() > hbm_budget)
            throw std::runtime_error("insert not admitted: the next version needs " +
                                     std::to_string(need) +
                                     " B of device memory (analytic model), above the "
                                     "session's HBM budget of " +
                                     std::to_string(hbm_budget) + " B");
    }
    insert_points(g, r.rows.data(), r.count, CORESG_HIP_HOST, 0, params, &r.insert_stats);
    std::vector<float>().swap(r.rows);
}

void coresg_hip_session::extract_batch(std::vector<std::shared_ptr<Request>>& batch,
                                       const Snapshot& snap) {
    Graph& g = snap.graph->g;
    const std::uint32_t n = g.n;
    std::vector<std::uint32_t> all;
    bool edges = false, labels = false;
    for (auto& r : batch) {
        all.insert(all.end(), r->mpts.begin(), r->mpts.end());
        edges = edges || r->want_edges;
        labels = labels || r->want_labels;
    }
    const std::size_t q = all.size(), estride = n ? n - 1 : 0;
    std::vector<coresg_hip_mst_info> info(q);
    std::vector<std::uint32_t> u, v, lab;
    std::vector<float> w;
    if (edges) {
        u.resize(q * estride);
        v.resize(q * estride);
        w.resize(q * estride);
    }
    if (labels) lab.resize(q * std::size_t(n));
    extract_msts(g, all.data(), static_cast<std::uint32_t>(q), info.data(),
                 edges ? u.data() : nullptr, edges ? v.data() : nullptr,
                 edges ? w.data() : nullptr, labels ? lab.data() : nullptr);
    std::size_t off = 0;
    for (auto& r : batch) {
        const std::size_t c = r->mpts.size();
        r->info.assign(info.begin() + off, info.begin() + off + c);
        if (r->want_edges) {
            r->u.assign(u.begin() + off * estride, u.begin() + (off + c) * estride);
            r->v.assign(v.begin() + off * estride, v.begin() + (off + c) * estride);
            r->w.assign(w.begin() + off * estride, w.begin() + (off + c) * estride);
        }
        if (r->want_labels)
            r->labels.assign(lab.begin() + off * n, lab.begin() + (off + c) * n);
        off += c;
    }
}

void coresg_hip_session::run_inserts() {
    (void)hipSetDevice(graph->g.device);
    for (;;) {
        std::shared_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_insert.wait(lk, [&] { return stop || !inserts.empty(); });
            if (inserts.empty()) return;  // stop requested and drained
            r = inserts.front();
            inserts.pop_front();
            busy_insert = true;
            r->state = CORESG_HIP_REQ_RUNNING;
            r->t_start = Clock::now();
            r->batch_requests = 1;
        }
        std::string err;
        std::shared_ptr<Snapshot> next, old;
        try {
            insert_one(*r);
            next = make_snapshot(version + 1);
        } catch (const std::exception& e) {
            err = e.what();
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            if (err.empty()) {
                ++version;
                old = std::move(published);
                published = next;
            }
            r->state = err.empty() ? CORESG_HIP_REQ_DONE : CORESG_HIP_REQ_FAILED;
            r->error = err;
            r->t_done = Clock::now();
            r->graph_n = graph->g.n;
            r->graph_version = version;
            busy_insert = false;
        }
        old.reset();   // freed here unless an extract still sweeps it
        next.reset();
        cv_done.notify_all();
    }
}

void coresg_hip_session::run_extracts() {
    (void)hipSetDevice(graph->g.device);
    for (;;) {
        std::vector<std::shared_ptr<Request>> batch;
        std::shared_ptr<Snapshot> snap;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_extract.wait(lk, [&] { return stop || !extracts.empty(); });
            if (extracts.empty()) return;
            std::size_t slots = 0;
            while (!extracts.empty() &&
                   (batch.empty() || slots + extracts.front()->mpts.size() <= kBatchSlots)) {
                slots += extracts.front()->mpts.size();
                batch.push_back(extracts.front());
                extracts.pop_front();
            }
            snap = published;
            busy_extract = true;
            const auto t = Clock::now();
            for (auto& r : batch) {
                r->state = CORESG_HIP_REQ_RUNNING;
                r->t_start = t;
                r->batch_requests = static_cast<std::uint32_t>(batch.size());
            }
        }
        std::string err;
        extracts_running.fetch_add(1);
        try {
            extract_batch(batch, *snap);
        } catch (const std::exception& e) {
            err = e.what();
        }
        extracts_running.fetch_sub(1);
        {
            std::lock_guard<std::mutex> lk(mu);
            const auto t = Clock::now();
            for (auto& r : batch) {
                r->state = err.empty() ? CORESG_HIP_REQ_DONE : CORESG_HIP_REQ_FAILED;
                r->error = err;
                r->t_done = t;
                r->graph_n = snap->n;
                r->graph_version = snap->version;
            }
            busy_extract = false;
        }
        snap.reset();  // may free a replaced version
        cv_done.notify_all();
    }
}

namespace {

coresg_hip_session* start_session(coresg_hip_graph* g, const coresg_hip_build_params* p,
                                  std::size_t budget) {
    std::unique_ptr<coresg_hip_session> s(new coresg_hip_session());
    s->graph = g;
    s->params = *p;
    s->hbm_budget = budget;
    s->published = s->make_snapshot(0);
    coresg_hip_session* self = s.get();
    g->g.yield_hook = [self] {
        const auto t0 = Clock::now();
        while (self->extracts_running.load() > 0 && ms_between(t0, Clock::now()) < kYieldMs)
            std::this_thread::sleep_for(std::chrono::microseconds(50));
    };
    coresg_hip_session* raw = s.release();
    raw->insert_worker = std::thread([raw] { raw->run_inserts(); });
    raw->extract_worker = std::thread([raw] { raw->run_extracts(); });
    return raw;
}

void fill_status(const Request& r, coresg_hip_request_status* st) {
    std::memset(st, 0, sizeof(*st));
    st->id = r.id;
    st->kind = r.kind;
    st->state = r.state;
    st->graph_n = r.graph_n;
    st->graph_version = r.graph_version;
    st->num_msts = static_cast<std::uint32_t>(r.mpts.size());
    st->batch_requests = r.batch_requests;
    if (r.state == CORESG_HIP_REQ_DONE || r.state == CORESG_HIP_REQ_FAILED) {
        st->queued_ms = ms_between(r.t_submit, r.t_start);
        st->run_ms = ms_between(r.t_start, r.t_done);
        st->latency_ms = ms_between(r.t_submit, r.t_done);
    }
    st->insert_stats = r.insert_stats;
    std::strncpy(st->error, r.error.c_str(), sizeof(st->error) - 1);
}

std::uint64_t enqueue(coresg_hip_session* s, std::shared_ptr<Request> r) {
    const bool ins = r->kind == CORESG_HIP_REQ_INSERT;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) throw std::runtime_error("session is closing");
        r->id = s->next_id++;
        r->t_submit = Clock::now();
        s->requests[r->id] = r;
        (ins ? s->inserts : s->extracts).push_back(r);
    }
    (ins ? s->cv_insert : s->cv_extract).notify_one();
    return r->id;
}

std::shared_ptr<Request> find(coresg_hip_session* s, std::uint64_t id) {
    auto it = s->requests.find(id);
    if (it == s->requests.end()) throw std::runtime_error("unknown request id");
    return it->second;
}

}  // namespace

extern "C" {

coresg_hip_session* coresg_hip_session_create(int device, const float* rows, uint32_t n,
                                              uint32_t dim, uint32_t k_max,
                                              const coresg_hip_build_params* p,
                                              size_t hbm_budget_bytes,
                                              coresg_hip_build_stats* st) {
    coresg_hip_graph* g = coresg_hip_graph_create(device, n, dim, k_max);
    if (!g) return nullptr;
    if (!coresg_hip_graph_set_points(g, rows, CORESG_HIP_HOST, 0) || !coresg_hip_build(g, p, st)) {
        const std::string msg = coresg_hip_last_error();
        coresg_hip_graph_destroy(g);
        set_last_error(msg);
        return nullptr;
    }
    coresg_hip_session* s = nullptr;
    if (!guarded([&] { s = start_session(g, p, hbm_budget_bytes); })) {
        const std::string msg = coresg_hip_last_error();
        coresg_hip_graph_destroy(g);
        set_last_error(msg);
        return nullptr;
    }
    return s;
}

coresg_hip_session* coresg_hip_session_adopt(coresg_hip_graph* g,
                                             const coresg_hip_build_params* p,
                                             size_t hbm_budget_bytes) {
    if (!g) return nullptr;
    coresg_hip_session* s = nullptr;
    if (!guarded([&] {
            CSG_CHECK(hipSetDevice(g->g.device));
            s = start_session(g, p, hbm_budget_bytes);
        }))
        return nullptr;
    return s;
}

uint64_t coresg_hip_session_submit_insert(coresg_hip_session* s, const float* rows,
                                          uint32_t count) {
    std::uint64_t id = 0;
    guarded([&] {
        auto r = std::make_shared<Request>();
        r->kind = CORESG_HIP_REQ_INSERT;
        r->count = count;
        const std::size_t dim = s->graph->g.dim;
        r->rows.assign(rows, rows + std::size_t(count) * dim);
        id = enqueue(s, r);
    });
    return id;
}

uint64_t coresg_hip_session_submit_extract(coresg_hip_session* s, const uint32_t* mpts,
                                           uint32_t count, int want_edges, int want_labels) {
    std::uint64_t id = 0;
    guarded([&] {
        if (count == 0) throw std::runtime_error("extract request without mpts values");
        auto r = std::make_shared<Request>();
        r->kind = CORESG_HIP_REQ_EXTRACT;
        r->mpts.assign(mpts, mpts + count);
        for (std::uint32_t m : r->mpts)
            if (m > s->graph->g.k) throw std::runtime_error("CoreSG::edges: mpts exceeds k_max");
        r->want_edges = want_edges != 0;
        r->want_labels = want_labels != 0;
        id = enqueue(s, r);
    });
    return id;
}

int coresg_hip_session_wait(coresg_hip_session* s, uint64_t id, double timeout_s,
                            coresg_hip_request_status* st) {
    int done = 0;
    const int ok = guarded([&] {
        std::unique_lock<std::mutex> lk(s->mu);
        auto r = find(s, id);
        auto finished = [&] {
            return r->state == CORESG_HIP_REQ_DONE || r->state == CORESG_HIP_REQ_FAILED;
        };
        if (timeout_s < 0)
            s->cv_done.wait(lk, finished);
        else
            s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished);
        done = finished() ? 1 : 0;
        if (st) fill_status(*r, st);
    });
    return ok ? done : -1;
}

int coresg_hip_session_extract_result(coresg_hip_session* s, uint64_t id,
                                      coresg_hip_mst_info* info, uint32_t* out_u,
                                      uint32_t* out_v, float* out_w, uint32_t* out_labels) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(s->mu);
        auto r = find(s, id);
        if (r->kind != CORESG_HIP_REQ_EXTRACT) throw std::runtime_error("not an extract request");
        if (r->state != CORESG_HIP_REQ_DONE) throw std::runtime_error("request not completed");
        if (info) std::copy(r->info.begin(), r->info.end(), info);
        const std::size_t n = r->graph_n, estride = n ? n - 1 : 0;
        for (std::size_t i = 0; i < r->info.size(); ++i) {
            const std::size_t t = r->info[i].num_edges;
            if ((out_u || out_v || out_w) && !r->want_edges)
                throw std::runtime_error("edges were not requested");
            if (out_u) std::copy_n(r->u.begin() + i * estride, t, out_u + i * estride);
            if (out_v) std::copy_n(r->v.begin() + i * estride, t, out_v + i * estride);
            if (out_w) std::copy_n(r->w.begin() + i * estride, t, out_w + i * estride);
        }
        if (out_labels) {
            if (!r->want_labels) throw std::runtime_error("labels were not requested");
            std::copy(r->labels.begin(), r->labels.end(), out_labels);
        }
    });
}

int coresg_hip_session_release(coresg_hip_session* s, uint64_t id) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(s->mu);
        auto r = find(s, id);
        if (r->state != CORESG_HIP_REQ_DONE && r->state != CORESG_HIP_REQ_FAILED)
            throw std::runtime_error("request still in flight");
        s->requests.erase(id);
    });
}

coresg_hip_graph* coresg_hip_session_graph(coresg_hip_session* s) {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] {
        return s->inserts.empty() && s->extracts.empty() && !s->busy_insert && !s->busy_extract;
    });
    return s->graph;
}

size_t coresg_hip_session_device_bytes(coresg_hip_session* s) {
    std::lock_guard<std::mutex> lk(s->mu);
    return coresg_hip_graph_device_bytes(s->graph) + s->snapshot_bytes_alive;
}

void coresg_hip_session_destroy(coresg_hip_session* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_insert.notify_all();
    s->cv_extract.notify_all();
    s->cv_mem.notify_all();
    if (s->insert_worker.joinable()) s->insert_worker.join();
    if (s->extract_worker.joinable()) s->extract_worker.join();
    s->published.reset();
    s->graph->g.yield_hook = nullptr;
    coresg_hip_graph_destroy(s->graph);
    delete s;
}

}  // extern "C"
