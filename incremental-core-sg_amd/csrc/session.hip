// Serving session: one device-resident CORE-SG, a request queue and a worker
// thread that owns the device.
//
// Replaces the reference's one-shot `coresg_cli update` / `extract`
// round trips (coresg_cli.cpp:507-570), which reload the CORE-SG from disk for
// every command, with a long-lived graph in HBM:
//   * insert requests run CoreSG::insert (coresg.cpp:325-389) on the resident
//     graph, admitted only while the graph's projected device footprint stays
//     within the session's HBM budget;
//   * extract requests run CoreSG::mst for their mpts list (coresg.cpp:408-417).
// Requests take effect in submission order: an extract sees exactly the
// inserts submitted before it. Consecutive extract requests that are waiting
// when the worker becomes free are admitted together into one batched sweep
// (up to kBatchSlots mpts values), so requests that arrive while others run
// share the next pass over the graph; each slot of a sweep is computed
// independently, so a result is the same alone and in a mix.

#include <algorithm>
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

}  // namespace

struct coresg_hip_session {
    coresg_hip_graph* graph = nullptr;  // owned
    coresg_hip_build_params params{};
    std::size_t hbm_budget = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<std::uint64_t, std::shared_ptr<Request>> requests;
    std::uint64_t next_id = 1, version = 0;
    bool stop = false, busy = false;
    std::thread worker;

    void run();
    void run_insert(Request& r);
    void run_extracts(std::vector<std::shared_ptr<Request>>& batch);
};

void coresg_hip_session::run_insert(Request& r) {
    Graph& g = graph->g;
    if (hbm_budget) {
        // the footprint grows with n: project it linearly from the current one
        const double now = double(coresg_hip_graph_device_bytes(graph));
        const double projected = now * double(g.n + r.count) / double(g.n ? g.n : 1);
        if (projected > double(hbm_budget))
            throw std::runtime_error("insert not admitted: projected device footprint " +
                                     std::to_string(std::size_t(projected)) +
                                     " B exceeds the session's HBM budget of " +
                                     std::to_string(hbm_budget) + " B");
    }
    insert_points(g, r.rows.data(), r.count, CORESG_HIP_HOST, 0, params, &r.insert_stats);
    prepare_edges(g);
    std::vector<float>().swap(r.rows);
}

void coresg_hip_session::run_extracts(std::vector<std::shared_ptr<Request>>& batch) {
    Graph& g = graph->g;
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

void coresg_hip_session::run() {
    (void)hipSetDevice(graph->g.device);
    for (;;) {
        std::vector<std::shared_ptr<Request>> batch;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !queue.empty(); });
            if (queue.empty()) return;  // stop requested and drained
            if (queue.front()->kind == CORESG_HIP_REQ_INSERT) {
                batch.push_back(queue.front());
                queue.pop_front();
            } else {
                std::size_t slots = 0;
                while (!queue.empty() && queue.front()->kind == CORESG_HIP_REQ_EXTRACT &&
                       (batch.empty() || slots + queue.front()->mpts.size() <= kBatchSlots)) {
                    slots += queue.front()->mpts.size();
                    batch.push_back(queue.front());
                    queue.pop_front();
                }
            }
            busy = true;
            const auto t = Clock::now();
            for (auto& r : batch) {
                r->state = CORESG_HIP_REQ_RUNNING;
                r->t_start = t;
                r->batch_requests = static_cast<std::uint32_t>(batch.size());
            }
        }
        std::string err;
        try {
            if (batch.front()->kind == CORESG_HIP_REQ_INSERT)
                run_insert(*batch.front());
            else
                run_extracts(batch);
        } catch (const std::exception& e) {
            err = e.what();
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            const auto t = Clock::now();
            if (err.empty() && batch.front()->kind == CORESG_HIP_REQ_INSERT) ++version;
            for (auto& r : batch) {
                r->state = err.empty() ? CORESG_HIP_REQ_DONE : CORESG_HIP_REQ_FAILED;
                r->error = err;
                r->t_done = t;
                r->graph_n = graph->g.n;
                r->graph_version = version;
            }
            busy = false;
        }
        cv_done.notify_all();
    }
}

namespace {

coresg_hip_session* start_session(coresg_hip_graph* g, const coresg_hip_build_params* p,
                                  std::size_t budget) {
    auto* s = new coresg_hip_session();
    s->graph = g;
    s->params = *p;
    s->hbm_budget = budget;
    s->worker = std::thread([s] { s->run(); });
    return s;
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
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) throw std::runtime_error("session is closing");
        r->id = s->next_id++;
        r->t_submit = Clock::now();
        s->requests[r->id] = r;
        s->queue.push_back(r);
    }
    s->cv_work.notify_one();
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
    if (!coresg_hip_graph_set_points(g, rows, CORESG_HIP_HOST, 0) || !coresg_hip_build(g, p, st) ||
        !coresg_hip_prepare_edges(g, nullptr, nullptr)) {
        const std::string msg = coresg_hip_last_error();
        coresg_hip_graph_destroy(g);
        set_last_error(msg);
        return nullptr;
    }
    return start_session(g, p, hbm_budget_bytes);
}

coresg_hip_session* coresg_hip_session_adopt(coresg_hip_graph* g,
                                             const coresg_hip_build_params* p,
                                             size_t hbm_budget_bytes) {
    if (!g || !coresg_hip_prepare_edges(g, nullptr, nullptr)) return nullptr;
    return start_session(g, p, hbm_budget_bytes);
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
    s->cv_done.wait(lk, [&] { return s->queue.empty() && !s->busy; });
    return s->graph;
}

void coresg_hip_session_destroy(coresg_hip_session* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    coresg_hip_graph_destroy(s->graph);
    delete s;
}

}  // extern "C"
