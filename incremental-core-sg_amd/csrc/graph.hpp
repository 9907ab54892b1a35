// Device-resident CORE-SG state shared by the build and extraction paths.
#pragma once

#include <algorithm>
#include <initializer_list>
#include <map>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace coresg_hip {

// One symmetric-adjacency entry: the other endpoint (bit 31 when seen from the
// edge's second endpoint), the edge's rank in the sorted list, its base
// weight. Written and read as one 12-byte access.
struct AdjEntry {
    std::uint32_t nbr, rank;
    float w;
};

// Device allocations, frees and stream creation/destruction on one thread and
// a stream capture on another are kept apart (a serving session builds the
// next graph version while the previous one is swept, and the sweep captures
// its Boruvka rounds): the runtime may synchronise the device inside a free.
inline std::mutex& device_api_mutex() {
    static std::mutex mu;
    return mu;
}

// A device-memory budget for the allocations one entry point makes on the
// calling thread (coresg_hip_set_device_budget): every DBuf allocation is
// checked against it before hipMalloc, and exceeding it fails the way a
// device out-of-memory does.
struct AllocBudget {
    std::size_t limit = 0;  // bytes; 0 = none
    std::size_t used = 0;
};
inline thread_local AllocBudget* t_alloc_budget = nullptr;

inline void budget_charge(std::size_t bytes) {
    AllocBudget* b = t_alloc_budget;
    if (!b || !b->limit) return;
    if (b->used + bytes > b->limit)
        throw HipError("hipMalloc(" + std::to_string(bytes) +
                           " B) -> hipErrorOutOfMemory: the device budget of " +
                           std::to_string(b->limit) + " B is exhausted (" +
                           std::to_string(b->used) + " B in use)",
                       hipErrorOutOfMemory);
    b->used += bytes;
}

// Owning device allocation (hipMalloc / hipFree) with its element count.
template <typename T>
struct DBuf {
    T* p = nullptr;
    std::size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) {
            std::lock_guard<std::mutex> lk(device_api_mutex());
            (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
    }
    // Ensure capacity for `count` elements (contents are not preserved). A
    // buffer that has to grow takes a quarter more than asked, so buffers
    // sized by a growing point count (a stream of inserts) are reallocated
    // once per quarter of growth instead of on every insert.
    void reserve(std::size_t count) {
        if (count <= n && p) return;
        const std::size_t want = p ? std::max(count, n + n / 4) : count;
        release();
        const std::size_t bytes = (want ? want : 1) * sizeof(T);
        budget_charge(bytes);
        std::lock_guard<std::mutex> lk(device_api_mutex());
        CSG_CHECK(hipMalloc(reinterpret_cast<void**>(&p), bytes));
        n = want;
    }
    // Take over another buffer's allocation (this one's is freed).
    void adopt(DBuf& o) {
        release();
        p = o.p;
        n = o.n;
        o.p = nullptr;
        o.n = 0;
    }
    // Grow to at least `count` elements keeping the first `keep` (copied on
    // `s`); capacity grows by half again at least, so a stream of inserts
    // reallocates O(log n) times.
    void grow(std::size_t count, std::size_t keep, hipStream_t s) {
        if (count <= n && p) return;
        const std::size_t cap = std::max(count, n + n / 2);
        T* q = nullptr;
        budget_charge((cap ? cap : 1) * sizeof(T));
        {
            std::lock_guard<std::mutex> lk(device_api_mutex());
            CSG_CHECK(hipMalloc(reinterpret_cast<void**>(&q), (cap ? cap : 1) * sizeof(T)));
        }
        if (p && keep)
            CSG_CHECK(hipMemcpyAsync(q, p, keep * sizeof(T), hipMemcpyDeviceToDevice, s));
        CSG_CHECK(hipStreamSynchronize(s));
        release();
        p = q;
        n = cap;
    }
    std::size_t bytes() const { return n * sizeof(T); }
};

// Pinned host allocation (hipHostMalloc / hipHostFree): the target of
// asynchronous device-to-host copies.
template <typename T>
struct HostBuf {
    T* p = nullptr;
    std::size_t n = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() { release(); }
    void release() {
        if (p) {
            std::lock_guard<std::mutex> lk(device_api_mutex());
            (void)hipHostFree(p);
        }
        p = nullptr;
        n = 0;
    }
    // Ensure capacity for `count` elements (contents are not preserved).
    void reserve(std::size_t count) {
        if (count <= n && p) return;
        release();
        std::lock_guard<std::mutex> lk(device_api_mutex());
        CSG_CHECK(hipHostMalloc(reinterpret_cast<void**>(&p), (count ? count : 1) * sizeof(T),
                                hipHostMallocDefault));
        n = count;
    }
};

// A non-blocking stream destroyed with its owner (also on an exception path).
inline void create_stream(hipStream_t* s) {
    std::lock_guard<std::mutex> lk(device_api_mutex());
    CSG_CHECK(hipStreamCreateWithFlags(s, hipStreamNonBlocking));
}
inline void destroy_stream(hipStream_t s) {
    if (!s) return;
    (void)hipStreamSynchronize(s);
    std::lock_guard<std::mutex> lk(device_api_mutex());
    (void)hipStreamDestroy(s);
}

struct OwnedStream {
    hipStream_t s = nullptr;
    OwnedStream() { create_stream(&s); }
    OwnedStream(const OwnedStream&) = delete;
    OwnedStream& operator=(const OwnedStream&) = delete;
    ~OwnedStream() { destroy_stream(s); }
};

// Stream capture bracket: ends the capture and discards its graph when left by
// an exception, so the stream never stays in capture mode after a failure.
struct CaptureGuard {
    hipStream_t s;
    bool open = false;
    std::unique_lock<std::mutex> api;  // no allocation or free elsewhere meanwhile
    explicit CaptureGuard(hipStream_t stream) : s(stream), api(device_api_mutex()) {
        CSG_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        open = true;
    }
    CaptureGuard(const CaptureGuard&) = delete;
    CaptureGuard& operator=(const CaptureGuard&) = delete;
    // Ends the capture; the caller owns the returned graph.
    hipGraph_t end() {
        hipGraph_t graph = nullptr;
        open = false;
        CSG_CHECK(hipStreamEndCapture(s, &graph));
        return graph;
    }
    ~CaptureGuard() {
        if (!open) return;
        hipGraph_t graph = nullptr;
        if (hipStreamEndCapture(s, &graph) == hipSuccess && graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
    }
};

// Two reusable events (no timing).
struct EventPair {
    hipEvent_t e[2] = {nullptr, nullptr};
    EventPair() = default;
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
    ~EventPair() {
        for (hipEvent_t& x : e)
            if (x) (void)hipEventDestroy(x);
    }
    void create() {
        for (hipEvent_t& x : e)
            if (!x) CSG_CHECK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
    }
};

// One NN-Descent round's counts as copied back to the host.
struct RoundReadback {
    unsigned long long upd;
    unsigned long long tally[3];       // join work (profiling)
    unsigned long long gram_tally[3];  // Gram-filter work: pairs, exact, tiles (profiling)
    unsigned int gram_audit[2];        // Gram-filter audit maxima (profiling)
};

// NN-Descent round scratch: sampled lists, bounded reverse lists, gathered
// candidates and the per-node proposal buffers (kept across builds).
struct DescentBuffers {
    DBuf<std::uint32_t> new_list, new_cnt, old_list, old_cnt;
    DBuf<std::uint32_t> rev_new, rev_new_cnt, rev_old, rev_old_cnt;
    DBuf<unsigned long long> cand, updates, partial, tally;
    DBuf<std::uint32_t> fill, gids, gcnt;
    DBuf<float> thr_node;  // each node's join threshold this round
    // reverse lists by target buckets: sampled (target, source) pairs, the
    // per-block bucket histograms (scanned in place), scan scratch, total
    DBuf<unsigned long long> rev_pairs;
    DBuf<std::uint32_t> rev_hist, rev_tmp, rev_total;
    DBuf<float> gthr;
    DBuf<unsigned long long> gram_tally;
    DBuf<unsigned int> gram_audit;
    // device-side convergence flag and the per-round counts read back
    DBuf<std::uint32_t> halt;
    // filter join: nodes set aside for its second pass, and their count
    DBuf<std::uint32_t> defer, defer_count;
    HostBuf<RoundReadback> readback;
    EventPair round_done;
};

// Launch shape of the sweep's Boruvka round kernel: lanes per vertex group
// (at least the slot count rounded up to a power of two) and the grid cap.
struct SweepShape {
    std::uint32_t lanes = 0, max_blocks = 0;
};

// Accumulated device time of one kernel (HIP events on the graph's stream).
struct KernelTimer {
    std::vector<hipEvent_t> pending;  // start/stop pairs not yet resolved
    double total_ms = 0.0;
    std::uint64_t launches = 0;
    std::uint64_t work[4] = {0, 0, 0, 0};  // kernel-specific work tallies
};

struct Graph {
    int device = 0;
    hipStream_t stream = nullptr;
    std::uint32_t n = 0, dim = 0, stride = 0, k = 0;
    // Called by an insert between its phases (search, scan, merge, each
    // refinement round, hubness, bridges); a serving session sets it to let
    // an extract sweep in flight run first. Empty: no pauses.
    std::function<void()> yield_hook;

    // Points and k-NN lists (reference SoA layout, metric units).
    DBuf<float> pts;
    DBuf<std::uint32_t> ids;
    DBuf<float> dist;
    DBuf<std::uint8_t> flags;
    DBuf<std::uint32_t> count;
    bool points_set = false, knn_set = false;

    // Bridges (host copy is authoritative; device copy for the edge build).
    std::vector<std::uint32_t> bu, bv;
    std::vector<float> bw;
    std::size_t repair_count = 0;
    std::uint32_t missing_mode = 0;  // MissingEdgeMode of the bridges (coresg.hpp:54-58)

    // Ownership of every slot (knn_graph.cpp:143-166), as computed for the
    // build's edge count; valid while the k-NN lists are unchanged, and reused
    // by prepare_edges.
    bool emit_ready = false;
    DBuf<std::uint8_t> emit_flags;
    // Deduplicated, (u,v)-sorted CORE-SG edge list.
    bool edges_ready = false;
    std::uint64_t num_edges = 0, knn_edges = 0;
    DBuf<std::uint32_t> eu, ev, eoff;
    DBuf<float> ew;
    // Symmetric CSR adjacency of the same edges: for vertex x, entries
    // adj_off[x] .. adj_off[x+1]-1 hold the other endpoint (bit 31 set when x
    // is the edge's second endpoint), the edge's rank in the sorted list and
    // its base weight, one 12-byte record each.
    DBuf<std::uint32_t> adj_off;
    DBuf<AdjEntry> adj;

    // Extraction scratch.
    DBuf<float> core;
    DBuf<std::uint32_t> comp, parent, newroot, minv, labels_tmp;
    DBuf<std::uint64_t> best, mst_keys;
    DBuf<std::uint8_t> active;  // Boruvka: vertex may still have an outgoing edge
    // Boruvka, sweeps of few slots: each vertex's best outgoing edge at its
    // last scan (key, far endpoint)
    DBuf<std::uint64_t> sweep_vkey;
    DBuf<std::uint32_t> sweep_vfar;
    DBuf<std::uint8_t> mst_flag;
    DBuf<std::uint32_t> counters;     // small device counters
    DBuf<std::uint64_t> counters64;
    std::vector<std::uint32_t> host_counters;
    DBuf<std::uint32_t> scan_tmp;
    DBuf<std::uint32_t> out_pos;
    // mpts sweep: per-block per-slot partials, per-slot results, slot mpts,
    // and one slot's outputs
    DBuf<std::uint64_t> slot_part, slot_out;
    DBuf<std::uint32_t> slot_mpts, label_buf, out_u, out_v, slot_rounds;
    // the captured Boruvka rounds (replayed while slots are live) and the
    // launch parameters it was captured with
    hipGraphExec_t sweep_exec = nullptr;
    std::vector<unsigned char> sweep_key;
    // the descent's rounds as one captured graph, and what it was made for
    hipGraphExec_t descent_exec = nullptr;
    std::vector<std::uint64_t> descent_key, descent_key_seen;
    DBuf<float> out_w;
    // autotuned sweep launch shapes of this graph, by slots (power of two)
    std::map<std::uint32_t, SweepShape> sweep_tune;

    DescentBuffers nnd;
    DBuf<unsigned long long> ins_partial;  // insert: per-chunk best lists
    DBuf<std::uint32_t> ins_hints;         // insert search: entry hints

    // Hubness / selection state (filled by the build).
    DBuf<std::int32_t> rdeg;
    DBuf<double> score;
    DBuf<std::uint32_t> conn;
    std::uint32_t num_components = 0;
    std::uint32_t score_k = 0;  // k of the score's degenerate case (0 = k)
    std::vector<std::uint32_t> selected;
    bool hub_ready = false;

    // Independent instances sharing the graph (batched uneven point sets):
    // instance i holds points [inst_off[i], inst_off[i+1]); empty = one
    // instance of all points. inst_of / inst_off_d are the device copies.
    std::vector<std::uint32_t> inst_off;
    DBuf<std::uint32_t> inst_of, inst_off_d;
    std::uint32_t num_instances() const {
        return inst_off.empty() ? 1u : static_cast<std::uint32_t>(inst_off.size() - 1);
    }

    // The plain form: every launch followed by a host synchronisation, the
    // sweep one mpts value per pass with every Boruvka round launched and
    // checked from the host, the descent's convergence read after each round.
    bool plain = false;
    // exact_mst: one launch per Prim step instead of one cooperative launch
    bool exact_stepped = false;
    // Profiling (audit: FP64 check of the Gram-filter join's distances).
    bool profiling = false;
    bool audit = false;
    std::map<std::string, KernelTimer> timers;

    ~Graph();
    void ensure_scan_tmp(std::size_t n);
    void timer_start(const char* name, hipEvent_t* start);
    void timer_stop(const char* name, hipEvent_t start);
    void resolve_timers();
    void add_work(const char* name, std::initializer_list<std::uint64_t> w);
    void max_work(const char* name, int index, std::uint64_t value);
};

// RAII bracket for a profiled launch.
struct ScopedKernelTimer {
    Graph& g;
    const char* name;
    hipEvent_t start = nullptr;
    ScopedKernelTimer(Graph& graph, const char* nm) : g(graph), name(nm) {
        if (g.profiling) g.timer_start(name, &start);
    }
    ~ScopedKernelTimer() {
        if (g.profiling && start) g.timer_stop(name, start);
    }
};

// extraction (extract.hip)
void prepare_edges(Graph& g);
void check_mpts(const Graph& g, std::uint32_t* mpts);
void extract_msts(Graph& g, const std::uint32_t* mpts, std::uint32_t count, void* info,
                  std::uint32_t* out_u, std::uint32_t* out_v, float* out_w,
                  std::uint32_t* out_labels);
void reweighted_edges(Graph& g, std::uint32_t mpts, std::uint32_t* u, std::uint32_t* v,
                      float* w);
// The sweep's round-kernel launch shape for Bp slots on g: tuned on g, else
// tuned on a graph of the same size bucket (log2 n, k), else the default.
SweepShape sweep_shape(const Graph& g, std::uint32_t Bp);
// Times the candidate launch shapes on g with the first min(count, 32) of
// `mpts` and keeps the fastest for g and its size bucket; `res` is a
// coresg_hip_tune_result.
void autotune_sweep(Graph& g, const std::uint32_t* mpts, std::uint32_t count, void* res);
std::string tune_table_text();
void tune_table_load(const char* text);
// Per-instance MST summaries of a batch of mpts values (info[i * count + q]).
void extract_instances(Graph& g, const std::uint32_t* mpts, std::uint32_t count, void* info);
// CoreSG::exact_mst over the resident points (exact.hip); info is a
// coresg_hip_mst_info.
void exact_mst(Graph& g, std::uint32_t mpts, void* info, std::uint32_t* out_u,
               std::uint32_t* out_v, float* out_w, std::uint32_t* out_labels);
// MissingEdgeMode::ExactMst bridges: the exact MST at k_max with base
// distances (exact.hip).
void exact_mst_bridges(Graph& g, std::vector<std::uint32_t>* bu, std::vector<std::uint32_t>* bv,
                       std::vector<float>* bw);
// Undirected k-NN edge count of every instance.
std::vector<std::uint64_t> knn_edges_per_instance(Graph& g);

}  // namespace coresg_hip
