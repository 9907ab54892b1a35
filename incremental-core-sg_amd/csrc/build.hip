// Build orchestration after the k-NN half: hubness, anti-hub selection, bridge
// MST and connectivity repair (CoreSG::build, src/coresg.cpp:278-305, with
// recompute_missing_edges :143-169 and repair_connectivity :171-258).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "build.hpp"
#include "hubness.hpp"

namespace coresg_hip {
namespace {

using Clock = std::chrono::steady_clock;
double since(Clock::time_point t) {
    return std::chrono::duration<double>(Clock::now() - t).count();
}

// src/hubness.cpp:16-40 (antihub_budget).
std::size_t antihub_budget(const coresg_hip_build_params& p, std::uint32_t dim,
                           std::uint64_t num_edges, std::uint32_t n) {
    if (n == 0) return 0;
    const double logs = (num_edges > 1) ? std::log(static_cast<double>(num_edges)) : 0.0;
    double raw = p.antihub_multiplier * static_cast<double>(dim) * logs;
    if (!(raw > 0.0)) raw = 0.0;
    std::size_t count = static_cast<std::size_t>(std::llround(raw));
    if (p.max_pair_ops > 0.0 && dim > 0) {
        const double m_max = std::sqrt(2.0 * p.max_pair_ops / static_cast<double>(dim));
        if (m_max >= 1.0 && static_cast<double>(count) > m_max)
            count = static_cast<std::size_t>(m_max);
    }
    if (p.max_selected && count > p.max_selected) count = p.max_selected;
    if (count < p.min_selected) count = p.min_selected;
    if (count > n) count = n;
    return count;
}

// Union-find by rank with path halving (include/coresg/mst.hpp:30-75); labels
// dense by first appearance (src/mst.cpp:48-59).
struct HostUF {
    std::vector<std::uint32_t> parent;
    std::vector<std::uint8_t> rank;
    std::uint32_t components;
    explicit HostUF(std::uint32_t n) : parent(n), rank(n, 0), components(n) {
        for (std::uint32_t i = 0; i < n; ++i) parent[i] = i;
    }
    std::uint32_t find(std::uint32_t x) {
        while (parent[x] != x) {
            parent[x] = parent[parent[x]];
            x = parent[x];
        }
        return x;
    }
    bool unite(std::uint32_t a, std::uint32_t b) {
        a = find(a);
        b = find(b);
        if (a == b) return false;
        if (rank[a] < rank[b]) std::swap(a, b);
        parent[b] = a;
        if (rank[a] == rank[b]) ++rank[a];
        --components;
        return true;
    }
    std::vector<std::uint32_t> labels() {
        const std::uint32_t n = static_cast<std::uint32_t>(parent.size());
        std::vector<std::uint32_t> out(n), remap(n, kNone);
        std::uint32_t next = 0;
        for (std::uint32_t i = 0; i < n; ++i) {
            const std::uint32_t r = find(i);
            if (remap[r] == kNone) remap[r] = next++;
            out[i] = remap[r];
        }
        return out;
    }
};

// Repair of the nodes [lo, hi) (one instance): reps of each of its components
// after the bridges, an MST over the reps, the edges that join two
// components. Returns the number of added edges; *after receives the range's
// component count after the repair.
std::size_t repair_range(Graph& g, const coresg_hip_build_params& p,
                         const std::vector<std::uint32_t>& comp, std::uint32_t ncomp_all,
                         const std::vector<double>& score, std::uint32_t lo, std::uint32_t hi,
                         std::uint32_t* before, std::uint32_t* after) {
    // the range's components, densely renumbered in first-appearance order
    std::vector<std::uint32_t> local(ncomp_all, kNone);
    std::uint32_t nloc = 0;
    for (std::uint32_t v = lo; v < hi; ++v)
        if (local[comp[v]] == kNone) local[comp[v]] = nloc++;
    *before = *after = nloc;
    if (nloc <= 1 || !p.ensure_connected) return 0;
    std::size_t per = p.reps_per_component ? p.reps_per_component : 1;
    if (p.max_reps && std::size_t(nloc) * per > p.max_reps)
        per = std::max<std::size_t>(1, p.max_reps / std::size_t(nloc));
    std::vector<std::vector<std::uint32_t>> buckets(nloc);
    for (std::uint32_t v = lo; v < hi; ++v) {
        std::vector<std::uint32_t>& b = buckets[local[comp[v]]];
        if (b.size() < per) {
            b.push_back(v);
        } else {
            std::size_t worst = 0;
            for (std::size_t t = 1; t < b.size(); ++t)
                if (score[b[t]] > score[b[worst]]) worst = t;
            if (score[v] < score[b[worst]]) b[worst] = v;
        }
    }
    std::vector<std::uint32_t> reps;
    for (const auto& b : buckets) reps.insert(reps.end(), b.begin(), b.end());
    if (p.max_reps && reps.size() > p.max_reps) {
        // The reference's cut (coresg.cpp:221-227): std::partial_sort by score
        // alone. Among equal scores its heap procedure decides which
        // representatives survive and in what order (the order feeds Prim's
        // first-index tie rule), so the same call on the same sequence is made
        // here rather than any sort.
        std::partial_sort(reps.begin(), reps.begin() + p.max_reps, reps.end(),
                          [&](std::uint32_t a, std::uint32_t b) { return score[a] < score[b]; });
        reps.resize(p.max_reps);
    }
    if (reps.size() < 2) return 0;
    std::vector<std::uint32_t> ru, rv;
    std::vector<float> rw;
    mst_over_subset(g, reps, &ru, &rv, &rw);
    HostUF uf(nloc);
    std::size_t added = 0;
    for (std::size_t i = 0; i < ru.size(); ++i) {
        if (uf.unite(local[comp[ru[i]]], local[comp[rv[i]]])) {
            g.bu.push_back(ru[i]);
            g.bv.push_back(rv[i]);
            g.bw.push_back(rw[i]);
            ++added;
        }
    }
    *after = uf.components;
    return added;
}

void repair_connectivity(Graph& g, const coresg_hip_build_params& p, coresg_hip_build_stats* st) {
    const std::uint32_t n = g.n, ncomp0 = g.num_components;
    const std::uint32_t ninst = g.num_instances();
    if (st) st->components_before_repair = st->components_after_repair = ncomp0;
    if (ncomp0 <= ninst || n == 0) return;  // every instance is one component

    std::vector<std::uint32_t> conn(n);
    CSG_CHECK(hipMemcpyAsync(conn.data(), g.conn.p, std::size_t(n) * 4, hipMemcpyDeviceToHost,
                             g.stream));
    CSG_CHECK(hipStreamSynchronize(g.stream));
    // labels_with(missing) (connectivity.cpp:352-374)
    HostUF cuf(ncomp0);
    for (std::size_t i = 0; i < g.bu.size(); ++i) {
        if (g.bu[i] >= n || g.bv[i] >= n) continue;
        cuf.unite(conn[g.bu[i]], conn[g.bv[i]]);
    }
    const std::vector<std::uint32_t> merged = cuf.labels();
    const std::uint32_t ncomp = cuf.components;
    std::vector<std::uint32_t> comp(n);
    for (std::uint32_t v = 0; v < n; ++v) comp[v] = merged[conn[v]];
    if (st) st->components_before_repair = st->components_after_repair = ncomp;
    if (ncomp <= ninst || !p.ensure_connected) return;

    std::vector<double> score(n);
    CSG_CHECK(hipMemcpyAsync(score.data(), g.score.p, std::size_t(n) * 8,
                             hipMemcpyDeviceToHost, g.stream));
    CSG_CHECK(hipStreamSynchronize(g.stream));
    std::size_t added = 0;
    std::uint32_t before_all = 0, after_all = 0;
    for (std::uint32_t i = 0; i < ninst; ++i) {
        const std::uint32_t lo = g.inst_off.empty() ? 0 : g.inst_off[i];
        const std::uint32_t hi = g.inst_off.empty() ? n : g.inst_off[i + 1];
        std::uint32_t before = 0, after = 0;
        added += repair_range(g, p, comp, ncomp, score, lo, hi, &before, &after);
        before_all += before;
        after_all += after;
    }
    g.repair_count = added;
    if (st) {
        st->repair_edges = added;
        st->components_before_repair = before_all;
        st->components_after_repair = after_all;
    }
}

}  // namespace

std::uint64_t count_knn_edges(Graph& g);  // extract.hip

void derive_hubness(Graph& g, const coresg_hip_build_params& p) {
    if (!g.knn_set) throw std::runtime_error("derive_hubness: k-NN lists must be set");
    g.score_k = p.score_k;
    std::uint32_t tagged = 0;
    hubness_compute(g, &tagged);
    if (g.inst_off.empty()) {
        const std::uint64_t knn_edges = count_knn_edges(g);
        g.selected = select_antihubs(g, antihub_budget(p, g.dim, knn_edges, g.n),
                                     p.min_per_component);
    }
}

void build_bridges(Graph& g, const coresg_hip_build_params& p, coresg_hip_build_stats* st) {
    if (!g.knn_set || !g.points_set)
        throw std::runtime_error("build_bridges: points and k-NN lists must be set");
    const auto t0 = Clock::now();
    std::uint32_t tagged = 0;
    g.score_k = p.score_k;
    hubness_compute(g, &tagged);
    g.bu.clear();
    g.bv.clear();
    g.bw.clear();
    g.repair_count = 0;
    std::uint64_t knn_edges = 0;
    double bridge_s = 0.0;
    if (p.missing_mode > 2) throw std::runtime_error("build_bridges: unknown missing_mode");
    g.missing_mode = p.missing_mode;
    if (p.missing_mode != 0 && !g.inst_off.empty())
        throw std::runtime_error("build_bridges: missing_mode 1/2 needs a single-instance graph");
    if (g.inst_off.empty()) {
        knn_edges = count_knn_edges(g);
        const std::size_t budget = antihub_budget(p, g.dim, knn_edges, g.n);
        g.selected = select_antihubs(g, budget, p.min_per_component);
        if (g.yield_hook) g.yield_hook();
        const auto t1 = Clock::now();
        // coresg.cpp:148-166: the anti-hub selection is made in every mode
        // (the hubness index is refreshed regardless); the bridges follow the
        // mode.
        if (p.missing_mode == 0) {
            mst_over_subset(g, g.selected, &g.bu, &g.bv, &g.bw);
        } else if (p.missing_mode == 1) {
            exact_mst_bridges(g, &g.bu, &g.bv, &g.bw);
        }
        bridge_s += since(t1);
    } else {
        // per instance, as if built alone: budget from the instance's own
        // edge count, its lowest nodes, the quota winners of its components
        // (components never cross instances, so one pass finds them all),
        // the bridge MST over its selection
        const std::vector<std::uint64_t> edges = knn_edges_per_instance(g);
        const std::vector<std::uint32_t> quota = quota_winners(g, p.min_per_component);
        std::vector<std::vector<std::uint32_t>> quota_of(g.num_instances());
        for (std::uint32_t w : quota) {
            if (w == kNone) continue;
            const std::size_t i = static_cast<std::size_t>(
                std::upper_bound(g.inst_off.begin(), g.inst_off.end(), w) - g.inst_off.begin() - 1);
            quota_of[i].push_back(w);
        }
        std::vector<std::size_t> budgets(g.num_instances());
        for (std::uint32_t i = 0; i < g.num_instances(); ++i) {
            knn_edges += edges[i];
            budgets[i] = antihub_budget(p, g.dim, edges[i], g.inst_off[i + 1] - g.inst_off[i]);
        }
        const std::vector<std::vector<std::uint32_t>> sels =
            merge_selections(g, select_budget_instances(g, budgets), quota_of);
        g.selected.clear();
        for (const auto& sel : sels) g.selected.insert(g.selected.end(), sel.begin(), sel.end());
        const auto t1 = Clock::now();
        mst_over_subsets(g, sels, &g.bu, &g.bv, &g.bw);
        bridge_s += since(t1);
    }
    const double hub_s = since(t0) - bridge_s;
    const auto t1 = Clock::now();
    if (st) {
        st->knn_edges = knn_edges;
        // set in the anti-hub mode only (coresg.cpp:151-155)
        st->antihubs_selected = p.missing_mode == 0 ? g.selected.size() : 0;
        st->repair_edges = 0;
    }
    repair_connectivity(g, p, st);
    g.edges_ready = false;
    if (st) {
        st->hubness_seconds = hub_s;
        st->bridge_seconds = bridge_s + since(t1);
        st->bridge_edges = g.bu.size();
    }
}

}  // namespace coresg_hip
