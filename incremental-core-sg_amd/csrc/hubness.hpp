// Device hubness, components and anti-hub selection.
#pragma once

#include <vector>

#include "graph.hpp"

namespace coresg_hip {

// Reverse degrees, component labels (dense, first appearance) and scores of
// every node into g.rdeg / g.conn / g.score; g.num_components set. `tagged`
// receives the number of nodes scoring below 1.
void hubness_compute(Graph& g, std::uint32_t* tagged);

// The `budget` lowest nodes by (score, rdeg, id) plus the `per` lowest of every
// component, in that order (HubnessIndex::select).
std::vector<std::uint32_t> select_antihubs(Graph& g, std::size_t budget, std::size_t per);
// Its parts: the `budget` lowest of the nodes [lo, lo + cnt) (global ids, any
// order); the `per` lowest of every component (ncomp * per, kNone padded);
// their union in (score, rdeg, id) order.
std::vector<std::uint32_t> select_budget(Graph& g, std::uint32_t lo, std::uint32_t cnt,
                                         std::size_t budget);
std::vector<std::uint32_t> quota_winners(Graph& g, std::size_t per);
// merge_selection for several instances at once (one gather).
std::vector<std::vector<std::uint32_t>> merge_selections(
    Graph& g, std::vector<std::vector<std::uint32_t>> sels,
    const std::vector<std::vector<std::uint32_t>>& quotas);
// select_budget for every instance at once (one block per instance).
std::vector<std::vector<std::uint32_t>> select_budget_instances(
    Graph& g, const std::vector<std::size_t>& budgets);
// mst_over_subset for several subsets at once: every subset's matrix in one
// launch, every subset's Prim in one launch (a block each); edges appended
// subset after subset.
void mst_over_subsets(Graph& g, const std::vector<std::vector<std::uint32_t>>& sets,
                      std::vector<std::uint32_t>* u, std::vector<std::uint32_t>* v,
                      std::vector<float>* w);
std::vector<std::uint32_t> merge_selection(Graph& g, std::vector<std::uint32_t> sel,
                                           const std::vector<std::uint32_t>& quota);

// MST over the subset `ids` in mutual-reachability space at k_max, edges in
// Prim acceptance order, stored with the base distance in global ids
// (CoreSG::mst_over_subset). Appends to u/v/w.
void mst_over_subset(Graph& g, const std::vector<std::uint32_t>& ids,
                     std::vector<std::uint32_t>* u, std::vector<std::uint32_t>* v,
                     std::vector<float>* w);

}  // namespace coresg_hip
