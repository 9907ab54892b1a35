// Build-side entry points (k-NN descent, hubness, bridges) and the
// reference-shaped host-pointer device calls.
#pragma once

#include "../../include/coresg_hip.h"
#include "graph.hpp"

namespace coresg_hip {

void build_knn(Graph& g, const coresg_hip_build_params& p, coresg_hip_build_stats* st);
void build_bridges(Graph& g, const coresg_hip_build_params& p, coresg_hip_build_stats* st);
// The derived state CoreSG::load recomputes (io.cpp:316-328): components,
// reverse degrees, scores and the anti-hub selection, bridges untouched.
void derive_hubness(Graph& g, const coresg_hip_build_params& p);
// CoreSG::insert (coresg.cpp:325-389): append rows, exact k-NN of the batch
// against every point with the batch offered to the existing lists, then
// hubness, selection, bridges and repair over the grown graph.
void insert_points(Graph& g, const float* rows, std::uint32_t count, int where, int padded,
                   const coresg_hip_build_params& p, coresg_hip_build_stats* st);
void sampled_recall(Graph& g, std::size_t queries, std::uint64_t seed, double* recall,
                    double* ratio);

void nndescent_host(int device, const coresg_hip_nnd_config& cfg, const float* vectors,
                    std::uint32_t* ids, float* dists, std::uint8_t* flags, std::uint32_t* counts,
                    std::uint64_t* updates);
void mrd_block_host(int device, const float* pts, std::size_t m, std::uint32_t dim,
                    std::uint32_t stride, int metric, const float* core, std::size_t r0,
                    std::size_t r1, float* out);
void mrd_row_min_host(int device, const float* pts, std::size_t m, std::uint32_t dim,
                      std::uint32_t stride, int metric, const float* core,
                      const std::uint32_t* component, std::size_t r0, std::size_t r1,
                      float* out_min, std::uint32_t* out_argmin);
void release_thread_cache();
// Rows [r0, r0 + rows) of the m x m MRD matrix of device-resident operands,
// launched on s (mrd.hip); `cosine`: the metric value is r / 2 instead of
// sqrt(r).
void launch_mrd(const float* pts, std::uint32_t m, std::uint32_t stride, const float* core,
                std::uint32_t r0, std::uint32_t rows, float* out, hipStream_t s,
                bool cosine = false);

}  // namespace coresg_hip
