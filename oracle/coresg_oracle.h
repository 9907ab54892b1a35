/* CPU restatement of the reference CORE-SG hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the oracle: tests, the smoke check and the CPU-baseline leg
 * of the benchmark call it to check (or time beside) the HIP path. It is never
 * linked into, loaded by or called from the product library.
 *
 * Every function restates one reference function in plain C, single-threaded,
 * following the reference's serial schedule (file:line cited per function in
 * coresg_oracle.c). Parity is pinned against golden vectors written by the
 * reference itself (oracle/ref_dump.cpp, tests/golden/).
 *
 * Conventions follow the reference: node ids are u32 with 0xFFFFFFFF as "none";
 * unfilled list slots hold +inf; distances are the true L2 metric (sqrt of the
 * squared distance); rows are zero-padded to `stride` = round_up(dim, 8).
 */
#ifndef CORESG_ORACLE_H
#define CORESG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_INVALID 0xFFFFFFFFu

/* ---- primitives ---------------------------------------------------------- */
float orc_l2sqr(const float* a, const float* b, uint32_t stride);
float orc_distance(const float* a, const float* b, uint32_t stride);
void orc_pair_distances(const float* data, uint32_t stride, const uint32_t* pairs, size_t npairs,
                        float* out_l2sqr, float* out_dist);
uint64_t orc_noderng_first(uint64_t seed, uint32_t node, uint32_t round, uint32_t count,
                           uint32_t bound, uint32_t* out);

/* ---- k-NN graph (SoA: ids/dist/flags n*k, count n) ----------------------- */
int orc_try_insert(uint32_t* ids, float* dist, uint8_t* flags, uint32_t* count, uint32_t k,
                   uint32_t node, uint32_t cand, float d, int is_new, uint32_t* evicted);
void orc_init_random(const float* data, uint32_t n, uint32_t stride, uint32_t k, uint64_t seed,
                     uint32_t* ids, float* dist, uint8_t* flags, uint32_t* count);
void orc_nndescent(const float* data, uint32_t n, uint32_t stride, uint32_t k, uint32_t sample,
                   uint32_t iters, double delta, uint32_t* ids, float* dist, uint8_t* flags,
                   uint32_t* count, uint64_t* stats /* rounds, updates, evals */);
void orc_brute_force(const float* data, uint32_t n, uint32_t stride, uint32_t k, uint32_t* ids,
                     float* dist, uint8_t* flags, uint32_t* count);
int orc_validate(const uint32_t* ids, const float* dist, const uint32_t* count, uint32_t n,
                 uint32_t k);

/* ---- extraction ----------------------------------------------------------- */
void orc_core_distances(const float* dist, const uint32_t* count, uint32_t n, uint32_t k,
                        uint32_t mpts, float* out);
size_t orc_undirected_edges(const uint32_t* ids, const float* dist, const uint32_t* count,
                            uint32_t n, uint32_t k, uint32_t* u, uint32_t* v, float* w);
size_t orc_mst_kruskal(uint32_t n, const uint32_t* eu, const uint32_t* ev, const float* ew,
                       size_t m, uint32_t* out_u, uint32_t* out_v, float* out_w,
                       uint32_t* out_labels, uint32_t* out_ncomp, double* out_weight);

/* ---- hubness / selection / components ------------------------------------ */
void orc_reverse_degrees(const uint32_t* ids, const uint32_t* count, uint32_t n, uint32_t k,
                         int32_t* rdeg);
double orc_score(const uint32_t* ids, const uint32_t* count, uint32_t n, uint32_t k,
                 uint32_t k_for_score, const int32_t* rdeg, uint32_t v);
size_t orc_antihub_budget(double multiplier, uint32_t dim, size_t num_edges, uint32_t n,
                          double max_pair_ops, size_t max_selected, size_t min_selected,
                          size_t* uncapped, int* capped);
size_t orc_select(uint32_t n, const double* score, const int32_t* rdeg, size_t budget,
                  const uint32_t* comp, uint32_t ncomp, size_t per, uint32_t* out);
uint32_t orc_components(const uint32_t* ids, const uint32_t* count, uint32_t n, uint32_t k,
                        uint32_t* labels);

/* ---- bridges -------------------------------------------------------------- */
void orc_mrd_block(const float* pts, size_t m, uint32_t stride, const float* core, size_t r0,
                   size_t r1, float* out);
size_t orc_prim_dense(const float* matrix, size_t m, uint32_t* out_u, uint32_t* out_v,
                      float* out_w);

/* ---- whole CORE-SG (build / load / save / extract) ------------------------ */
typedef struct orc_params {
    uint32_t k_max;
    uint32_t sample;
    uint32_t iters;
    double delta;
    uint64_t nnd_seed;
    uint32_t brute_force_limit;
    double antihub_multiplier;
    double max_pair_ops;
    size_t min_selected;
    size_t max_selected;
    size_t min_per_component;
    int ensure_connected;
    size_t reps_per_component;
    size_t max_reps;
} orc_params;

typedef struct orc_sg orc_sg;

void orc_default_params(orc_params* p);
orc_sg* orc_sg_build(const float* rows, uint32_t n, uint32_t dim, const orc_params* p);
orc_sg* orc_sg_load(const char* path, const orc_params* p);
int orc_sg_save(const orc_sg* sg, const char* path);
void orc_sg_free(orc_sg* sg);

uint32_t orc_sg_n(const orc_sg* sg);
uint32_t orc_sg_k(const orc_sg* sg);
uint32_t orc_sg_dim(const orc_sg* sg);
uint32_t orc_sg_stride(const orc_sg* sg);
const float* orc_sg_points(const orc_sg* sg);
void orc_sg_graph(const orc_sg* sg, const uint32_t** ids, const float** dist,
                  const uint8_t** flags, const uint32_t** count);
size_t orc_sg_bridges(const orc_sg* sg, const uint32_t** u, const uint32_t** v, const float** w);
size_t orc_sg_repair_count(const orc_sg* sg);
size_t orc_sg_selected(const orc_sg* sg, const uint32_t** sel);
const double* orc_sg_scores(const orc_sg* sg);
const int32_t* orc_sg_rdeg(const orc_sg* sg);
const uint32_t* orc_sg_conn_labels(const orc_sg* sg, uint32_t* ncomp);
size_t orc_sg_knn_edges(const orc_sg* sg);
void orc_sg_stats(const orc_sg* sg, double* out /* knn_s, hub_s, bridge_s, total_s, rounds,
                                                    updates, evals */);

/* Recompute selection-derived bridges (and repair) from the current graph,
 * as CoreSG::build does after the k-NN half (coresg.cpp:278-305). */
void orc_sg_rebuild_bridges(orc_sg* sg);

/* Full CORE-SG edge list reweighted for mpts (CoreSG::edges). Returns |E|;
 * pass NULL outputs to count only. */
size_t orc_sg_edges(const orc_sg* sg, uint32_t mpts, uint32_t* u, uint32_t* v, float* w);
/* CoreSG::mst(mpts) with the single-threaded (Kruskal) schedule. */
size_t orc_sg_mst(const orc_sg* sg, uint32_t mpts, uint32_t* out_u, uint32_t* out_v,
                  float* out_w, uint32_t* out_labels, uint32_t* out_ncomp, double* out_weight);
size_t orc_sg_exact_mst(const orc_sg* sg, uint32_t mpts, uint32_t* out_u, uint32_t* out_v,
                        float* out_w, uint32_t* out_labels, uint32_t* out_ncomp,
                        double* out_weight);
/* eval.cpp sampled_recall: returns recall, writes distance ratio. */
double orc_sg_sampled_recall(const orc_sg* sg, size_t queries, uint64_t seed, double* ratio);

/* std::mt19937_64 (the standard-specified engine), exposed for tests. */
void orc_mt64_first(uint64_t seed, size_t count, uint64_t* out);

#ifdef __cplusplus
}
#endif
#endif
