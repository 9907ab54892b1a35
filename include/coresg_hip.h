/* C ABI of the gfx950 CORE-SG library (libcoresg_hip.so).
 *
 * Plain C types only: POD structs, raw pointers and sizes. Functions that can
 * fail return 1 on success and 0 on failure; the message of the most recent
 * failure on the calling thread is coresg_hip_last_error(). A message that
 * contains "out of memory" marks a device allocation failure (the host
 * side treats it as the reference does: record it, optionally fall back).
 *
 * Two groups of entry points:
 *
 *  1. Drop-in replacements for the reference's device ABI
 *     (include/coresg/cuda_api.hpp), same arguments and meaning, host
 *     pointers in and out. A `Backend` implementation calls these exactly
 *     where the reference's CUDA backend calls the cuda:: functions.
 *
 *  2. A device-resident CORE-SG (`coresg_hip_graph`): points, k-NN lists,
 *     bridges and the deduplicated edge list stay in HBM across calls, so a
 *     build and a whole mpts sweep run without host round trips. These extend
 *     the reference boundary, which only covers the k-NN descent and the MRD
 *     block; each names the reference host function whose result it
 *     reproduces.
 */
#ifndef CORESG_HIP_H
#define CORESG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CORESG_HIP_ABI_VERSION 1

/* Metric codes; replaces coresg::cuda::MetricCode (cuda_api.hpp:32-35). */
enum { CORESG_HIP_METRIC_L2 = 0, CORESG_HIP_METRIC_COSINE = 1 };

/* Replaces coresg::cuda::DeviceProps (cuda_api.hpp:17-25). `major`/`minor`
 * carry the gfx major/minor (gfx950 -> 9, 5). */
typedef struct coresg_hip_device_props {
    int id;
    char name[256];
    char arch[64];
    size_t total_mem;
    size_t free_mem;
    int cu_count;
    int major;
    int minor;
} coresg_hip_device_props;

/* Replaces coresg::cuda::NndConfig (cuda_api.hpp:40-51). */
typedef struct coresg_hip_nnd_config {
    uint32_t n;
    uint32_t dim;
    uint32_t stride;
    uint32_t k;
    uint32_t sample;
    uint32_t iters;
    uint64_t seed;
    int metric;
    float delta;
    int random_fill;
} coresg_hip_nnd_config;

int coresg_hip_abi_version(void);

/* ---- 1. reference device ABI --------------------------------------------- */

/* cuda_api.hpp:53 (device_count). */
int coresg_hip_device_count(void);
/* cuda_api.hpp:54 (get_device_props). */
int coresg_hip_get_device_props(int device, coresg_hip_device_props* out);
/* cuda_api.hpp:57 (last_error). */
const char* coresg_hip_last_error(void);
/* cuda_api.hpp:62-64 (nndescent_gpu): Dynamic NN-Descent on `device`, host
 * graph arrays refined in place (metric units), updates counted. */
int coresg_hip_nndescent(int device, const coresg_hip_nnd_config* cfg, const float* host_vectors,
                         uint32_t* host_ids, float* host_dists, uint8_t* host_flags,
                         uint32_t* host_counts, uint64_t* out_updates);
/* A device-memory budget for coresg_hip_nndescent on `device` (0 = none):
 * every device allocation the call makes is checked against it first, and a
 * call that would exceed it fails with an "out of memory" message, which the
 * Backend treats as the reference treats a device OOM (record it, fall back
 * to the host unless oom_fallback is off; cuda_backend.cpp:88-101). No
 * counterpart in cuda_api.hpp; the reference plans device memory per instance
 * (resource.cpp:72-84). */
int coresg_hip_set_device_budget(int device, size_t bytes);
/* cuda_api.hpp:68-70 (mrd_block_gpu): rows [r0,r1) of the m x m
 * mutual-reachability matrix into host_out ((r1-r0)*m floats). */
int coresg_hip_mrd_block(int device, const float* host_pts, size_t m, uint32_t dim,
                         uint32_t stride, int metric, const float* host_core, size_t r0,
                         size_t r1, float* host_out);
/* cuda_api.hpp:74-78 (mrd_row_min_gpu): per row in [r0,r1) the minimum MRD to
 * a column in another component (host_component may be NULL: any j != i),
 * first index on ties; +inf / 0xFFFFFFFF when none. */
int coresg_hip_mrd_row_min(int device, const float* host_pts, size_t m, uint32_t dim,
                           uint32_t stride, int metric, const float* host_core,
                           const uint32_t* host_component, size_t r0, size_t r1,
                           float* host_min, uint32_t* host_argmin);
/* cuda_api.hpp:82 (enable_peer_access). One GPU per process here: a no-op
 * that reports success for a single listed device. */
int coresg_hip_enable_peer_access(const int* devices, size_t count);
/* cuda_api.hpp:84 (device_memory). */
int coresg_hip_device_memory(int device, size_t* free_bytes, size_t* total_bytes);
/* cuda_api.hpp:86 (release_thread_cache): drop this thread's cached MRD
 * operands, and those of threads that have exited since (an exiting thread
 * makes no HIP call; its cache is freed here or by the next MRD call). */
void coresg_hip_release_thread_cache(void);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc): coresg_hip_mrd_block's result, with host operands staged
 * through one fixed 32 MiB array per GPU that the library holds until the
 * process ends, instead of device buffers it allocates. Points (m * stride),
 * core (m) and the rows ((r1 - r0) * m), each rounded up to 64 floats, must
 * fit coresg_hip_mrd_block_noalloc_capacity() floats; larger calls fail with a
 * message naming the limit. Calls on one GPU are serialised by a host lock
 * (they share the array); the null stream is used. */
int coresg_hip_mrd_block_noalloc(int device, const float* host_pts, size_t m, uint32_t dim,
                                 uint32_t stride, int metric, const float* host_core, size_t r0,
                                 size_t r1, float* host_out);
/* Floats of the fixed array behind coresg_hip_mrd_block_noalloc. */
size_t coresg_hip_mrd_block_noalloc_capacity(void);

/* ---- 2. device-resident CORE-SG ------------------------------------------ */

typedef struct coresg_hip_graph coresg_hip_graph;

/* Where a buffer argument lives. */
enum { CORESG_HIP_HOST = 0, CORESG_HIP_DEVICE = 1 };

/* Build parameters: the struct defaults of CoreSGParams / HubnessParams /
 * NNDescentParams (coresg.hpp:63-106, hubness.hpp:72-119, nndescent.hpp:45-71)
 * with the CLI's sample/iters (coresg_cli.cpp:154-158). */
typedef struct coresg_hip_build_params {
    uint32_t sample;
    uint32_t iters;
    double delta;
    uint64_t seed;
    uint32_t brute_force_limit;
    double antihub_multiplier;
    double max_pair_ops;
    size_t min_selected;
    size_t max_selected;
    size_t min_per_component;
    int ensure_connected;
    size_t reps_per_component;
    size_t max_reps;
    /* local join: per-node proposal buffer capacity per round (16..512;
     * 0 = library default, 512) */
    uint32_t join_capacity;
    /* local-join distances: 0 = library choice (exact for rows of up to 64
     * floats, the FP16 filter for wider rows), 1 = exact FP32 for every
     * pair, 2 = FP16 Gram on the matrix cores (gfx950 16x16x32) as a filter,
     * 3 = BF16 Gram as a filter. The filters use a rigorous error bound to
     * skip only pairs that cannot pass either threshold test and compute every
     * other pair in exact FP32, so all modes propose the same candidates.
     * Strides the filter join is not built for (other than 8, 16, 24, 32, 48,
     * 64, 96, 128) take the exact join. */
    uint32_t gram;
    /* bridge set (MissingEdgeMode, coresg.hpp:54-58; coresg.cpp:143-169):
     * 0 = AntiHub (default), 1 = ExactMst (the exact MST at k_max, stored
     * with base distances), 2 = None. Modes 1 and 2 apply to single-instance
     * graphs. */
    uint32_t missing_mode;
    /* insert: where a new point's neighbours among the existing points come
     * from. 0 = an exact scan of every point (the lists are the exact k
     * nearest); 1 = a beam search over the existing lists, entered at the
     * closest of up to 256 selected anti-hubs and 8 random points, with a
     * 64-entry beam (the reference's anchored seeding,
     * incremental_knn.cpp:118-216); the batch's own pairs are scanned
     * exactly in both modes, and every visited existing point is offered the
     * new point (try_insert semantics). */
    uint32_t insert_search;
    /* the k of the score's degenerate case (HubnessParams::k_for_score,
     * hubness.hpp:83; 0 = the graph's k_max, which CoreSG::build sets) */
    uint32_t score_k;
} coresg_hip_build_params;

typedef struct coresg_hip_build_stats {
    double knn_seconds;
    double hubness_seconds;
    double bridge_seconds;
    double total_seconds;
    uint32_t rounds;
    uint64_t updates;
    uint64_t knn_edges;
    uint64_t antihubs_selected;
    uint64_t bridge_edges;
    uint64_t repair_edges;
    uint32_t components_before_repair;
    uint32_t components_after_repair;
} coresg_hip_build_stats;

/* One extracted MST (MstResult, mst.hpp:77-85). */
typedef struct coresg_hip_mst_info {
    uint32_t mpts;
    uint32_t num_edges;
    uint32_t num_components;
    uint32_t rounds;
    double weight;
} coresg_hip_mst_info;

void coresg_hip_default_build_params(coresg_hip_build_params* p);

coresg_hip_graph* coresg_hip_graph_create(int device, uint32_t n, uint32_t dim, uint32_t k_max);
void coresg_hip_graph_destroy(coresg_hip_graph* g);
/* Rows: n x dim floats (padded=0) or n x stride floats (padded=1). */
int coresg_hip_graph_set_points(coresg_hip_graph* g, const float* rows, int where, int padded);
/* k-NN lists in the reference SoA layout (knn_graph.hpp:188-195). */
int coresg_hip_graph_set_knn(coresg_hip_graph* g, const uint32_t* ids, const float* dist,
                             const uint8_t* flags, const uint32_t* count, int where);
int coresg_hip_graph_get_knn(coresg_hip_graph* g, uint32_t* ids, float* dist, uint8_t* flags,
                             uint32_t* count);
/* Bridge edges (CoreSG::missing_edges, coresg.hpp:153), base distances. */
int coresg_hip_graph_set_bridges(coresg_hip_graph* g, const uint32_t* u, const uint32_t* v,
                                 const float* w, size_t m, size_t repair_count);
size_t coresg_hip_graph_get_bridges(coresg_hip_graph* g, uint32_t* u, uint32_t* v, float* w,
                                    size_t cap, size_t* repair_count);
/* Device pointers of the resident arrays (for callers that keep data in HBM). */
int coresg_hip_graph_device_arrays(coresg_hip_graph* g, float** pts, uint32_t** ids,
                                   float** dist, uint8_t** flags, uint32_t** count);
size_t coresg_hip_graph_device_bytes(const coresg_hip_graph* g);

/* k-NN half: init_random + Dynamic NN-Descent on device (replaces
 * IncrementalKnn::build's single-instance path, incremental_knn.cpp:1012-1050,
 * and CpuBackend::run_knn_job, cpu_backend.cpp:47-59), or exact brute force
 * when n <= brute_force_limit (incremental_knn.cpp:997-1007). */
int coresg_hip_build_knn(coresg_hip_graph* g, const coresg_hip_build_params* p,
                         coresg_hip_build_stats* st);
/* Hubness, anti-hub selection, bridge MST and repair (CoreSG::build after
 * the k-NN half, coresg.cpp:278-305). */
int coresg_hip_build_bridges(coresg_hip_graph* g, const coresg_hip_build_params* p,
                             coresg_hip_build_stats* st);
/* Both halves (CoreSG::build, coresg.cpp:260-323). */
int coresg_hip_build(coresg_hip_graph* g, const coresg_hip_build_params* p,
                     coresg_hip_build_stats* st);

/* CoreSG::insert (coresg.cpp:325-389; IncrementalKnn::insert,
 * incremental_knn.cpp:1155-1252): append `count` rows (layout as
 * set_points) to a built CORE-SG. The new points' lists are their exact k
 * nearest among all points, by (distance, id) as brute_force_knn orders them
 * (nndescent.cpp:242-263); every new point is offered to the existing lists it
 * belongs in (try_insert semantics, knn_graph.cpp:37-55); then hubness,
 * selection, bridges and repair are recomputed over the grown graph (the
 * reference refreshes them incrementally with identical results,
 * test_hubness.cpp:95-163). With n <= brute_force_limit every list is rebuilt
 * exactly, as the reference does (incremental_knn.cpp:1169-1180). */
int coresg_hip_insert(coresg_hip_graph* g, const float* rows, uint32_t count, int where,
                      int padded, const coresg_hip_build_params* p, coresg_hip_build_stats* st);
/* Current number of points (grows with coresg_hip_insert). */
uint32_t coresg_hip_graph_size(const coresg_hip_graph* g);

/* The derived hubness state alone, from the k-NN lists (points need not be
 * set; the graph's dim is the budget dimension): ConnectivityTracker::rebuild
 * (connectivity.cpp:24-139), HubnessIndex::build (hubness.cpp:44-158) and
 * rescore_all with the anti-hub budget and quota (hubness.cpp:16-40,
 * 424-521) under p's antihub_multiplier, max_pair_ops, min/max_selected and
 * min_per_component. Read back with coresg_hip_graph_get_hubness /
 * _get_selected. */
int coresg_hip_graph_derive_hubness(coresg_hip_graph* g, const coresg_hip_build_params* p);
size_t coresg_hip_graph_get_selected(coresg_hip_graph* g, uint32_t* out, size_t cap);
int coresg_hip_graph_get_hubness(coresg_hip_graph* g, int32_t* rdeg, double* score,
                                 uint32_t* conn_labels, uint32_t* num_components);

/* Deduplicated CORE-SG edge list (knn_graph.cpp:127-200 + bridges), built on
 * device once and kept resident, sorted by (u, v). */
int coresg_hip_prepare_edges(coresg_hip_graph* g, uint64_t* num_edges, uint64_t* knn_edges);
/* CoreSG::edges(mpts) (coresg.cpp:391-406) in (u, v) order. */
int coresg_hip_edges(coresg_hip_graph* g, uint32_t mpts, uint32_t* u, uint32_t* v, float* w);
/* CoreSG::mst for each mpts (coresg.cpp:408-417; the MST is the unique one
 * under edge_less, types.hpp:38-42, i.e. the reference's Kruskal tree).
 * Outputs per mpts i: info[i]; when non-NULL, out_u/v/w + i*(n-1) receive
 * the edges in (u, v) order and out_labels + i*n the dense component labels
 * (first appearance, mst.cpp:48-59). */
int coresg_hip_extract(coresg_hip_graph* g, const uint32_t* mpts, uint32_t count,
                       coresg_hip_mst_info* info, uint32_t* out_u, uint32_t* out_v,
                       float* out_w, uint32_t* out_labels);

/* CoreSG::exact_mst (coresg.cpp:419-430): Prim over the complete mutual-
 * reachability graph of the resident points (mst_prim_virtual, mst.cpp:172-220,
 * with CpuBackend::mrd_block rows, cpu_backend.cpp:61-90) at min(mpts, k_max),
 * keeping Prim's tie rules. out_u/v/w receive up to n-1 edges in acceptance
 * order (u < v), out_labels n dense labels by first appearance; any output may
 * be NULL. info->rounds = Prim steps. */
int coresg_hip_exact_mst(coresg_hip_graph* g, uint32_t mpts, coresg_hip_mst_info* info,
                         uint32_t* out_u, uint32_t* out_v, float* out_w, uint32_t* out_labels);
/* How coresg_hip_exact_mst launches its Prim steps: 0 (default) one
 * cooperative launch that runs every step with its blocks exchanging keys on
 * the device; 1 one launch per step (the kernel boundary as the step's
 * barrier, batches of launches replayed from a captured hipGraph), for callers
 * that avoid cooperative launches. Both give the same edges. No counterpart
 * in the reference. */
int coresg_hip_graph_set_exact_launch(coresg_hip_graph* g, int mode);

/* CoreSG::save (io.cpp:232-260): the resident graph in the reference's
 * .coresg format, byte for byte (metric L2, the bridges' missing-edge mode). */
int coresg_hip_graph_save(coresg_hip_graph* g, const char* path);
/* CoreSG::load (io.cpp:262-330): a graph on `device` from a .coresg file
 * (L2 only); the derived hubness state -- components, reverse degrees, scores
 * and the anti-hub selection -- is recomputed with the default parameters, as
 * the reference's load does. NULL on failure (coresg_hip_last_error). */
coresg_hip_graph* coresg_hip_graph_load(int device, const char* path);
/* write_edges_tsv (io.cpp:131-138): "u\tv\tweight" then one line per edge,
 * the weight as %.9e. Host-only; no device needed. */
int coresg_hip_write_edges_tsv(const char* path, const uint32_t* u, const uint32_t* v,
                               const float* w, size_t m);
/* write_labels_tsv (io.cpp:140-145): "id\tlabel" then one line per point. */
int coresg_hip_write_labels_tsv(const char* path, const uint32_t* labels, size_t n);

/* ---- autotuning ----------------------------------------------------------- */

/* The launch shape chosen for the mpts sweep's Borůvka round kernel. */
typedef struct coresg_hip_tune_result {
    uint32_t slots;      /* mpts values per pass the shape was tuned for */
    uint32_t lanes;      /* lanes serving one vertex (4..64, >= slots rounded up) */
    uint32_t max_blocks; /* grid cap of the round kernel */
    uint32_t candidates; /* shapes timed */
    double best_ms;      /* one extract of the slots with the chosen shape */
    double default_ms;   /* the same with the untuned shape */
} coresg_hip_tune_result;

/* Times every candidate launch shape of the sweep's round kernel (lanes per
 * vertex x grid cap) on this graph with the first min(count, 32) values of
 * `mpts` and keeps the fastest, for this graph and for every later graph of
 * the same size bucket (floor(log2 n), k_max) and slot count. Results never
 * depend on the shape. */
int coresg_hip_autotune_sweep(coresg_hip_graph* g, const uint32_t* mpts, uint32_t count,
                              coresg_hip_tune_result* out);
/* The process-wide table of tuned shapes as text, one line per bucket:
 * "log2n k_max slots lanes max_blocks". Writes up to cap bytes (NUL
 * terminated when room) and returns the length of the whole table. */
size_t coresg_hip_tune_table(char* buf, size_t cap);
/* Adds the lines of such a table (a table saved by an earlier process). */
int coresg_hip_tune_table_load(const char* text);

/* ---- measurement ---------------------------------------------------------- */

/* A probe of known algorithmic bytes for calibrating hardware byte counters
 * against the path's access pattern: `gathers` reads of whole random rows of a
 * rows x stride float array on `device` (stride * 4 bytes each, the pattern of
 * the local join's and the insert search's row reads), each writing one float.
 * *ms receives the kernel's device time. No counterpart in the reference. */
int coresg_hip_probe_gather(int device, size_t rows, uint32_t stride, size_t gathers,
                            double* ms);
/* The lane layout of the matrix-core instruction the filter join uses
 * (gfx950 v_mfma_f32_16x16x32_{f16,bf16}): c (16 x 16) = a b^T for a, b of
 * 16 x 32 floats (row-major, rounded to BF16 when bf16 != 0, else FP16),
 * computed by one wavefront with the join's operand and result layout. No
 * counterpart in the reference. */
int coresg_hip_probe_mfma(int device, int bf16, const float* a, const float* b, float* c);

/* ---- 3. serving session -------------------------------------------------- */

/* One device-resident CORE-SG served by continuous batching. The reference
 * serves inserts and MST queries as separate CLI runs that reload the CORE-SG
 * each time (coresg_cli.cpp:507-570); here the graph stays in HBM and two
 * device workers run concurrently:
 *   - inserts are applied one at a time in submission order to a working
 *     graph (CoreSG::insert); each completed insert publishes a new version
 *     (graph_version counts the inserts applied);
 *   - extracts run on the latest published version while the next one is
 *     built; every extract waiting when the extract worker becomes free is
 *     admitted into the next fused sweep (up to 32 mpts values). Every mpts
 *     is computed independently, so a result is the same alone and in a mix,
 *     and equals CoreSG::mst on the version it reports (graph_version,
 *     graph_n) -- an extract does not wait for inserts submitted before it.
 * An insert is admitted while the analytic device footprint of the next
 * version, plus older versions still held by in-flight extracts, stays within
 * hbm_budget_bytes (0 = no limit); it waits while those older versions are
 * released and fails, with a message naming the budget, only when it cannot
 * fit even then. */
typedef struct coresg_hip_session coresg_hip_session;

enum { CORESG_HIP_REQ_INSERT = 1, CORESG_HIP_REQ_EXTRACT = 2 };
enum {
    CORESG_HIP_REQ_QUEUED = 0,
    CORESG_HIP_REQ_RUNNING = 1,
    CORESG_HIP_REQ_DONE = 2,
    CORESG_HIP_REQ_FAILED = 3
};

typedef struct coresg_hip_request_status {
    uint64_t id;
    int kind;
    int state;
    uint32_t graph_n;       /* points in the version it produced (insert) or read (extract) */
    uint64_t graph_version; /* inserts applied in that version */
    uint32_t num_msts;      /* extract: mpts values requested */
    uint32_t batch_requests; /* requests served by the same device pass */
    double queued_ms;       /* submit -> start */
    double run_ms;          /* start -> done */
    double latency_ms;      /* submit -> done */
    coresg_hip_build_stats insert_stats;
    char error[256];
} coresg_hip_request_status;

/* Build a CORE-SG from host rows (CoreSG::build) and start serving it. */
coresg_hip_session* coresg_hip_session_create(int device, const float* rows, uint32_t n,
                                              uint32_t dim, uint32_t k_max,
                                              const coresg_hip_build_params* p,
                                              size_t hbm_budget_bytes,
                                              coresg_hip_build_stats* st);
/* Serve an existing built graph; the session takes ownership of it. */
coresg_hip_session* coresg_hip_session_adopt(coresg_hip_graph* g, const coresg_hip_build_params* p,
                                             size_t hbm_budget_bytes);
/* Queue a request (rows / mpts are copied); returns its id, 0 on failure. */
uint64_t coresg_hip_session_submit_insert(coresg_hip_session* s, const float* rows,
                                          uint32_t count);
uint64_t coresg_hip_session_submit_extract(coresg_hip_session* s, const uint32_t* mpts,
                                           uint32_t count, int want_edges, int want_labels);
/* Wait up to timeout_s (< 0: no limit): 1 finished (done or failed), 0 not
 * yet, -1 error (unknown id). `st` receives the request's status. */
int coresg_hip_session_wait(coresg_hip_session* s, uint64_t id, double timeout_s,
                            coresg_hip_request_status* st);
/* A completed extract's results, laid out as coresg_hip_extract's. */
int coresg_hip_session_extract_result(coresg_hip_session* s, uint64_t id,
                                      coresg_hip_mst_info* info, uint32_t* out_u,
                                      uint32_t* out_v, float* out_w, uint32_t* out_labels);
/* Forget a finished request. */
int coresg_hip_session_release(coresg_hip_session* s, uint64_t id);
/* The served graph, once the queue has drained (for inspection only; submit
 * nothing while it is in use). */
coresg_hip_graph* coresg_hip_session_graph(coresg_hip_session* s);
/* Device bytes the session holds now: the working graph plus every published
 * version still alive. */
size_t coresg_hip_session_device_bytes(coresg_hip_session* s);
/* Finish the queued requests, stop the workers, free the graph. */
void coresg_hip_session_destroy(coresg_hip_session* s);

/* ---- 4. batched independent instances ------------------------------------ */

/* Many independent point sets in one graph (the batched configuration of
 * uneven instances): instance i holds points [offsets[i], offsets[i+1]),
 * offsets[0] = 0 and offsets[count] = the graph's point count. Every CoreSG
 * step then runs per instance, as if each were built alone (CoreSG::build,
 * coresg.cpp:260-323, per instance): init_random draws inside the instance
 * with the instance-local node index (bit-exact against a separate build),
 * the descent's candidates never leave an instance, the anti-hub budget,
 * selection, bridges and repair are per instance; the descent's early stop is
 * taken over all instances together. One sweep over the shared edge list
 * gives every instance's MSTs. Call before building; count = 0 clears. */
int coresg_hip_graph_set_instances(coresg_hip_graph* g, const uint32_t* offsets, uint32_t count);
/* CoreSG::mst per instance and mpts value (coresg.cpp:408-417): info[i * count
 * + q] for instance i and mpts[q] (edges, components, exact weight, rounds).
 * The union's edges and labels (coresg_hip_extract) hold every instance's
 * MST edges in global ids, grouped by instance. */
int coresg_hip_extract_instances(coresg_hip_graph* g, const uint32_t* mpts, uint32_t count,
                                 coresg_hip_mst_info* info);

/* Sampled k-NN recall on device (eval.cpp:24-97; queries chosen on the
 * host by the same mt19937_64 draw). */
int coresg_hip_sampled_recall(coresg_hip_graph* g, size_t queries, uint64_t seed,
                              double* recall, double* distance_ratio);

/* The plain one-launch-per-step form of every graph operation, the shape of
 * the reference's CUDA driver loop (nndescent_kernels.cu:407-418: launch,
 * synchronise, read back, decide): with on != 0 every kernel launch is
 * followed by a host synchronisation, the descent's convergence is read after
 * each round before the next is enqueued, and the sweep runs one mpts value
 * per pass with each Boruvka round launched and checked from the host (no
 * captured graphs). Results are identical to the default (fused) form; only
 * the scheduling differs. Off by default. */
int coresg_hip_graph_set_plain(coresg_hip_graph* g, int on);

/* Per-kernel device timing with HIP events recorded on the graph's own
 * stream around each launch of the named kernels (off by default). */
int coresg_hip_graph_set_profiling(coresg_hip_graph* g, int on);
int coresg_hip_graph_kernel_times(coresg_hip_graph* g, const char* kernel, double* avg_ms,
                                  uint64_t* launches);
/* Work tallies accumulated while profiling, per kernel name (four u64):
 * "nnd_join": {members staged, exact distances evaluated, proposals
 * written, nodes}; "boruvka_round1": {adjacency entries, vertices, 0, 0};
 * "nnd_gram" (Gram-filter joins): {pairs visited, pairs computed exactly,
 * the largest relative error of a fast distance against FP64 (float bits),
 * the largest ratio of the bound's lower distance to the FP64 distance
 * (float bits; at most 1 when the bound holds)}; "insert_gram" (Gram-filter
 * insert scans): {pairs visited, pairs computed exactly, 0, 0}. */
int coresg_hip_graph_kernel_work(coresg_hip_graph* g, const char* kernel, uint64_t* work4);
int coresg_hip_graph_reset_timers(coresg_hip_graph* g);
/* While profiling, also audit the Gram-filter join: an FP64 distance for
 * every pair it visits, the worst relative error of the fast distance and the
 * worst bound ratio recorded under "nnd_gram" (off by default; slow). */
int coresg_hip_graph_set_audit(coresg_hip_graph* g, int on);

#ifdef __cplusplus
}
#endif
#endif
