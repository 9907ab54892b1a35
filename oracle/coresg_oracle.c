/* CPU restatement of the reference CORE-SG hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Used by tests/, the smoke check and the CPU-baseline leg of bench.py as the
 * checker. Nothing in the product library links or calls it.
 *
 * Every function is checked against the reference's own vectors or closed
 * forms (tests/test_oracle_golden.py).
 *
 * Each function names the reference function it restates (path:line relative
 * to the reference repository). The schedule is the reference's serial one
 * (`--threads 1`), which is the only schedule whose k-NN lists are
 * deterministic there. Built with -ffp-contract=off: every fused multiply-add
 * below is written out with fmaf() where the reference's AVX2 code issues one.
 */
#include "coresg_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ========================================================================== */
/* primitives                                                                 */
/* ========================================================================== */

/* include/coresg/distance.hpp:127-147 (l2sqr_avx2) with hsum256 at :116-125.
 * Two 8-lane accumulators over 16-float chunks, one 8-float step into acc0,
 * then ((v0+v4)+(v1+v5)) + ((v2+v6)+(v3+v7)) of v = acc0 + acc1. Rows are
 * padded to a multiple of 8 (dataset.hpp:21), so the scalar tail never runs
 * for a padded stride. */
float orc_l2sqr(const float* a, const float* b, uint32_t d) {
    float acc0[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    float acc1[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t i = 0;
    for (; i + 16 <= d; i += 16) {
        for (int j = 0; j < 8; ++j) {
            const float d0 = a[i + j] - b[i + j];
            const float d1 = a[i + 8 + j] - b[i + 8 + j];
            acc0[j] = fmaf(d0, d0, acc0[j]);
            acc1[j] = fmaf(d1, d1, acc1[j]);
        }
    }
    for (; i + 8 <= d; i += 8) {
        for (int j = 0; j < 8; ++j) {
            const float d0 = a[i + j] - b[i + j];
            acc0[j] = fmaf(d0, d0, acc0[j]);
        }
    }
    float v[8];
    for (int j = 0; j < 8; ++j) v[j] = acc0[j] + acc1[j];
    const float l0 = v[0] + v[4], l1 = v[1] + v[5], l2 = v[2] + v[6], l3 = v[3] + v[7];
    float s = (l0 + l1) + (l2 + l3);
    for (; i < d; ++i) {
        const float t = a[i] - b[i];
        s = s + t * t;
    }
    return s;
}

/* distance.hpp:196-201 (MetricOp<L2>::finalize = sqrt) and :243-248. */
float orc_distance(const float* a, const float* b, uint32_t stride) {
    return sqrtf(orc_l2sqr(a, b, stride));
}

void orc_pair_distances(const float* data, uint32_t stride, const uint32_t* pairs, size_t npairs,
                        float* out_l2sqr, float* out_dist) {
    for (size_t i = 0; i < npairs; ++i) {
        const float* a = data + (size_t)pairs[2 * i] * stride;
        const float* b = data + (size_t)pairs[2 * i + 1] * stride;
        const float r = orc_l2sqr(a, b, stride);
        if (out_l2sqr) out_l2sqr[i] = r;
        if (out_dist) out_dist[i] = sqrtf(r);
    }
}

/* src/nndescent.cpp:12-29: splitmix64 (Vigna) and the per-node NodeRng. */
static uint64_t splitmix64(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

typedef struct {
    uint64_t s;
} node_rng;

static void node_rng_init(node_rng* r, uint64_t seed, uint32_t node, uint32_t round) {
    r->s = seed ^ (0x9E3779B97F4A7C15ull * ((uint64_t)node + 1)) ^
           (0xC2B2AE3D27D4EB4Full * ((uint64_t)round + 1));
    splitmix64(&r->s);
}
static uint32_t node_rng_next(node_rng* r) { return (uint32_t)(splitmix64(&r->s) >> 32); }
static uint32_t node_rng_below(node_rng* r, uint32_t bound) {
    return bound ? node_rng_next(r) % bound : 0;
}

uint64_t orc_noderng_first(uint64_t seed, uint32_t node, uint32_t round, uint32_t count,
                           uint32_t bound, uint32_t* out) {
    node_rng r;
    node_rng_init(&r, seed, node, round);
    const uint64_t state = r.s;
    for (uint32_t i = 0; i < count; ++i) out[i] = bound ? node_rng_below(&r, bound) : node_rng_next(&r);
    return state;
}

/* std::mt19937_64 (ISO C++ [rand.predef]: the 64-bit Mersenne Twister of
 * Matsumoto & Nishimura with the standard's parameters and seeding). */
typedef struct {
    uint64_t mt[312];
    int idx;
} mt64;

static void mt64_seed(mt64* g, uint64_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 312; ++i)
        g->mt[i] = 6364136223846793005ull * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->idx = 312;
}

static uint64_t mt64_next(mt64* g) {
    if (g->idx >= 312) {
        const uint64_t upper = 0xFFFFFFFF80000000ull, lower = 0x7FFFFFFFull;
        for (int i = 0; i < 312; ++i) {
            const uint64_t x = (g->mt[i] & upper) | (g->mt[(i + 1) % 312] & lower);
            uint64_t xa = x >> 1;
            if (x & 1ull) xa ^= 0xB5026F5AA96619E9ull;
            g->mt[i] = g->mt[(i + 156) % 312] ^ xa;
        }
        g->idx = 0;
    }
    uint64_t y = g->mt[g->idx++];
    y ^= (y >> 29) & 0x5555555555555555ull;
    y ^= (y << 17) & 0x71D67FFFEDA60000ull;
    y ^= (y << 37) & 0xFFF7EEE000000000ull;
    y ^= y >> 43;
    return y;
}

void orc_mt64_first(uint64_t seed, size_t count, uint64_t* out) {
    mt64* g = (mt64*)malloc(sizeof(mt64));
    mt64_seed(g, seed);
    for (size_t i = 0; i < count; ++i) out[i] = mt64_next(g);
    free(g);
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ========================================================================== */
/* k-NN graph                                                                 */
/* ========================================================================== */

/* src/knn_graph.cpp:37-81 (insert_unlocked): reject self / invalid, reject
 * `!(d < D[k-1])` when full, reject duplicates, then shift-insert after every
 * entry with an equal distance. Reports the evicted id. */
int orc_try_insert(uint32_t* ids, float* dist, uint8_t* flags, uint32_t* count, uint32_t k,
                   uint32_t node, uint32_t cand, float d, int is_new, uint32_t* evicted) {
    if (evicted) *evicted = ORC_INVALID;
    if (cand == node || cand == ORC_INVALID) return 0;
    const size_t base = (size_t)node * k;
    uint32_t* I = ids + base;
    float* D = dist + base;
    uint8_t* F = flags + base;
    const uint32_t cnt = count[node];
    if (cnt == k && !(d < D[k - 1])) return 0;
    for (uint32_t j = 0; j < cnt; ++j)
        if (I[j] == cand) return 0;
    uint32_t start;
    if (cnt < k) {
        start = cnt;
        count[node] = cnt + 1;
    } else {
        if (evicted) *evicted = I[k - 1];
        start = k - 1;
    }
    uint32_t pos = start;
    while (pos > 0 && D[pos - 1] > d) {
        D[pos] = D[pos - 1];
        I[pos] = I[pos - 1];
        F[pos] = F[pos - 1];
        --pos;
    }
    D[pos] = d;
    I[pos] = cand;
    F[pos] = is_new ? 1u : 0u;
    return 1;
}

static void graph_reset(uint32_t* ids, float* dist, uint8_t* flags, uint32_t* count, uint32_t n,
                        uint32_t k) {
    /* knn_graph.cpp:11-20: unfilled slots are kInvalidIdx / +inf. */
    const size_t total = (size_t)n * k;
    for (size_t i = 0; i < total; ++i) {
        ids[i] = ORC_INVALID;
        dist[i] = INFINITY;
        flags[i] = 0;
    }
    for (uint32_t i = 0; i < n; ++i) count[i] = 0;
}

/* src/nndescent.cpp:211-240 (run_init_random), serial schedule. */
void orc_init_random(const float* data, uint32_t n, uint32_t stride, uint32_t k, uint64_t seed,
                     uint32_t* ids, float* dist, uint8_t* flags, uint32_t* count) {
    if (n < 2) return;
    for (uint32_t node = 0; node < n; ++node) {
        node_rng rng;
        node_rng_init(&rng, seed, node, 0);
        const uint32_t needed = k - count[node];
        const uint32_t attempts = needed + needed / 2 + 4;
        const float* v = data + (size_t)node * stride;
        for (uint32_t t = 0; t < attempts && count[node] < k; ++t) {
            const uint32_t cand = node_rng_below(&rng, n);
            if (cand == node) continue;
            const float d = sqrtf(orc_l2sqr(v, data + (size_t)cand * stride, stride));
            orc_try_insert(ids, dist, flags, count, k, node, cand, d, 1, NULL);
        }
    }
}

/* src/nndescent.cpp:271-359 (ReverseCache), serial: bounded per-node lists of
 * `sample` ids with a fill counter; a drain takes min(count, sample) entries
 * and resets the counter. */
typedef struct {
    uint32_t n, sample;
    uint32_t *old_ids, *new_ids, *old_cnt, *new_cnt;
} rev_cache;

static void rev_push(uint32_t* ids, uint32_t* cnt, uint32_t sample, uint32_t node, uint32_t src) {
    if (sample == 0) return;
    if (cnt[node] >= sample) return;
    const uint32_t pos = cnt[node]++;
    ids[(size_t)node * sample + pos] = src;
}

static uint32_t rev_drain(uint32_t* ids, uint32_t* cnt, uint32_t sample, uint32_t node,
                          uint32_t* out) {
    if (sample == 0) return 0;
    uint32_t c = cnt[node];
    cnt[node] = 0;
    if (c > sample) c = sample;
    uint32_t m = 0;
    for (uint32_t j = 0; j < c; ++j) {
        const uint32_t v = ids[(size_t)node * sample + j];
        if (v != ORC_INVALID) out[m++] = v;
    }
    return m;
}

/* src/nndescent.cpp:31-50 (try_pair) for L2: compare in squared space against
 * to_rank(D[k-1]) = D*D, finalize with sqrt, insert both directions. */
static void try_pair(const float* data, uint32_t stride, uint32_t* ids, float* dist,
                     uint8_t* flags, uint32_t* count, uint32_t k, uint32_t u1, uint32_t u2,
                     uint64_t* evals, uint64_t* updates) {
    if (u1 == u2) return;
    const float r = orc_l2sqr(data + (size_t)u1 * stride, data + (size_t)u2 * stride, stride);
    ++*evals;
    const float t1 = dist[(size_t)u1 * k + k - 1];
    const float t2 = dist[(size_t)u2 * k + k - 1];
    const int want1 = r < t1 * t1;
    const int want2 = r < t2 * t2;
    if (!want1 && !want2) return;
    const float d = sqrtf(r);
    if (want1 && orc_try_insert(ids, dist, flags, count, k, u1, u2, d, 1, NULL)) ++*updates;
    if (want2 && orc_try_insert(ids, dist, flags, count, k, u2, u1, d, 1, NULL)) ++*updates;
}

/* src/nndescent.cpp:148-209 (run_descent) + :52-146 (local_join), Closest
 * sampling, no join_prefix, no block filter, serial node order. */
void orc_nndescent(const float* data, uint32_t n, uint32_t stride, uint32_t k, uint32_t sample,
                   uint32_t iters, double delta, uint32_t* ids, float* dist, uint8_t* flags,
                   uint32_t* count, uint64_t* stats) {
    uint64_t rounds = 0, total_updates = 0, total_evals = 0;
    if (n == 0) {
        if (stats) stats[0] = stats[1] = stats[2] = 0;
        return;
    }
    const uint32_t max_new = sample, max_old = 3u * sample;
    rev_cache rc;
    rc.n = n;
    rc.sample = sample;
    const size_t cells = (size_t)n * (sample ? sample : 1);
    rc.old_ids = (uint32_t*)malloc(cells * 4);
    rc.new_ids = (uint32_t*)malloc(cells * 4);
    rc.old_cnt = (uint32_t*)calloc(n, 4);
    rc.new_cnt = (uint32_t*)calloc(n, 4);
    for (size_t i = 0; i < cells; ++i) rc.old_ids[i] = rc.new_ids[i] = ORC_INVALID;
    uint32_t* nnew = (uint32_t*)malloc((max_new + sample + 1) * 4);
    uint32_t* nold = (uint32_t*)malloc((max_old + sample + 1) * 4);
    uint32_t* drained = (uint32_t*)malloc((sample + 1) * 4);
    const double stop_threshold = delta * (double)k * (double)n;

    for (uint32_t round = 0; round < iters; ++round) {
        uint64_t round_updates = 0, round_evals = 0;
        for (uint32_t node = 0; node < n; ++node) {
            uint32_t cn = 0, co = 0;
            const uint32_t cnt = count[node];
            uint32_t* I = ids + (size_t)node * k;
            uint8_t* F = flags + (size_t)node * k;
            for (uint32_t j = 0; j < cnt; ++j) {
                if (F[j]) {
                    if (cn < max_new) {
                        nnew[cn++] = I[j];
                        F[j] = 0;
                    }
                } else if (co < max_old) {
                    nold[co++] = I[j];
                }
                if (cn >= max_new && co >= max_old) break;
            }
            for (uint32_t j = 0; j < co; ++j)
                if (nold[j] < n) rev_push(rc.old_ids, rc.old_cnt, sample, nold[j], node);
            for (uint32_t j = 0; j < cn; ++j)
                if (nnew[j] < n) rev_push(rc.new_ids, rc.new_cnt, sample, nnew[j], node);
            uint32_t m = rev_drain(rc.old_ids, rc.old_cnt, sample, node, drained);
            for (uint32_t j = 0; j < m; ++j) nold[co++] = drained[j];
            m = rev_drain(rc.new_ids, rc.new_cnt, sample, node, drained);
            for (uint32_t j = 0; j < m; ++j) nnew[cn++] = drained[j];
            if (cn == 0) continue;
            for (uint32_t a = 0; a < cn; ++a) {
                const uint32_t u1 = nnew[a];
                if (u1 >= n) continue;
                for (uint32_t b = a + 1; b < cn; ++b) {
                    const uint32_t u2 = nnew[b];
                    if (u2 >= n) continue;
                    try_pair(data, stride, ids, dist, flags, count, k, u1, u2, &round_evals,
                             &round_updates);
                }
                for (uint32_t b = 0; b < co; ++b) {
                    const uint32_t u2 = nold[b];
                    if (u2 >= n) continue;
                    try_pair(data, stride, ids, dist, flags, count, k, u1, u2, &round_evals,
                             &round_updates);
                }
            }
        }
        total_updates += round_updates;
        total_evals += round_evals;
        rounds = round + 1;
        if ((double)round_updates <= stop_threshold) break;
    }
    free(rc.old_ids);
    free(rc.new_ids);
    free(rc.old_cnt);
    free(rc.new_cnt);
    free(nnew);
    free(nold);
    free(drained);
    if (stats) {
        stats[0] = rounds;
        stats[1] = total_updates;
        stats[2] = total_evals;
    }
}

/* src/nndescent.cpp:242-263 (run_brute_force), j ascending: ties keep the
 * lower id because a full list rejects an equal distance. */
void orc_brute_force(const float* data, uint32_t n, uint32_t stride, uint32_t k, uint32_t* ids,
                     float* dist, uint8_t* flags, uint32_t* count) {
    for (uint32_t node = 0; node < n; ++node) {
        const float* v = data + (size_t)node * stride;
        for (uint32_t j = 0; j < n; ++j) {
            if (j == node) continue;
            const float r = orc_l2sqr(v, data + (size_t)j * stride, stride);
            const float t = dist[(size_t)node * k + k - 1];
            if (count[node] == k && !(r < t * t)) continue;
            orc_try_insert(ids, dist, flags, count, k, node, j, sqrtf(r), 0, NULL);
        }
    }
}

/* src/knn_graph.cpp:237-273 (validate): sorted, no self loop, no duplicate. */
int orc_validate(const uint32_t* ids, const float* dist, const uint32_t* count, uint32_t n,
                 uint32_t k) {
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t c = count[i];
        if (c > k) return 0;
        const uint32_t* I = ids + (size_t)i * k;
        const float* D = dist + (size_t)i * k;
        for (uint32_t j = 0; j < c; ++j) {
            if (I[j] == i || I[j] == ORC_INVALID) return 0;
            for (uint32_t t = 0; t < j; ++t)
                if (I[t] == I[j]) return 0;
            if (j > 0 && D[j] < D[j - 1]) return 0;
        }
    }
    return 1;
}

/* ========================================================================== */
/* extraction                                                                 */
/* ========================================================================== */

/* include/coresg/knn_graph.hpp:157-163 (core_dist): 1-based mpts, clamped to
 * the filled length, +inf for an empty list. */
static float core_dist(const float* dist, const uint32_t* count, uint32_t k, uint32_t node,
                       uint32_t mpts) {
    const uint32_t c = count[node];
    if (c == 0) return INFINITY;
    uint32_t pos = (mpts == 0) ? 0 : mpts - 1;
    if (pos >= c) pos = c - 1;
    return dist[(size_t)node * k + pos];
}

void orc_core_distances(const float* dist, const uint32_t* count, uint32_t n, uint32_t k,
                        uint32_t mpts, float* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = core_dist(dist, count, k, i, mpts);
}

/* src/knn_graph.cpp:132-169 (collect_undirected) serial, and the Edge
 * constructor's u<v canonicalisation (types.hpp:24-31). Node p owns (p,q) if
 * q>p, or if q<p and p is not in q's list (membership scan short-circuited by
 * the sorted-list bound). Weight = p's stored distance. */
size_t orc_undirected_edges(const uint32_t* ids, const float* dist, const uint32_t* count,
                            uint32_t n, uint32_t k, uint32_t* u, uint32_t* v, float* w) {
    size_t m = 0;
    for (uint32_t p = 0; p < n; ++p) {
        const uint32_t* I = ids + (size_t)p * k;
        const float* D = dist + (size_t)p * k;
        const uint32_t c = count[p];
        for (uint32_t j = 0; j < c; ++j) {
            const uint32_t q = I[j];
            if (q == ORC_INVALID || q == p) continue;
            if (q < p) {
                const uint32_t cq = count[q];
                if (cq > 0 && !(D[j] > dist[(size_t)q * k + cq - 1])) {
                    const uint32_t* J = ids + (size_t)q * k;
                    int mutual = 0;
                    for (uint32_t t = 0; t < cq; ++t) {
                        if (J[t] == p) {
                            mutual = 1;
                            break;
                        }
                    }
                    if (mutual) continue;
                }
            }
            if (u) {
                u[m] = p < q ? p : q;
                v[m] = p < q ? q : p;
                w[m] = D[j];
            }
            ++m;
        }
    }
    return m;
}

/* include/coresg/mst.hpp:30-75 (UnionFind: union by rank, path halving) and
 * src/mst.cpp:48-59 (labels dense by first appearance). */
typedef struct {
    uint32_t* parent;
    uint8_t* rank;
    uint32_t n, components;
} union_find;

static void uf_init(union_find* uf, uint32_t n) {
    uf->parent = (uint32_t*)malloc((size_t)(n ? n : 1) * 4);
    uf->rank = (uint8_t*)calloc(n ? n : 1, 1);
    for (uint32_t i = 0; i < n; ++i) uf->parent[i] = i;
    uf->n = n;
    uf->components = n;
}
static void uf_free(union_find* uf) {
    free(uf->parent);
    free(uf->rank);
}
static uint32_t uf_find(union_find* uf, uint32_t x) {
    while (uf->parent[x] != x) {
        uf->parent[x] = uf->parent[uf->parent[x]];
        x = uf->parent[x];
    }
    return x;
}
static int uf_unite(union_find* uf, uint32_t a, uint32_t b) {
    a = uf_find(uf, a);
    b = uf_find(uf, b);
    if (a == b) return 0;
    if (uf->rank[a] < uf->rank[b]) {
        const uint32_t t = a;
        a = b;
        b = t;
    }
    uf->parent[b] = a;
    if (uf->rank[a] == uf->rank[b]) ++uf->rank[a];
    --uf->components;
    return 1;
}
static void uf_labels(union_find* uf, uint32_t* out) {
    uint32_t* remap = (uint32_t*)malloc((size_t)(uf->n ? uf->n : 1) * 4);
    for (uint32_t i = 0; i < uf->n; ++i) remap[i] = ORC_INVALID;
    uint32_t next = 0;
    for (uint32_t i = 0; i < uf->n; ++i) {
        const uint32_t r = uf_find(uf, i);
        if (remap[r] == ORC_INVALID) remap[r] = next++;
        out[i] = remap[r];
    }
    free(remap);
}

typedef struct {
    uint32_t u, v;
    float w;
} orc_edge;

/* include/coresg/types.hpp:38-42 (edge_less): (w, u, v). */
static int edge_cmp(const void* pa, const void* pb) {
    const orc_edge* a = (const orc_edge*)pa;
    const orc_edge* b = (const orc_edge*)pb;
    if (a->w != b->w) return a->w < b->w ? -1 : 1;
    if (a->u != b->u) return a->u < b->u ? -1 : 1;
    if (a->v != b->v) return a->v < b->v ? -1 : 1;
    return 0;
}

/* src/mst.cpp:61-77 (mst_kruskal) + finish() at :36-44: sort by edge_less,
 * union by rank, stop at n-1 edges, labels, component count, double sum of
 * weights in acceptance order (types.hpp:58-62). */
size_t orc_mst_kruskal(uint32_t n, const uint32_t* eu, const uint32_t* ev, const float* ew,
                       size_t m, uint32_t* out_u, uint32_t* out_v, float* out_w,
                       uint32_t* out_labels, uint32_t* out_ncomp, double* out_weight) {
    orc_edge* e = (orc_edge*)malloc((m ? m : 1) * sizeof(orc_edge));
    for (size_t i = 0; i < m; ++i) {
        e[i].u = eu[i] < ev[i] ? eu[i] : ev[i];
        e[i].v = eu[i] < ev[i] ? ev[i] : eu[i];
        e[i].w = ew[i];
    }
    qsort(e, m, sizeof(orc_edge), edge_cmp);
    union_find uf;
    uf_init(&uf, n);
    size_t t = 0;
    double weight = 0.0;
    for (size_t i = 0; i < m; ++i) {
        if (e[i].u >= n || e[i].v >= n) continue;
        if (uf_unite(&uf, e[i].u, e[i].v)) {
            if (out_u) {
                out_u[t] = e[i].u;
                out_v[t] = e[i].v;
                out_w[t] = e[i].w;
            }
            weight += (double)e[i].w;
            ++t;
            if (t + 1 == (size_t)n) break;
        }
    }
    if (out_labels) uf_labels(&uf, out_labels);
    if (out_ncomp) *out_ncomp = uf.components;
    if (out_weight) *out_weight = weight;
    uf_free(&uf);
    free(e);
    return t;
}

/* ========================================================================== */
/* hubness / selection / components                                          */
/* ========================================================================== */

/* src/hubness.cpp:86-100: rdeg[q] = #{p : q in N(p)}. */
void orc_reverse_degrees(const uint32_t* ids, const uint32_t* count, uint32_t n, uint32_t k,
                         int32_t* rdeg) {
    for (uint32_t i = 0; i < n; ++i) rdeg[i] = 0;
    for (uint32_t p = 0; p < n; ++p) {
        const uint32_t c = count[p];
        for (uint32_t j = 0; j < c; ++j) {
            const uint32_t q = ids[(size_t)p * k + j];
            if (q < n) ++rdeg[q];
        }
    }
}

/* src/hubness.cpp:307-352 (compute_score, ArithmeticMean): int64 neighbour sum,
 * degenerate rules, then rd / (sum / deg) in double. */
double orc_score(const uint32_t* ids, const uint32_t* count, uint32_t n, uint32_t k,
                 uint32_t k_for_score, const int32_t* rdeg, uint32_t v) {
    const uint32_t deg = count[v];
    const int32_t rd = rdeg[v];
    int64_t sum = 0;
    for (uint32_t j = 0; j < deg; ++j) {
        const uint32_t q = ids[(size_t)v * k + j];
        if (q >= n) continue;
        sum += rdeg[q];
    }
    if (rd > 0 && sum <= 0) return (double)k_for_score;
    if (rd <= 0 && sum > 0) return 0.0;
    if (rd <= 0 && sum <= 0) return 1.0;
    if (deg == 0) return 1.0;
    const double denom = (double)sum / (double)deg;
    if (!(denom > 0.0)) return (double)k_for_score;
    return (double)rd / denom;
}

/* src/hubness.cpp:16-40 (antihub_budget). */
size_t orc_antihub_budget(double multiplier, uint32_t dim, size_t num_edges, uint32_t n,
                          double max_pair_ops, size_t max_selected, size_t min_selected,
                          size_t* uncapped, int* capped) {
    if (n == 0) return 0;
    const double logs = (num_edges > 1) ? log((double)num_edges) : 0.0;
    double raw = multiplier * (double)dim * logs;
    if (!(raw > 0.0)) raw = 0.0;
    size_t count = (size_t)llround(raw);
    if (uncapped) *uncapped = count;
    const size_t requested = count;
    if (max_pair_ops > 0.0 && dim > 0) {
        const double m_max = sqrt(2.0 * max_pair_ops / (double)dim);
        if (m_max >= 1.0 && (double)count > m_max) count = (size_t)m_max;
    }
    if (max_selected && count > max_selected) count = max_selected;
    if (count < min_selected) count = min_selected;
    if (count > n) count = n;
    if (capped) *capped = count < requested;
    return count;
}

/* Ordering of src/hubness.cpp:438-445: score, then reverse degree, then id. */
static const double* g_sel_score;
static const int32_t* g_sel_rdeg;
static int sel_less(uint32_t a, uint32_t b) {
    if (g_sel_score[a] != g_sel_score[b]) return g_sel_score[a] < g_sel_score[b];
    if (g_sel_rdeg[a] != g_sel_rdeg[b]) return g_sel_rdeg[a] < g_sel_rdeg[b];
    return a < b;
}
static int sel_cmp(const void* pa, const void* pb) {
    const uint32_t a = *(const uint32_t*)pa, b = *(const uint32_t*)pb;
    if (a == b) return 0;
    return sel_less(a, b) ? -1 : 1;
}

/* src/hubness.cpp:424-478 (select): the `budget` lowest by the order above
 * (nth_element + sort is the sorted prefix of the total order), then the
 * per-component quota of the `per` lowest of every component, appended and
 * re-sorted when anything was added. Returns the selected count. */
size_t orc_select(uint32_t n, const double* score, const int32_t* rdeg, size_t budget,
                  const uint32_t* comp, uint32_t ncomp, size_t per, uint32_t* out) {
    g_sel_score = score;
    g_sel_rdeg = rdeg;
    uint32_t* order = (uint32_t*)malloc((size_t)(n ? n : 1) * 4);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    qsort(order, n, 4, sel_cmp);
    if (budget > n) budget = n;
    size_t m = 0;
    for (size_t i = 0; i < budget; ++i) out[m++] = order[i];
    free(order);
    if (per && comp && ncomp > 0) {
        uint8_t* chosen = (uint8_t*)calloc(n ? n : 1, 1);
        for (size_t i = 0; i < m; ++i) chosen[out[i]] = 1;
        uint32_t* buckets = (uint32_t*)malloc((size_t)ncomp * per * 4);
        uint32_t* bsize = (uint32_t*)calloc(ncomp, 4);
        for (uint32_t v = 0; v < n; ++v) {
            const uint32_t c = comp[v];
            if (c >= ncomp) continue;
            uint32_t* b = buckets + (size_t)c * per;
            if (bsize[c] < per) {
                b[bsize[c]++] = v;
            } else {
                size_t worst = 0;
                for (size_t t = 1; t < bsize[c]; ++t)
                    if (sel_less(b[worst], b[t])) worst = t;
                if (sel_less(v, b[worst])) b[worst] = v;
            }
        }
        size_t added = 0;
        for (uint32_t c = 0; c < ncomp; ++c) {
            for (uint32_t t = 0; t < bsize[c]; ++t) {
                const uint32_t v = buckets[(size_t)c * per + t];
                if (!chosen[v]) {
                    chosen[v] = 1;
                    out[m++] = v;
                    ++added;
                }
            }
        }
        if (added) qsort(out, m, 4, sel_cmp);
        free(chosen);
        free(buckets);
        free(bsize);
    }
    return m;
}

/* src/connectivity.cpp:99-139 (rebuild) + :64-74 (finalize): components of
 * the adjacency read as undirected; labels dense by first appearance. The
 * partition, and therefore the labelling, does not depend on union order, so
 * a serial union-find gives the reference's labels. */
uint32_t orc_components(const uint32_t* ids, const uint32_t* count, uint32_t n, uint32_t k,
                        uint32_t* labels) {
    union_find uf;
    uf_init(&uf, n);
    for (uint32_t p = 0; p < n; ++p) {
        const uint32_t c = count[p];
        for (uint32_t j = 0; j < c; ++j) {
            const uint32_t q = ids[(size_t)p * k + j];
            if (q < n && q != p) uf_unite(&uf, p, q);
        }
    }
    uf_labels(&uf, labels);
    const uint32_t nc = uf.components;
    uf_free(&uf);
    return nc;
}

/* ========================================================================== */
/* bridges                                                                    */
/* ========================================================================== */

/* src/cpu_backend.cpp:61-90 (CpuBackend::mrd_block, L2): rows [r0,r1) of the
 * m x m matrix, 0 on the diagonal, max(d, max(c_i, c_j)) elsewhere. */
void orc_mrd_block(const float* pts, size_t m, uint32_t stride, const float* core, size_t r0,
                   size_t r1, float* out) {
    for (size_t i = 0; i < r1 - r0; ++i) {
        const float* a = pts + (r0 + i) * stride;
        const float ci = core[r0 + i];
        float* row = out + i * m;
        for (size_t j = 0; j < m; ++j) {
            if (r0 + i == j) {
                row[j] = 0.0f;
                continue;
            }
            const float d = sqrtf(orc_l2sqr(a, pts + j * stride, stride));
            const float cm = ci < core[j] ? core[j] : ci;
            row[j] = d < cm ? cm : d;
        }
    }
}

/* src/mst.cpp:172-220 (mst_prim_virtual) over a dense matrix (mst_dense,
 * :158-162): grow from node 0, strict improvement, first index wins ties.
 * Edges are canonicalised u<v. */
size_t orc_prim_dense(const float* matrix, size_t m, uint32_t* out_u, uint32_t* out_v,
                      float* out_w) {
    if (m < 2) return 0;
    float* best = (float*)malloc(m * 4);
    uint32_t* from = (uint32_t*)malloc(m * 4);
    uint8_t* in_tree = (uint8_t*)calloc(m, 1);
    for (size_t i = 0; i < m; ++i) {
        best[i] = INFINITY;
        from[i] = ORC_INVALID;
    }
    size_t t = 0;
    uint32_t current = 0;
    in_tree[0] = 1;
    best[0] = 0.0f;
    for (size_t step = 1; step < m; ++step) {
        const float* row = matrix + (size_t)current * m;
        for (size_t j = 0; j < m; ++j) {
            if (in_tree[j]) continue;
            if (row[j] < best[j]) {
                best[j] = row[j];
                from[j] = current;
            }
        }
        uint32_t pick = ORC_INVALID;
        float pick_w = INFINITY;
        for (size_t j = 0; j < m; ++j) {
            if (in_tree[j]) continue;
            if (best[j] < pick_w || (best[j] == pick_w && pick == ORC_INVALID)) {
                pick_w = best[j];
                pick = (uint32_t)j;
            }
        }
        if (pick == ORC_INVALID) break;
        in_tree[pick] = 1;
        if (from[pick] != ORC_INVALID && pick_w < INFINITY) {
            const uint32_t a = from[pick], b = pick;
            out_u[t] = a < b ? a : b;
            out_v[t] = a < b ? b : a;
            out_w[t] = pick_w;
            ++t;
        }
        current = pick;
    }
    free(best);
    free(from);
    free(in_tree);
    return t;
}

/* ========================================================================== */
/* whole CORE-SG                                                              */
/* ========================================================================== */

struct orc_sg {
    uint32_t n, k, dim, stride;
    float* pts;
    uint32_t* ids;
    float* dist;
    uint8_t* flags;
    uint32_t* count;
    orc_edge* bridges;
    size_t nbridges, repair_count;
    uint32_t* selected;
    size_t nselected;
    double* score;
    int32_t* rdeg;
    uint32_t* conn;
    uint32_t ncomp;
    size_t knn_edges;
    orc_params p;
    double stats[7];
};

void orc_default_params(orc_params* p) {
    /* Struct defaults of CoreSGParams/HubnessParams/NNDescentParams
     * (coresg.hpp:63-106, hubness.hpp:72-119, nndescent.hpp:45-71) and the
     * CLI's sample/iters defaults (coresg_cli.cpp:154-158). */
    p->k_max = 32;
    p->sample = 16;
    p->iters = 10;
    p->delta = 0.001;
    p->nnd_seed = 42;
    p->brute_force_limit = 0;
    p->antihub_multiplier = 1.0;
    p->max_pair_ops = 2e9;
    p->min_selected = 2;
    p->max_selected = 0;
    p->min_per_component = 2;
    p->ensure_connected = 1;
    p->reps_per_component = 4;
    p->max_reps = 4096;
}

static orc_sg* sg_alloc(uint32_t n, uint32_t k, uint32_t dim, const orc_params* p) {
    orc_sg* sg = (orc_sg*)calloc(1, sizeof(orc_sg));
    sg->n = n;
    sg->k = k;
    sg->dim = dim;
    sg->stride = (dim + 7u) & ~7u;
    sg->p = *p;
    sg->p.k_max = k;
    const size_t nk = (size_t)(n ? n : 1) * k;
    sg->pts = (float*)calloc((size_t)(n ? n : 1) * sg->stride, 4);
    sg->ids = (uint32_t*)malloc(nk * 4);
    sg->dist = (float*)malloc(nk * 4);
    sg->flags = (uint8_t*)malloc(nk);
    sg->count = (uint32_t*)malloc((size_t)(n ? n : 1) * 4);
    graph_reset(sg->ids, sg->dist, sg->flags, sg->count, n, k);
    return sg;
}

void orc_sg_free(orc_sg* sg) {
    if (!sg) return;
    free(sg->pts);
    free(sg->ids);
    free(sg->dist);
    free(sg->flags);
    free(sg->count);
    free(sg->bridges);
    free(sg->selected);
    free(sg->score);
    free(sg->rdeg);
    free(sg->conn);
    free(sg);
}

/* src/coresg.cpp:97-141 (mst_over_subset): gather, core at k_max, dense MRD
 * matrix, Prim, bridges re-stored with the base distance in global ids. */
static size_t mst_over_subset(const orc_sg* sg, const uint32_t* sel, size_t m, orc_edge* out) {
    if (m < 2) return 0;
    const uint32_t stride = sg->stride;
    float* pts = (float*)malloc(m * stride * 4);
    float* core = (float*)malloc(m * 4);
    for (size_t i = 0; i < m; ++i) {
        memcpy(pts + i * stride, sg->pts + (size_t)sel[i] * stride, stride * 4);
        core[i] = core_dist(sg->dist, sg->count, sg->k, sel[i], sg->k);
    }
    float* mat = (float*)malloc(m * m * 4);
    orc_mrd_block(pts, m, stride, core, 0, m, mat);
    uint32_t* lu = (uint32_t*)malloc(m * 4);
    uint32_t* lv = (uint32_t*)malloc(m * 4);
    float* lw = (float*)malloc(m * 4);
    const size_t t = orc_prim_dense(mat, m, lu, lv, lw);
    for (size_t i = 0; i < t; ++i) {
        const float base = sqrtf(orc_l2sqr(pts + (size_t)lu[i] * stride,
                                           pts + (size_t)lv[i] * stride, stride));
        const uint32_t a = sel[lu[i]], b = sel[lv[i]];
        out[i].u = a < b ? a : b;
        out[i].v = a < b ? b : a;
        out[i].w = base;
    }
    free(pts);
    free(core);
    free(mat);
    free(lu);
    free(lv);
    free(lw);
    return t;
}

/* src/connectivity.cpp:352-374 (labels_with): contract k-NN components and
 * union the extra edges; labels dense over component order. */
static uint32_t labels_with(const orc_sg* sg, const orc_edge* extra, size_t nextra,
                            uint32_t* out) {
    if (nextra == 0) {
        memcpy(out, sg->conn, (size_t)sg->n * 4);
        return sg->ncomp;
    }
    union_find uf;
    uf_init(&uf, sg->ncomp);
    for (size_t i = 0; i < nextra; ++i) {
        if (extra[i].u >= sg->n || extra[i].v >= sg->n) continue;
        const uint32_t a = sg->conn[extra[i].u], b = sg->conn[extra[i].v];
        if (a < sg->ncomp && b < sg->ncomp) uf_unite(&uf, a, b);
    }
    uint32_t* merged = (uint32_t*)malloc((size_t)(sg->ncomp ? sg->ncomp : 1) * 4);
    uf_labels(&uf, merged);
    for (uint32_t v = 0; v < sg->n; ++v) {
        const uint32_t lab = sg->conn[v];
        out[v] = lab < sg->ncomp ? merged[lab] : lab;
    }
    const uint32_t nc = uf.components;
    free(merged);
    uf_free(&uf);
    return nc;
}

/* std::partial_sort(first, first + mid, last, comp) as GCC 11's libstdc++
 * implements it (bits/stl_algo.h: __partial_sort = __heap_select then
 * __sort_heap; bits/stl_heap.h: __make_heap, __adjust_heap, __push_heap,
 * __pop_heap), restated over an index array ordered by a double key. The
 * reference calls it with `score[a] < score[b]` (src/coresg.cpp:222-225); with
 * tied scores the surviving set and its order are those of this heap
 * procedure, so it is restated step for step rather than by any sort. */
static int ps_less(const double* key, uint32_t a, uint32_t b) { return key[a] < key[b]; }

static void ps_push_heap(uint32_t* f, ptrdiff_t hole, ptrdiff_t top, uint32_t value,
                         const double* key) {
    ptrdiff_t parent = (hole - 1) / 2;
    while (hole > top && ps_less(key, f[parent], value)) {
        f[hole] = f[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    f[hole] = value;
}

static void ps_adjust_heap(uint32_t* f, ptrdiff_t hole, ptrdiff_t len, uint32_t value,
                           const double* key) {
    const ptrdiff_t top = hole;
    ptrdiff_t child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (ps_less(key, f[child], f[child - 1])) child--;
        f[hole] = f[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        f[hole] = f[child - 1];
        hole = child - 1;
    }
    ps_push_heap(f, hole, top, value, key);
}

/* __pop_heap(first, last, result): the top goes to *result, *result's old
 * value sinks from the root of [first, last). */
static void ps_pop_heap(uint32_t* f, ptrdiff_t len, uint32_t* result, const double* key) {
    const uint32_t value = *result;
    *result = f[0];
    ps_adjust_heap(f, 0, len, value, key);
}

static void partial_sort_by_key(uint32_t* f, size_t mid, size_t last, const double* key) {
    const ptrdiff_t len = (ptrdiff_t)mid;
    if (len >= 2) { /* __make_heap */
        for (ptrdiff_t parent = (len - 2) / 2;; --parent) {
            ps_adjust_heap(f, parent, len, f[parent], key);
            if (parent == 0) break;
        }
    }
    for (size_t i = mid; i < last; ++i) /* __heap_select */
        if (ps_less(key, f[i], f[0])) ps_pop_heap(f, len, &f[i], key);
    for (ptrdiff_t l = len; l > 1;) { /* __sort_heap */
        --l;
        ps_pop_heap(f, l, &f[l], key);
    }
}

/* src/coresg.cpp:143-169 (recompute_missing_edges, AntiHub) and :171-258
 * (repair_connectivity), including the representative cut above max_reps by
 * the reference's partial_sort (restated above). */
static void recompute_bridges(orc_sg* sg) {
    free(sg->bridges);
    const size_t cap = sg->nselected + 1;
    sg->bridges = (orc_edge*)malloc(cap * sizeof(orc_edge));
    sg->nbridges = mst_over_subset(sg, sg->selected, sg->nselected, sg->bridges);
    sg->repair_count = 0;

    const uint32_t n = sg->n;
    if (n == 0) return;
    uint32_t* comp = (uint32_t*)malloc((size_t)n * 4);
    const uint32_t ncomp = labels_with(sg, sg->bridges, sg->nbridges, comp);
    if (ncomp <= 1 || !sg->p.ensure_connected) {
        free(comp);
        return;
    }
    size_t per = sg->p.reps_per_component ? sg->p.reps_per_component : 1;
    if (sg->p.max_reps && (size_t)ncomp * per > sg->p.max_reps) {
        per = sg->p.max_reps / (size_t)ncomp;
        if (per < 1) per = 1;
    }
    uint32_t* buckets = (uint32_t*)malloc((size_t)ncomp * per * 4);
    uint32_t* bsize = (uint32_t*)calloc(ncomp, 4);
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t c = comp[v];
        uint32_t* b = buckets + (size_t)c * per;
        if (bsize[c] < per) {
            b[bsize[c]++] = v;
        } else {
            size_t worst = 0;
            for (size_t t = 1; t < bsize[c]; ++t)
                if (sg->score[b[t]] > sg->score[b[worst]]) worst = t;
            if (sg->score[v] < sg->score[b[worst]]) b[worst] = v;
        }
    }
    size_t nreps = 0;
    uint32_t* reps = (uint32_t*)malloc((size_t)ncomp * per * 4 + 4);
    for (uint32_t c = 0; c < ncomp; ++c)
        for (uint32_t t = 0; t < bsize[c]; ++t) reps[nreps++] = buckets[(size_t)c * per + t];
    if (sg->p.max_reps && nreps > sg->p.max_reps) {
        partial_sort_by_key(reps, sg->p.max_reps, nreps, sg->score);
        nreps = sg->p.max_reps;
    }
    if (nreps >= 2) {
        orc_edge* rep_mst = (orc_edge*)malloc(nreps * sizeof(orc_edge));
        const size_t t = mst_over_subset(sg, reps, nreps, rep_mst);
        union_find uf;
        uf_init(&uf, ncomp);
        sg->bridges = (orc_edge*)realloc(sg->bridges, (sg->nbridges + t + 1) * sizeof(orc_edge));
        for (size_t i = 0; i < t; ++i) {
            if (uf_unite(&uf, comp[rep_mst[i].u], comp[rep_mst[i].v])) {
                sg->bridges[sg->nbridges++] = rep_mst[i];
                ++sg->repair_count;
            }
        }
        uf_free(&uf);
        free(rep_mst);
    }
    free(reps);
    free(buckets);
    free(bsize);
    free(comp);
}

/* Derived state: components, reverse degrees, scores, selection
 * (src/coresg.cpp:278-300 in build, src/io.cpp:316-328 in load). */
static void derive_hubness(orc_sg* sg) {
    const uint32_t n = sg->n, k = sg->k;
    free(sg->conn);
    free(sg->rdeg);
    free(sg->score);
    free(sg->selected);
    sg->conn = (uint32_t*)malloc((size_t)(n ? n : 1) * 4);
    sg->ncomp = orc_components(sg->ids, sg->count, n, k, sg->conn);
    sg->rdeg = (int32_t*)malloc((size_t)(n ? n : 1) * 4);
    orc_reverse_degrees(sg->ids, sg->count, n, k, sg->rdeg);
    sg->knn_edges = orc_undirected_edges(sg->ids, sg->dist, sg->count, n, k, NULL, NULL, NULL);
    sg->score = (double*)malloc((size_t)(n ? n : 1) * 8);
    for (uint32_t v = 0; v < n; ++v)
        sg->score[v] = orc_score(sg->ids, sg->count, n, k, k, sg->rdeg, v);
    const size_t budget =
        orc_antihub_budget(sg->p.antihub_multiplier, sg->dim, sg->knn_edges, n,
                           sg->p.max_pair_ops, sg->p.max_selected, sg->p.min_selected, NULL, NULL);
    const size_t per = sg->p.min_per_component;
    sg->selected = (uint32_t*)malloc((budget + (size_t)sg->ncomp * per + 1) * 4);
    sg->nselected = orc_select(n, sg->score, sg->rdeg, budget, sg->conn, sg->ncomp, per,
                               sg->selected);
}

/* src/coresg.cpp:260-323 (CoreSG::build) with one instance
 * (incremental_knn.cpp:1012-1050 -> cpu_backend.cpp:47-59), or brute force
 * when n <= brute_force_limit (incremental_knn.cpp:997-1007). */
orc_sg* orc_sg_build(const float* rows, uint32_t n, uint32_t dim, const orc_params* p) {
    const double t0 = now_s();
    orc_sg* sg = sg_alloc(n, p->k_max, dim, p);
    for (uint32_t i = 0; i < n; ++i) memcpy(sg->pts + (size_t)i * sg->stride, rows + (size_t)i * dim, dim * 4);
    uint64_t st[3] = {0, 0, 0};
    if (p->brute_force_limit && n <= p->brute_force_limit) {
        orc_brute_force(sg->pts, n, sg->stride, sg->k, sg->ids, sg->dist, sg->flags, sg->count);
    } else {
        orc_init_random(sg->pts, n, sg->stride, sg->k, p->nnd_seed, sg->ids, sg->dist, sg->flags,
                        sg->count);
        orc_nndescent(sg->pts, n, sg->stride, sg->k, p->sample, p->iters, p->delta, sg->ids,
                      sg->dist, sg->flags, sg->count, st);
    }
    const double t1 = now_s();
    derive_hubness(sg);
    const double t2 = now_s();
    recompute_bridges(sg);
    const double t3 = now_s();
    sg->stats[0] = t1 - t0;
    sg->stats[1] = t2 - t1;
    sg->stats[2] = t3 - t2;
    sg->stats[3] = t3 - t0;
    sg->stats[4] = (double)st[0];
    sg->stats[5] = (double)st[1];
    sg->stats[6] = (double)st[2];
    return sg;
}

/* src/io.cpp:195-261 (save) / :263-330 (load): "CORESG01", u32 k_max, dim,
 * stride, metric, mode, u64 n, u64 floats + padded rows, then u64-counted
 * ids/dist/flags/count vectors, the bridge Edge vector (12-byte records) and
 * the u64 repair count. */
static int write_vec(FILE* f, const void* data, size_t count, size_t elem) {
    const uint64_t c = count;
    if (fwrite(&c, 8, 1, f) != 1) return 0;
    if (count && fwrite(data, elem, count, f) != count) return 0;
    return 1;
}

int orc_sg_save(const orc_sg* sg, const char* path) {
    FILE* f = fopen(path, "wb");
    if (!f) return 0;
    const char magic[8] = {'C', 'O', 'R', 'E', 'S', 'G', '0', '1'};
    int ok = fwrite(magic, 1, 8, f) == 8;
    const uint32_t hdr[5] = {sg->k, sg->dim, sg->stride, 0u, 0u};
    ok = ok && fwrite(hdr, 4, 5, f) == 5;
    const uint64_t n = sg->n, floats = (uint64_t)sg->n * sg->stride;
    ok = ok && fwrite(&n, 8, 1, f) == 1 && fwrite(&floats, 8, 1, f) == 1;
    ok = ok && (floats == 0 || fwrite(sg->pts, 4, floats, f) == floats);
    const size_t nk = (size_t)sg->n * sg->k;
    ok = ok && write_vec(f, sg->ids, nk, 4) && write_vec(f, sg->dist, nk, 4) &&
         write_vec(f, sg->flags, nk, 1) && write_vec(f, sg->count, sg->n, 4) &&
         write_vec(f, sg->bridges, sg->nbridges, sizeof(orc_edge));
    const uint64_t rc = sg->repair_count;
    ok = ok && fwrite(&rc, 8, 1, f) == 1;
    fclose(f);
    return ok;
}

static int read_vec(FILE* f, void* data, size_t expect, size_t elem) {
    uint64_t c = 0;
    if (fread(&c, 8, 1, f) != 1 || c != expect) return 0;
    return expect == 0 || fread(data, elem, expect, f) == expect;
}

orc_sg* orc_sg_load(const char* path, const orc_params* p) {
    FILE* f = fopen(path, "rb");
    if (!f) return NULL;
    char magic[8];
    uint32_t hdr[5];
    uint64_t n = 0, floats = 0;
    if (fread(magic, 1, 8, f) != 8 || memcmp(magic, "CORESG01", 8) != 0 ||
        fread(hdr, 4, 5, f) != 5 || fread(&n, 8, 1, f) != 1 || fread(&floats, 8, 1, f) != 1) {
        fclose(f);
        return NULL;
    }
    orc_params q = *p;
    q.k_max = hdr[0];
    orc_sg* sg = sg_alloc((uint32_t)n, hdr[0], hdr[1], &q);
    int ok = sg->stride == hdr[2] && floats == n * hdr[2];
    ok = ok && (floats == 0 || fread(sg->pts, 4, floats, f) == floats);
    const size_t nk = (size_t)n * hdr[0];
    ok = ok && read_vec(f, sg->ids, nk, 4) && read_vec(f, sg->dist, nk, 4) &&
         read_vec(f, sg->flags, nk, 1) && read_vec(f, sg->count, n, 4);
    uint64_t nb = 0;
    ok = ok && fread(&nb, 8, 1, f) == 1;
    if (ok) {
        sg->bridges = (orc_edge*)malloc((nb + 1) * sizeof(orc_edge));
        ok = nb == 0 || fread(sg->bridges, sizeof(orc_edge), nb, f) == nb;
        sg->nbridges = nb;
    }
    uint64_t rc = 0;
    ok = ok && fread(&rc, 8, 1, f) == 1;
    sg->repair_count = rc;
    fclose(f);
    if (!ok) {
        orc_sg_free(sg);
        return NULL;
    }
    derive_hubness(sg);
    return sg;
}

void orc_sg_rebuild_bridges(orc_sg* sg) {
    derive_hubness(sg);
    recompute_bridges(sg);
}

uint32_t orc_sg_n(const orc_sg* sg) { return sg->n; }
uint32_t orc_sg_k(const orc_sg* sg) { return sg->k; }
uint32_t orc_sg_dim(const orc_sg* sg) { return sg->dim; }
uint32_t orc_sg_stride(const orc_sg* sg) { return sg->stride; }
const float* orc_sg_points(const orc_sg* sg) { return sg->pts; }
void orc_sg_graph(const orc_sg* sg, const uint32_t** ids, const float** dist,
                  const uint8_t** flags, const uint32_t** count) {
    *ids = sg->ids;
    *dist = sg->dist;
    *flags = sg->flags;
    *count = sg->count;
}
size_t orc_sg_bridges(const orc_sg* sg, const uint32_t** u, const uint32_t** v, const float** w) {
    /* Edge records are {u32 u, u32 v, f32 w}; expose them as strided views. */
    *u = sg->nbridges ? &sg->bridges[0].u : NULL;
    *v = sg->nbridges ? &sg->bridges[0].v : NULL;
    *w = sg->nbridges ? &sg->bridges[0].w : NULL;
    return sg->nbridges;
}
size_t orc_sg_repair_count(const orc_sg* sg) { return sg->repair_count; }
size_t orc_sg_selected(const orc_sg* sg, const uint32_t** sel) {
    *sel = sg->selected;
    return sg->nselected;
}
const double* orc_sg_scores(const orc_sg* sg) { return sg->score; }
const int32_t* orc_sg_rdeg(const orc_sg* sg) { return sg->rdeg; }
const uint32_t* orc_sg_conn_labels(const orc_sg* sg, uint32_t* ncomp) {
    *ncomp = sg->ncomp;
    return sg->conn;
}
size_t orc_sg_knn_edges(const orc_sg* sg) { return sg->knn_edges; }
void orc_sg_stats(const orc_sg* sg, double* out) { memcpy(out, sg->stats, sizeof(sg->stats)); }

/* src/coresg.cpp:391-406 (CoreSG::edges) = knn_graph.cpp:219-235
 * (mutual_reachability_edges) + reweighted bridges. mpts 0 -> 1; mpts above
 * k_max is an error (returns (size_t)-1). */
size_t orc_sg_edges(const orc_sg* sg, uint32_t mpts, uint32_t* u, uint32_t* v, float* w) {
    if (mpts == 0) mpts = 1;
    if (mpts > sg->k) return (size_t)-1;
    const size_t m = orc_undirected_edges(sg->ids, sg->dist, sg->count, sg->n, sg->k, u, v, w);
    if (!u) return m + sg->nbridges;
    float* core = (float*)malloc((size_t)(sg->n ? sg->n : 1) * 4);
    orc_core_distances(sg->dist, sg->count, sg->n, sg->k, mpts, core);
    for (size_t i = 0; i < m; ++i) {
        const float c = core[u[i]] < core[v[i]] ? core[v[i]] : core[u[i]];
        w[i] = w[i] < c ? c : w[i];
    }
    for (size_t i = 0; i < sg->nbridges; ++i) {
        const orc_edge e = sg->bridges[i];
        const float c = core[e.u] < core[e.v] ? core[e.v] : core[e.u];
        u[m + i] = e.u;
        v[m + i] = e.v;
        w[m + i] = e.w < c ? c : e.w;
    }
    free(core);
    return m + sg->nbridges;
}

/* src/coresg.cpp:408-410 (CoreSG::mst) -> mst.cpp:153-156 with one thread,
 * i.e. Kruskal. Returns the MST edge count, or (size_t)-1 for a bad mpts. */
size_t orc_sg_mst(const orc_sg* sg, uint32_t mpts, uint32_t* out_u, uint32_t* out_v,
                  float* out_w, uint32_t* out_labels, uint32_t* out_ncomp, double* out_weight) {
    const size_t m = orc_sg_edges(sg, mpts, NULL, NULL, NULL);
    if (m == (size_t)-1) return m;
    uint32_t* u = (uint32_t*)malloc((m + 1) * 4);
    uint32_t* v = (uint32_t*)malloc((m + 1) * 4);
    float* w = (float*)malloc((m + 1) * 4);
    orc_sg_edges(sg, mpts, u, v, w);
    const size_t t = orc_mst_kruskal(sg->n, u, v, w, m, out_u, out_v, out_w, out_labels,
                                     out_ncomp, out_weight);
    free(u);
    free(v);
    free(w);
    return t;
}

/* src/coresg.cpp:419-430 (CoreSG::exact_mst): core distances for
 * min(mpts, k_max) (coresg.cpp:68-70; mpts 0 reads column 0, knn_graph.hpp:157-163), then mst_prim_virtual (mst.cpp:172-220) whose row i is
 * CpuBackend::mrd_block's row (cpu_backend.cpp:61-90), i.e. orc_mrd_block.
 * Edges in Prim acceptance order, canonicalised u<v; labels and component
 * count from a union-find over them (mst.cpp:211-212, :43-46); weight = the
 * double sum in acceptance order (types.hpp:58-62). Returns the edge count. */
size_t orc_sg_exact_mst(const orc_sg* sg, uint32_t mpts, uint32_t* out_u, uint32_t* out_v,
                        float* out_w, uint32_t* out_labels, uint32_t* out_ncomp,
                        double* out_weight) {
    if (mpts > sg->k) mpts = sg->k;
    const uint32_t n = sg->n;
    float* core = (float*)malloc((n ? n : 1) * 4);
    orc_core_distances(sg->dist, sg->count, n, sg->k, mpts, core);
    float* row = (float*)malloc((n ? n : 1) * 4);
    float* best = (float*)malloc((n ? n : 1) * 4);
    uint32_t* from = (uint32_t*)malloc((n ? n : 1) * 4);
    uint8_t* in_tree = (uint8_t*)calloc(n ? n : 1, 1);
    union_find uf;
    uf_init(&uf, n);
    size_t t = 0;
    double weight = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        best[i] = INFINITY;
        from[i] = ORC_INVALID;
    }
    if (n >= 2) {
        uint32_t current = 0;
        in_tree[0] = 1;
        best[0] = 0.0f;
        for (uint32_t step = 1; step < n; ++step) {
            orc_mrd_block(sg->pts, n, sg->stride, core, current, (size_t)current + 1, row);
            for (uint32_t j = 0; j < n; ++j) {
                if (in_tree[j]) continue;
                if (row[j] < best[j]) {
                    best[j] = row[j];
                    from[j] = current;
                }
            }
            uint32_t pick = ORC_INVALID;
            float pick_w = INFINITY;
            for (uint32_t j = 0; j < n; ++j) {
                if (in_tree[j]) continue;
                if (best[j] < pick_w || (best[j] == pick_w && pick == ORC_INVALID)) {
                    pick_w = best[j];
                    pick = j;
                }
            }
            if (pick == ORC_INVALID) break;
            in_tree[pick] = 1;
            if (from[pick] != ORC_INVALID && pick_w < INFINITY) {
                const uint32_t a = from[pick], b = pick;
                out_u[t] = a < b ? a : b;
                out_v[t] = a < b ? b : a;
                out_w[t] = pick_w;
                weight += (double)pick_w;
                uf_unite(&uf, a, b);
                ++t;
            }
            current = pick;
        }
    }
    if (out_labels) uf_labels(&uf, out_labels);
    if (out_ncomp) *out_ncomp = n ? uf.components : 0;
    if (out_weight) *out_weight = weight;
    uf_free(&uf);
    free(core);
    free(row);
    free(best);
    free(from);
    free(in_tree);
    return t;
}

/* src/eval.cpp:24-97 (sampled_recall): queries drawn by a partial
 * Fisher-Yates over mt19937_64(seed); exact top-k of each query by squared
 * distance, i.e. the k smallest (rank, id) pairs; recall = mean hit fraction;
 * ratio = sum of approximate k-th distances / sum of exact ones. */
double orc_sg_sampled_recall(const orc_sg* sg, size_t queries, uint64_t seed, double* ratio) {
    const uint32_t n = sg->n, k = sg->k, stride = sg->stride;
    if (n == 0 || k == 0 || queries == 0) {
        if (ratio) *ratio = 0.0;
        return 0.0;
    }
    uint32_t* qs = (uint32_t*)malloc((size_t)n * 4);
    for (uint32_t i = 0; i < n; ++i) qs[i] = i;
    mt64* g = (mt64*)malloc(sizeof(mt64));
    mt64_seed(g, seed);
    size_t nq = n;
    if (queries < n) {
        for (size_t i = 0; i < queries; ++i) {
            const size_t j = i + (size_t)(mt64_next(g) % (uint64_t)(n - i));
            const uint32_t t = qs[i];
            qs[i] = qs[j];
            qs[j] = t;
        }
        nq = queries;
    }
    free(g);
    float* br = (float*)malloc((size_t)k * 4);
    uint32_t* bi = (uint32_t*)malloc((size_t)k * 4);
    double sum_hit = 0.0, sum_true = 0.0, sum_approx = 0.0;
    for (size_t t = 0; t < nq; ++t) {
        const uint32_t q = qs[t];
        const float* vq = sg->pts + (size_t)q * stride;
        uint32_t filled = 0;
        for (uint32_t u = 0; u < n; ++u) {
            if (u == q) continue;
            const float d = orc_l2sqr(vq, sg->pts + (size_t)u * stride, stride);
            /* sorted insertion keeps the k smallest (d, id); ids ascend, so an
             * equal d never displaces an earlier id. */
            if (filled < k) {
                uint32_t pos = filled++;
                while (pos > 0 && br[pos - 1] > d) {
                    br[pos] = br[pos - 1];
                    bi[pos] = bi[pos - 1];
                    --pos;
                }
                br[pos] = d;
                bi[pos] = u;
            } else if (d < br[k - 1]) {
                uint32_t pos = k - 1;
                while (pos > 0 && br[pos - 1] > d) {
                    br[pos] = br[pos - 1];
                    bi[pos] = bi[pos - 1];
                    --pos;
                }
                br[pos] = d;
                bi[pos] = u;
            }
        }
        const uint32_t ca = sg->count[q] < k ? sg->count[q] : k;
        size_t h = 0;
        for (uint32_t j = 0; j < ca; ++j) {
            const uint32_t id = sg->ids[(size_t)q * k + j];
            for (uint32_t s = 0; s < filled; ++s)
                if (bi[s] == id) {
                    ++h;
                    break;
                }
        }
        sum_hit += filled ? (double)h / (double)filled : 0.0;
        sum_true += filled ? (double)sqrtf(br[filled - 1]) : 0.0;
        sum_approx += ca ? (double)sg->dist[(size_t)q * k + ca - 1] : 0.0;
    }
    free(br);
    free(bi);
    free(qs);
    if (ratio) *ratio = sum_true > 0.0 ? sum_approx / sum_true : 1.0;
    return sum_hit / (double)nq;
}
