// Golden-vector harness over the reference library.
//
// This file is test infrastructure. It links the reference's own CPU sources
// (compiled in place by oracle/Makefile, never copied) and dumps the values its
// functions produce, so the CPU restatement in oracle/coresg_oracle.c and the
// HIP kernels can be checked against the real thing on identical inputs.
//
// Output format ("tagged arrays"): a sequence of records
//   char name[32] | u32 dtype | u64 count | count * sizeof(dtype) bytes
// dtype: 1=u8 4=u32 5=i32 8=u64 9=f32 10=f64.  tests/golden/refdump.py reads it.
//
// Commands (all single-threaded, so the reference schedule is deterministic):
//   l2     DATA DIM PAIRS OUT        l2sqr + metric_distance for index pairs
//   init   DATA DIM K SEED OUT       init_random on an empty graph
//   nnd    DATA DIM K SAMPLE ITERS SEED OUT   init_random + nndescent
//   brute  DATA DIM K OUT            brute_force_knn
//   state  CORESG OUT                hubness / connectivity / edges after load()
//   mst    CORESG MPTS OUT           CoreSG::mst(mpts), Kruskal order
//   exact  CORESG MPTS OUT           CoreSG::exact_mst(mpts), Prim acceptance order
//   mrd    DATA DIM IDS CORE OUT     CPU mrd_block over a gathered subset
//   recall CORESG QUERIES SEED OUT   sampled_recall against brute force
//   reps   DATA DIM K MAXREPS MINPC BFLIMIT OUT.coresg
//                                    CoreSG::build with max_reps / the
//                                    anti-hub quota / brute-force limit set,
//                                    saved (the repair's representative cut)

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "coresg/backend.hpp"
#include "coresg/coresg.hpp"
#include "coresg/eval.hpp"
#include "coresg/knn_graph.hpp"
#include "coresg/mst.hpp"
#include "coresg/nndescent.hpp"

using namespace coresg;

namespace {

struct Tagged {
    std::ofstream out;
    explicit Tagged(const std::string& path) : out(path, std::ios::binary) {
        if (!out) throw std::runtime_error("cannot write " + path);
    }
    void put(const char* name, std::uint32_t dtype, std::size_t elem, const void* data,
             std::size_t count) {
        char nm[32] = {0};
        std::strncpy(nm, name, 31);
        out.write(nm, 32);
        out.write(reinterpret_cast<const char*>(&dtype), 4);
        std::uint64_t c = count;
        out.write(reinterpret_cast<const char*>(&c), 8);
        if (count) out.write(reinterpret_cast<const char*>(data), std::streamsize(count * elem));
    }
    template <typename T>
    void vec(const char* name, std::uint32_t dtype, const std::vector<T>& v) {
        put(name, dtype, sizeof(T), v.data(), v.size());
    }
};

std::vector<float> read_f32(const std::string& path) {
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (!in) throw std::runtime_error("cannot open " + path);
    const std::size_t bytes = static_cast<std::size_t>(in.tellg());
    in.seekg(0);
    std::vector<float> v(bytes / 4);
    in.read(reinterpret_cast<char*>(v.data()), std::streamsize(v.size() * 4));
    return v;
}

template <typename T>
std::vector<T> read_raw(const std::string& path) {
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (!in) throw std::runtime_error("cannot open " + path);
    const std::size_t bytes = static_cast<std::size_t>(in.tellg());
    in.seekg(0);
    std::vector<T> v(bytes / sizeof(T));
    in.read(reinterpret_cast<char*>(v.data()), std::streamsize(v.size() * sizeof(T)));
    return v;
}

Dataset load_data(const std::string& path, std::uint32_t dim) {
    std::vector<float> raw = read_f32(path);
    Dataset ds(dim, MetricKind::L2);
    ds.append(raw);
    return ds;
}

void dump_graph(Tagged& t, const KnnGraph& g) {
    t.vec("ids", 4, g.ids_storage());
    t.vec("dist", 9, g.dist_storage());
    t.vec("flags", 1, g.flag_storage());
    t.vec("count", 4, g.count_storage());
}

void dump_edges(Tagged& t, const char* prefix, const std::vector<Edge>& e) {
    std::vector<std::uint32_t> u(e.size()), v(e.size());
    std::vector<float> w(e.size());
    for (std::size_t i = 0; i < e.size(); ++i) {
        u[i] = e[i].u;
        v[i] = e[i].v;
        w[i] = e[i].w;
    }
    const std::string p(prefix);
    t.vec((p + "_u").c_str(), 4, u);
    t.vec((p + "_v").c_str(), 4, v);
    t.vec((p + "_w").c_str(), 9, w);
}

CoreSGParams single_thread_params() {
    CoreSGParams p;
    p.threads = 1;
    p.verbose = false;
    p.backend.prefer_gpu = false;
    return p;
}

int run(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string cmd = argv[1];
    if (cmd == "l2" && argc == 6) {
        const std::uint32_t dim = std::strtoul(argv[3], nullptr, 10);
        Dataset ds = load_data(argv[2], dim);
        std::vector<std::uint32_t> pairs = read_raw<std::uint32_t>(argv[4]);
        std::vector<float> r(pairs.size() / 2), d(pairs.size() / 2);
        for (std::size_t i = 0; i + 1 < pairs.size(); i += 2) {
            r[i / 2] = l2sqr(ds.row(pairs[i]), ds.row(pairs[i + 1]), ds.stride());
            d[i / 2] = metric_distance(MetricKind::L2, ds.row(pairs[i]), ds.row(pairs[i + 1]),
                                       ds.stride());
        }
        Tagged t(argv[5]);
        t.vec("l2sqr", 9, r);
        t.vec("dist", 9, d);
        return 0;
    }
    if (cmd == "init" && argc == 7) {
        const std::uint32_t dim = std::strtoul(argv[3], nullptr, 10);
        const std::uint32_t k = std::strtoul(argv[4], nullptr, 10);
        const std::uint64_t seed = std::strtoull(argv[5], nullptr, 10);
        Dataset ds = load_data(argv[2], dim);
        KnnGraph g(static_cast<idx_t>(ds.size()), k);
        init_random(ds.view(), MetricKind::L2, &g, seed, nullptr);
        Tagged t(argv[6]);
        dump_graph(t, g);
        return 0;
    }
    if (cmd == "nnd" && argc == 9) {
        const std::uint32_t dim = std::strtoul(argv[3], nullptr, 10);
        NNDescentParams p;
        p.k = std::strtoul(argv[4], nullptr, 10);
        p.sample = std::strtoul(argv[5], nullptr, 10);
        p.iters = std::strtoul(argv[6], nullptr, 10);
        p.seed = std::strtoull(argv[7], nullptr, 10);
        Dataset ds = load_data(argv[2], dim);
        KnnGraph g(static_cast<idx_t>(ds.size()), p.k);
        init_random(ds.view(), MetricKind::L2, &g, p.seed, nullptr);
        NNDescentStats st = nndescent(ds.view(), MetricKind::L2, &g, p, nullptr);
        Tagged t(argv[8]);
        dump_graph(t, g);
        std::vector<std::uint64_t> s = {st.rounds, st.updates, st.distance_evals};
        t.vec("stats", 8, s);
        return 0;
    }
    if (cmd == "brute" && argc == 6) {
        const std::uint32_t dim = std::strtoul(argv[3], nullptr, 10);
        const std::uint32_t k = std::strtoul(argv[4], nullptr, 10);
        Dataset ds = load_data(argv[2], dim);
        KnnGraph g(static_cast<idx_t>(ds.size()), k);
        brute_force_knn(ds.view(), MetricKind::L2, &g, nullptr);
        Tagged t(argv[5]);
        dump_graph(t, g);
        return 0;
    }
    if (cmd == "state" && argc == 4) {
        CoreSG sg(single_thread_params());
        sg.load(argv[2]);
        const idx_t n = sg.knn().size();
        std::vector<std::int32_t> rdeg(n);
        std::vector<double> score(n);
        for (idx_t v = 0; v < n; ++v) {
            rdeg[v] = static_cast<std::int32_t>(sg.hubness().reverse_degree(v));
            score[v] = sg.hubness().reverse_score(v);
        }
        Tagged t(argv[3]);
        t.vec("rdeg", 5, rdeg);
        t.vec("score", 10, score);
        t.vec("selected", 4, sg.hubness().selected());
        t.vec("conn_labels", 4, sg.connectivity().labels());
        std::vector<std::uint32_t> nc = {sg.connectivity().components()};
        t.vec("conn_components", 4, nc);
        dump_edges(t, "undirected", sg.knn().undirected_edges(nullptr));
        dump_edges(t, "missing", sg.missing_edges());
        std::vector<std::uint64_t> rc = {sg.repair_edge_count()};
        t.vec("repair_count", 8, rc);
        return 0;
    }
    if (cmd == "mst" && argc == 5) {
        CoreSG sg(single_thread_params());
        sg.load(argv[2]);
        const std::uint32_t mpts = std::strtoul(argv[3], nullptr, 10);
        MstResult r = sg.mst(mpts);
        Tagged t(argv[4]);
        dump_edges(t, "mst", r.edges);
        t.vec("component", 4, r.component);
        std::vector<std::uint32_t> nc = {r.num_components};
        t.vec("num_components", 4, nc);
        std::vector<double> w = {r.weight};
        t.vec("weight", 10, w);
        dump_edges(t, "edges", sg.edges(mpts));
        return 0;
    }
    if (cmd == "exact" && argc == 5) {
        CoreSG sg(single_thread_params());
        sg.load(argv[2]);
        const std::uint32_t mpts = std::strtoul(argv[3], nullptr, 10);
        MstResult r = sg.exact_mst(mpts);
        Tagged t(argv[4]);
        dump_edges(t, "mst", r.edges);
        t.vec("component", 4, r.component);
        std::vector<std::uint32_t> nc = {r.num_components};
        t.vec("num_components", 4, nc);
        std::vector<double> w = {r.weight};
        t.vec("weight", 10, w);
        return 0;
    }
    if (cmd == "mrd" && argc == 7) {
        const std::uint32_t dim = std::strtoul(argv[3], nullptr, 10);
        Dataset ds = load_data(argv[2], dim);
        std::vector<std::uint32_t> ids = read_raw<std::uint32_t>(argv[4]);
        std::vector<float> core = read_raw<float>(argv[5]);
        const std::vector<float> pts = ds.gather(ids);
        const std::size_t m = ids.size();
        std::vector<float> mat(m * m);
        std::unique_ptr<Backend> be = make_cpu_backend(1, nullptr);
        be->mrd_block(0, pts.data(), m, ds.dim(), ds.stride(), MetricKind::L2, core.data(), 0, m,
                      mat.data());
        MstResult r = mst_dense(mat.data(), m);
        Tagged t(argv[6]);
        t.vec("matrix", 9, mat);
        dump_edges(t, "prim", r.edges);
        return 0;
    }
    if (cmd == "recall" && argc == 6) {
        CoreSG sg(single_thread_params());
        sg.load(argv[2]);
        const std::size_t q = std::strtoull(argv[3], nullptr, 10);
        const std::uint64_t seed = std::strtoull(argv[4], nullptr, 10);
        RecallResult r = sampled_recall(sg.data(), sg.knn(), sg.knn().k(), q, seed, nullptr);
        Tagged t(argv[5]);
        std::vector<double> v = {r.recall, r.distance_ratio, double(r.nodes)};
        t.vec("recall", 10, v);
        return 0;
    }
    if (cmd == "reps" && argc == 9) {
        const std::uint32_t dim = std::strtoul(argv[3], nullptr, 10);
        CoreSGParams p = single_thread_params();
        p.k_max = std::strtoul(argv[4], nullptr, 10);
        p.max_reps = std::strtoull(argv[5], nullptr, 10);
        p.hub.min_per_component = std::strtoull(argv[6], nullptr, 10);
        p.incr.brute_force_limit = std::strtoull(argv[7], nullptr, 10);
        std::vector<float> raw = read_f32(argv[2]);
        CoreSG sg(p);
        sg.build(raw.data(), raw.size() / dim, dim);
        sg.save(argv[8]);
        return 0;
    }
    std::fprintf(stderr, "ref_dump: bad command line\n");
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_dump: %s\n", e.what());
        return 1;
    }
}
