// The reference's on-disk formats for a resident CORE-SG.
//
//   * .coresg (CoreSG::save / CoreSG::load, src/io.cpp:232-330): magic
//     "CORESG01"; u32 k_max, dim, stride, metric, missing-edge mode; u64 n;
//     u64 float count and the padded rows; the u64-counted SoA vectors ids,
//     dist, flags, count; the u64-counted 12-byte bridge records {u, v, w};
//     u64 repair count. Little-endian, no padding. A loaded graph recomputes
//     the derived hubness state (components, degrees, scores, selection) as
//     the reference's load does, with the default parameters.
//   * MST edges as TSV (write_edges_tsv, io.cpp:131-138): header "u\tv\tweight",
//     the weight in scientific notation with 9 digits; labels as TSV
//     (write_labels_tsv, io.cpp:140-145).
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/coresg_hip.h"
#include "build.hpp"
#include "handle.hpp"

namespace coresg_hip {
namespace {

constexpr char kMagic[8] = {'C', 'O', 'R', 'E', 'S', 'G', '0', '1'};

struct File {
    std::FILE* f = nullptr;
    std::string path;
    File(const char* p, const char* mode) : path(p ? p : "") {
        if (!p) throw std::runtime_error("null path");
        f = std::fopen(p, mode);
        if (!f) throw std::runtime_error("cannot open " + path);
    }
    ~File() {
        if (f) std::fclose(f);
    }
    void write(const void* d, std::size_t bytes) {
        if (bytes && std::fwrite(d, 1, bytes, f) != bytes)
            throw std::runtime_error("write failed for " + path);
    }
    void read(void* d, std::size_t bytes) {
        if (bytes && std::fread(d, 1, bytes, f) != bytes)
            throw std::runtime_error("truncated coresg file " + path);
    }
    template <typename T>
    void pod(const T& v) { write(&v, sizeof(T)); }
    template <typename T>
    T get() {
        T v;
        read(&v, sizeof(T));
        return v;
    }
    void close() {
        if (f && std::fclose(f) != 0) {
            f = nullptr;
            throw std::runtime_error("write failed for " + path);
        }
        f = nullptr;
    }
};

template <typename T>
void write_device_vec(File& out, const T* dev, std::size_t count, hipStream_t s) {
    std::vector<T> h(count);
    if (count) {
        CSG_CHECK(hipMemcpyAsync(h.data(), dev, count * sizeof(T), hipMemcpyDeviceToHost, s));
        CSG_CHECK(hipStreamSynchronize(s));
    }
    out.pod(std::uint64_t(count));
    out.write(h.data(), count * sizeof(T));
}

template <typename T>
std::vector<T> read_vec(File& in, std::uint64_t expect, const char* what) {
    const std::uint64_t count = in.get<std::uint64_t>();
    if (count != expect)
        throw std::runtime_error(std::string("coresg file: ") + what + " has " +
                                 std::to_string(count) + " entries, expected " +
                                 std::to_string(expect));
    std::vector<T> v(count);
    in.read(v.data(), count * sizeof(T));
    return v;
}

}  // namespace

void save_graph(Graph& g, const char* path) {
    if (!g.points_set || !g.knn_set)
        throw std::runtime_error("save: points and k-NN lists must be set");
    File out(path, "wb");
    out.write(kMagic, 8);
    out.pod(std::uint32_t(g.k));
    out.pod(std::uint32_t(g.dim));
    out.pod(std::uint32_t(g.stride));
    out.pod(std::uint32_t(CORESG_HIP_METRIC_L2));
    out.pod(std::uint32_t(g.missing_mode));
    out.pod(std::uint64_t(g.n));
    const std::size_t nk = std::size_t(g.n) * g.k;
    write_device_vec(out, g.pts.p, std::size_t(g.n) * g.stride, g.stream);  // u64 floats + rows
    write_device_vec(out, g.ids.p, nk, g.stream);
    write_device_vec(out, g.dist.p, nk, g.stream);
    write_device_vec(out, g.flags.p, nk, g.stream);
    write_device_vec(out, g.count.p, std::size_t(g.n), g.stream);
    const std::uint64_t m = g.bu.size();
    out.pod(m);
    for (std::size_t i = 0; i < m; ++i) {
        out.pod(g.bu[i]);
        out.pod(g.bv[i]);
        out.pod(g.bw[i]);
    }
    out.pod(std::uint64_t(g.repair_count));
    out.close();
}

// Reads a .coresg into a new graph on `device`; the caller owns the result.
coresg_hip_graph* load_graph(int device, const char* path) {
    File in(path, "rb");
    char magic[8] = {0};
    in.read(magic, 8);
    if (std::memcmp(magic, kMagic, 8) != 0)
        throw std::runtime_error(std::string(path) + " is not a coresg file");
    const std::uint32_t k = in.get<std::uint32_t>(), dim = in.get<std::uint32_t>(),
                        stride = in.get<std::uint32_t>(), metric = in.get<std::uint32_t>(),
                        mode = in.get<std::uint32_t>();
    const std::uint64_t n = in.get<std::uint64_t>();
    if (metric != CORESG_HIP_METRIC_L2)
        throw std::runtime_error("load: only L2 graphs are held on the device");
    if (k == 0 || dim == 0 || stride != ((dim + 7u) & ~7u) || n >= 0x7FFFFFFFull)
        throw std::runtime_error("load: inconsistent coresg header in " + std::string(path));
    const std::uint64_t nk = n * k;
    const std::vector<float> pts = read_vec<float>(in, n * stride, "points");
    const std::vector<std::uint32_t> ids = read_vec<std::uint32_t>(in, nk, "ids");
    const std::vector<float> dist = read_vec<float>(in, nk, "dist");
    const std::vector<std::uint8_t> flags = read_vec<std::uint8_t>(in, nk, "flags");
    const std::vector<std::uint32_t> count = read_vec<std::uint32_t>(in, n, "count");
    const std::uint64_t m = in.get<std::uint64_t>();
    std::vector<std::uint32_t> bu(m), bv(m);
    std::vector<float> bw(m);
    for (std::uint64_t i = 0; i < m; ++i) {
        bu[i] = in.get<std::uint32_t>();
        bv[i] = in.get<std::uint32_t>();
        bw[i] = in.get<float>();
        if (bu[i] >= n || bv[i] >= n) throw std::runtime_error("load: bridge endpoint out of range");
    }
    const std::uint64_t repair = in.get<std::uint64_t>();
    if (std::fgetc(in.f) != EOF) throw std::runtime_error("load: trailing bytes in " + std::string(path));

    std::unique_ptr<coresg_hip_graph> h(new coresg_hip_graph());
    Graph& g = h->g;
    g.device = device;
    g.n = static_cast<std::uint32_t>(n);
    g.dim = dim;
    g.stride = stride;
    g.k = k;
    CSG_CHECK(hipSetDevice(device));
    create_stream(&g.stream);
    g.pts.reserve(n * stride);
    g.ids.reserve(nk);
    g.dist.reserve(nk);
    g.flags.reserve(nk);
    g.count.reserve(n);
    CSG_CHECK(hipMemcpyAsync(g.pts.p, pts.data(), pts.size() * 4, hipMemcpyHostToDevice, g.stream));
    CSG_CHECK(hipMemcpyAsync(g.ids.p, ids.data(), nk * 4, hipMemcpyHostToDevice, g.stream));
    CSG_CHECK(hipMemcpyAsync(g.dist.p, dist.data(), nk * 4, hipMemcpyHostToDevice, g.stream));
    CSG_CHECK(hipMemcpyAsync(g.flags.p, flags.data(), nk, hipMemcpyHostToDevice, g.stream));
    CSG_CHECK(hipMemcpyAsync(g.count.p, count.data(), n * 4, hipMemcpyHostToDevice, g.stream));
    CSG_CHECK(hipStreamSynchronize(g.stream));
    g.points_set = g.knn_set = true;
    g.bu = std::move(bu);
    g.bv = std::move(bv);
    g.bw = std::move(bw);
    g.repair_count = static_cast<std::size_t>(repair);
    g.missing_mode = mode;
    coresg_hip_build_params p;
    coresg_hip_default_build_params(&p);
    // a CoreSG constructed with default parameters sets k_for_score to their
    // k_max (32, coresg.cpp:51) and load() keeps it whatever the file's k_max
    p.score_k = 32;
    derive_hubness(g, p);
    return h.release();
}

void write_edges_tsv(const char* path, const std::uint32_t* u, const std::uint32_t* v,
                     const float* w, std::size_t m) {
    File out(path, "w");
    if (std::fputs("u\tv\tweight\n", out.f) < 0) throw std::runtime_error("write failed");
    for (std::size_t i = 0; i < m; ++i)
        if (std::fprintf(out.f, "%u\t%u\t%.9e\n", u[i], v[i], double(w[i])) < 0)
            throw std::runtime_error("write failed for " + out.path);
    out.close();
}

void write_labels_tsv(const char* path, const std::uint32_t* labels, std::size_t n) {
    File out(path, "w");
    if (std::fputs("id\tlabel\n", out.f) < 0) throw std::runtime_error("write failed");
    for (std::size_t i = 0; i < n; ++i)
        if (std::fprintf(out.f, "%zu\t%u\n", i, labels[i]) < 0)
            throw std::runtime_error("write failed for " + out.path);
    out.close();
}

}  // namespace coresg_hip

extern "C" {

int coresg_hip_graph_save(coresg_hip_graph* g, const char* path) {
    return coresg_hip::guarded([&] {
        CSG_CHECK(hipSetDevice(g->g.device));
        coresg_hip::save_graph(g->g, path);
    });
}

coresg_hip_graph* coresg_hip_graph_load(int device, const char* path) {
    coresg_hip_graph* out = nullptr;
    const int ok = coresg_hip::guarded([&] { out = coresg_hip::load_graph(device, path); });
    return ok ? out : nullptr;
}

int coresg_hip_write_edges_tsv(const char* path, const uint32_t* u, const uint32_t* v,
                               const float* w, size_t m) {
    return coresg_hip::guarded_host([&] { coresg_hip::write_edges_tsv(path, u, v, w, m); });
}

int coresg_hip_write_labels_tsv(const char* path, const uint32_t* labels, size_t n) {
    return coresg_hip::guarded_host([&] { coresg_hip::write_labels_tsv(path, labels, n); });
}

}  // extern "C"
