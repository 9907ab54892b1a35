// The opaque ABI handle of a resident graph and the ABI's error convention,
// shared by the translation units that implement extern "C" entry points.
#pragma once

#include <new>
#include <string>

#include "graph.hpp"

struct coresg_hip_graph {
    coresg_hip::Graph g;
};

namespace coresg_hip {

void set_last_error(const std::string& msg);
const char* get_last_error();

// Runs `f`, converting exceptions into the 0-return + last_error convention.
// HIP keeps the last failing call's code per thread until it is read, and the
// launch checks read it (hipGetLastError): an error a caller of this library
// left behind, or one a handled failure left, would otherwise surface at the
// next unrelated launch. Every entry point therefore starts with it cleared.
template <typename F>
int guarded(F&& f) {
    (void)hipGetLastError();
    struct PlainReset {
        PlainReset() { t_plain_stream = nullptr; }
        ~PlainReset() { t_plain_stream = nullptr; }
    } plain_reset;
    try {
        f();
        return 1;
    } catch (const HipError& e) {
        std::string msg = e.what();
        if (e.code == hipErrorOutOfMemory) msg += " (out of memory)";
        set_last_error(msg);
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed (out of memory)");
    } catch (const std::exception& e) {
        set_last_error(e.what());
    }
    return 0;
}

// guarded() for host-only entry points (no HIP call at all, so they also run
// on a machine without a GPU).
template <typename F>
int guarded_host(F&& f) {
    try {
        f();
        return 1;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed (out of memory)");
    } catch (const std::exception& e) {
        set_last_error(e.what());
    }
    return 0;
}

}  // namespace coresg_hip
