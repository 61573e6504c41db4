// Process-wide runtime settings and resources of the library: options,
// the capture stream used to record Krylov iterations as HIP graphs, and
// the pinned host buffer that per-RHS scalars are read back through.

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "common.hpp"

namespace {

std::mutex g_mu;
// graphs on; matrix-free kernels: FP64 and FP16-split matrix cores; overlapped product phases
int g_options[MPA_OPT_COUNT] = {1, 0, 0, 1};
std::map<int, hipStream_t> g_capture;  // per device
std::map<int, SideStream> g_side;      // per device
std::mutex g_side_mu;                  // held by a SideLease (apart from g_mu: its holder reads options)
void* g_pinned = nullptr;
size_t g_pinned_bytes = 0;

}  // namespace

int mpa_option(int key) {
    std::lock_guard<std::mutex> lock(g_mu);
    return (key >= 0 && key < MPA_OPT_COUNT) ? g_options[key] : 0;
}

hipStream_t mpa_capture_stream() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_capture.find(dev);
    if (it != g_capture.end()) return it->second;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
    g_capture[dev] = s;
    return s;
}

SideLease::SideLease() : lock_(g_side_mu), side_(nullptr) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    auto it = g_side.find(dev);
    if (it == g_side.end()) {
        SideStream s{};
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess) return;
        if (hipEventCreateWithFlags(&s.fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s.join, hipEventDisableTiming) != hipSuccess) {
            if (s.fork) (void)hipEventDestroy(s.fork);
            (void)hipStreamDestroy(s.stream);
            return;
        }
        it = g_side.emplace(dev, s).first;
    }
    side_ = &it->second;
}

PinnedLease::PinnedLease(size_t bytes) : lock_(g_mu), ptr_(nullptr) {
    if (g_pinned_bytes < bytes) {
        if (g_pinned) (void)hipHostFree(g_pinned);
        g_pinned = nullptr;
        g_pinned_bytes = 0;
        size_t want = bytes < 65536 ? 65536 : bytes * 2;
        if (hipHostMalloc(&g_pinned, want, hipHostMallocDefault) == hipSuccess) g_pinned_bytes = want;
        else g_pinned = nullptr;
    }
    ptr_ = g_pinned;
}

extern "C" {

int mpa_set_option(int key, int value) {
    if (key < 0 || key >= MPA_OPT_COUNT) {
        mpa_set_error("set_option: unknown key %d", key);
        return MPA_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    g_options[key] = value;
    return MPA_OK;
}

int mpa_get_option(int key) { return mpa_option(key); }

}  // extern "C"
