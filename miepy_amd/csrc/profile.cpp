// Optional per-kernel timing with HIP events, recorded on the stream each
// kernel group is launched on.  Disabled by default; when enabled, every
// instrumented launch group (one matvec, one assembly) records a start/stop
// event pair, and mpa_profile_read() returns the count and summed duration.

#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "common.hpp"

namespace {

struct Rec {
    int kind;
    hipEvent_t start, stop;
};

std::mutex g_mu;
bool g_enabled = false;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;

hipEvent_t take() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

}  // namespace

ProfScope::ProfScope(int kind, hipStream_t stream) : kind_(kind), stream_(stream), start_(nullptr) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_enabled) return;
    start_ = take();
    if (start_) (void)hipEventRecord((hipEvent_t)start_, stream_);
}

ProfScope::~ProfScope() {
    if (!start_) return;
    std::lock_guard<std::mutex> lock(g_mu);
    hipEvent_t stop = take();
    if (!stop) return;
    (void)hipEventRecord(stop, stream_);
    g_recs.push_back({kind_, (hipEvent_t)start_, stop});
}

bool mpa_profile_active() {
    std::lock_guard<std::mutex> lock(g_mu);
    return g_enabled;
}

extern "C" {

int mpa_profile_enable(int on) {
    std::lock_guard<std::mutex> lock(g_mu);
    for (auto& r : g_recs) {
        (void)hipEventSynchronize(r.stop);
        g_pool.push_back(r.start);
        g_pool.push_back(r.stop);
    }
    g_recs.clear();
    g_enabled = on != 0;
    return MPA_OK;
}

int mpa_profile_read(int kind, int* count, double* total_ms) {
    std::lock_guard<std::mutex> lock(g_mu);
    int c = 0;
    double t = 0;
    for (auto& r : g_recs) {
        if (r.kind != kind) continue;
        if (hipEventSynchronize(r.stop) != hipSuccess) {
            mpa_set_error("profile: event synchronize failed");
            return MPA_ERR_HIP;
        }
        float ms = 0;
        if (hipEventElapsedTime(&ms, r.start, r.stop) != hipSuccess) {
            mpa_set_error("profile: elapsed time failed");
            return MPA_ERR_HIP;
        }
        c++;
        t += ms;
    }
    if (count) *count = c;
    if (total_ms) *total_ms = t;
    return MPA_OK;
}

}  // extern "C"
