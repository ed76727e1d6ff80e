// Named timing ranges recorded with HIP events on the library stream: the
// counterpart of the reference's NVTX ranges (nvtxStartRange/EndRange,
// mod_nvtx.f90:50-101), used to measure per-kernel durations in-process.
#include <string>
#include <vector>

#include "sod2d_internal.hpp"

namespace sod {

namespace {
struct Rec {
    int id;
    hipEvent_t a, b;
};
bool g_on = false;
std::vector<std::string> g_names;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;

hipEvent_t get_event() {
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

int prof_id(const char* name) {
    for (size_t i = 0; i < g_names.size(); ++i)
        if (g_names[i] == name) return (int)i;
    g_names.emplace_back(name);
    return (int)g_names.size() - 1;
}

bool prof_set(bool on) {
    const bool was = g_on;
    g_on = on;
    return was;
}

ProfScope::ProfScope(int id_) : id(id_), rec(-1) {
    if (!g_on) return;
    hipEvent_t a = get_event(), b = get_event();
    if (!a || !b) return;
    hipEventRecord(a, rt().stream);
    g_recs.push_back(Rec{id, a, b});
    rec = (int)g_recs.size() - 1;
}

ProfScope::~ProfScope() {
    if (rec >= 0) hipEventRecord(g_recs[rec].b, rt().stream);
}

}  // namespace sod

using namespace sod;

extern "C" {

int sod_profile_enable(int on) {
    g_on = on != 0;
    return SOD_OK;
}

int sod_profile_reset(void) {
    SOD_HIP(hipStreamSynchronize(rt().stream));
    for (Rec& r : g_recs) {
        g_pool.push_back(r.a);
        g_pool.push_back(r.b);
    }
    g_recs.clear();
    return SOD_OK;
}

int sod_profile_get(const char* name, double* total_ms, int* count) {
    *total_ms = 0.0;
    *count = 0;
    SOD_HIP(hipStreamSynchronize(rt().stream));
    for (size_t i = 0; i < g_names.size(); ++i) {
        if (g_names[i] != name) continue;
        for (Rec& r : g_recs) {
            if (r.id != (int)i) continue;
            float ms = 0.f;
            SOD_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            *total_ms += ms;
            *count += 1;
        }
    }
    return SOD_OK;
}

}  // extern "C"
