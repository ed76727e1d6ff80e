// Runtime of libsod2d_hip.so: initialisation, error reporting, the stream,
// the numerical parameters, the host->device "present" table that mirrors
// the reference's OpenACC data regions, pointer resolution and scratch.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <vector>

#include "sod2d_internal.hpp"

namespace sod {

int upload_tables();   // tables.hip
void fused_release();     // k_fused.hip
void dt_cache_release();  // k_misc.hip
int fused_invalidate(const void* lo, size_t bytes);      // k_fused.hip
int dt_cache_invalidate(const void* lo, size_t bytes);   // k_misc.hip
int noalloc_setup();      // k_noalloc.hip

namespace {
thread_local char g_err[2048] = "";

struct Block {
    size_t bytes;
    void* dev;
};
std::map<uintptr_t, Block> g_present;   // host start -> block

struct Slot {
    void* p = nullptr;
    size_t bytes = 0;
};
std::vector<Slot> g_scratch;
std::set<std::pair<const void*, int>> g_tables_ok;
}  // namespace

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

Runtime& rt() {
    static Runtime r;
    return r;
}

int ensure_init() {
    if (!rt().initialised) return fail(SOD_ERR_NOT_INIT, "sod_init(3) has not been called");
    return SOD_OK;
}

int finish_call() {
    SOD_HIP(hipGetLastError());
    if (rt().sync_on_return) SOD_HIP(hipStreamSynchronize(rt().stream));
    return SOD_OK;
}

int resolve(const void* p, const char* what, const void** out) {
    *out = nullptr;
    if (p == nullptr) return fail(SOD_ERR_ARG, "argument %s is NULL", what);
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    if (!g_present.empty()) {
        auto it = g_present.upper_bound(a);
        if (it != g_present.begin()) {
            --it;
            if (a < it->first + it->second.bytes) {
                *out = static_cast<const char*>(it->second.dev) + (a - it->first);
                return SOD_OK;
            }
        }
    }
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(SOD_ERR_NOT_PRESENT,
                    "argument %s (%p) is host memory that is not present on the device; "
                    "register it with sod_enter_data()", what, p);
    }
    if (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
        attr.type == hipMemoryTypeUnified) {
        *out = p;
        return SOD_OK;
    }
    // host memory of any kind (pageable or pinned) must go through the
    // present table, as OpenACC's "present" clause requires
    return fail(SOD_ERR_NOT_PRESENT,
                "argument %s (%p) is host memory that is not present on the device; "
                "register it with sod_enter_data()", what, p);
}

int scratch(size_t bytes, void** out, int slot) {
    if ((int)g_scratch.size() <= slot) g_scratch.resize(slot + 1);
    Slot& s = g_scratch[slot];
    if (s.bytes < bytes) {
        if (s.p) SOD_HIP(hipFree(s.p));
        s.p = nullptr;
        s.bytes = 0;
        SOD_HIP(hipMalloc(&s.p, bytes));
        s.bytes = bytes;
    }
    *out = s.p;
    return SOD_OK;
}

static int copy_to_host(void* dst, const void* src, size_t bytes) {
    SOD_HIP(hipStreamSynchronize(rt().stream));
    SOD_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return SOD_OK;
}

// The kernels use the library's own p=3 tables; the caller's tables must be
// the reference's tables (a2ijk = gmsh2ijk, FP32-rounded GLL nodes).  Integer
// tables must match exactly, the derivative table to 1e-12 relative.
// One caller table against the library's: kind 0 the derivative table
// dlxigp_ip (1e-12 relative), 1-4 invAtoIJK, gmshAtoI/J/K (exact).
static int check_table(const void* p, int kind) {
    const HostTables& t = host_tables();
    const void* d = nullptr;
    SOD_TRY(resolve(p, "table", &d));
    if (kind == 0) {
        double h[768];
        SOD_TRY(copy_to_host(h, d, sizeof(h)));
        for (int i = 0; i < 768; ++i) {
            double ref = t.dlxigp_ip[i];
            if (std::fabs(h[i] - ref) > 1e-12 * (1.0 + std::fabs(ref)))
                return fail(SOD_ERR_ARG, "dlxigp_ip differs from the p=3 GLL table at %d (%.17g vs %.17g)", i, h[i],
                            ref);
        }
    } else {
        int h[64];
        SOD_TRY(copy_to_host(h, d, sizeof(h)));
        const int* ref = kind == 1 ? t.invAtoIJK : kind == 2 ? t.gmshAtoI : kind == 3 ? t.gmshAtoJ : t.gmshAtoK;
        if (std::memcmp(h, ref, sizeof(h)) != 0)
            return fail(SOD_ERR_ARG, "node-order table %d differs from a2ijk = gmsh2ijk for p=3", kind);
    }
    return SOD_OK;
}

int check_tables(const double* dlxigp_ip, const int* invAtoIJK, const int* gmshAtoI,
                 const int* gmshAtoJ, const int* gmshAtoK) {
    struct Item {
        const void* p;
        int kind;
    } items[5] = {{dlxigp_ip, 0}, {invAtoIJK, 1}, {gmshAtoI, 2}, {gmshAtoJ, 3}, {gmshAtoK, 4}};
    for (const Item& it : items) {
        if (it.p == nullptr) {
            if (it.kind == 0) continue;   // caller does not use the derivative table
            return fail(SOD_ERR_ARG, "a p=3 table argument is NULL");
        }
        if (g_tables_ok.count({it.p, it.kind})) continue;
        SOD_TRY(check_table(it.p, it.kind));
        g_tables_ok.insert({it.p, it.kind});
    }
    return SOD_OK;
}

int check_tables_noalloc(const int* invAtoIJK, const int* gmshAtoI, const int* gmshAtoJ, const int* gmshAtoK,
                         bool capturing) {
    const void* items[4] = {invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK};
    for (int k = 0; k < 4; ++k) {
        if (items[k] == nullptr) return fail(SOD_ERR_ARG, "a p=3 table argument is NULL");
        if (g_tables_ok.count({items[k], k + 1})) continue;
        // the check copies the table to the host, which a stream under
        // capture does not allow
        if (capturing)
            return fail(SOD_ERR_ARG,
                        "node-order table %d at %p has not been checked: under stream capture the tables must be "
                        "ones an earlier uncaptured call (any entry point taking them) has checked",
                        k + 1, items[k]);
        SOD_TRY(check_table(items[k], k + 1));
    }
    return SOD_OK;
}

}  // namespace sod

using namespace sod;

extern "C" {

const char* sod_last_error(void) { return g_err; }

void sod_default_numerical_params(sod_numerical_params* p) {
    p->ce = 0.1;
    p->cmax = 0.5;
    p->cglob = 1.0;
    p->c_rho = 1.0;
    p->c_ener = 1.0;
    p->Rgas_inf = 1.0;   // mod_numerical_params.f90:45 default; cases override
    p->flag_normalise_entropy = 1;
    p->flag_total_enthalpy = 0;
    p->flag_rk_order = 4;
    p->flag_les = 0;
    p->flag_real_diff = 1;
    p->flag_diff_suth = 1;
}

int sod_init(int porder) {
    if (porder != kP)
        return fail(SOD_ERR_PORDER, "libsod2d_hip is built for porder=3, got porder=%d", porder);
    Runtime& r = rt();
    if (r.initialised) return SOD_OK;
    int ndev = 0;
    SOD_HIP(hipGetDeviceCount(&ndev));
    if (ndev < 1) return fail(SOD_ERR_HIP, "no HIP device visible");
    SOD_HIP(hipGetDevice(&r.device));
    SOD_TRY(upload_tables());
    // the allocation-free entry points' scratch on this device (they report
    // a failure of this lookup; nothing else depends on it)
    (void)noalloc_setup();
    sod_default_numerical_params(&r.params);
    r.stream = nullptr;
    r.initialised = true;
    return SOD_OK;
}

int sod_finalize(void) {
    if (!rt().initialised) return SOD_OK;
    SOD_HIP(hipDeviceSynchronize());
    for (auto& kv : g_present) SOD_HIP(hipFree(kv.second.dev));
    g_present.clear();
    for (auto& s : g_scratch)
        if (s.p) SOD_HIP(hipFree(s.p));
    g_scratch.clear();
    g_tables_ok.clear();
    fused_release();
    dt_cache_release();
    rt().initialised = false;
    return SOD_OK;
}

int sod_set_stream(void* s) {
    rt().stream = static_cast<hipStream_t>(s);
    return SOD_OK;
}

int sod_set_sync(int sync_on_return) {
    rt().sync_on_return = sync_on_return != 0;
    return SOD_OK;
}

int sod_synchronize(void) {
    SOD_TRY(ensure_init());
    SOD_HIP(hipStreamSynchronize(rt().stream));
    return SOD_OK;
}

int sod_set_numerical_params(const sod_numerical_params* p) {
    SOD_TRY(ensure_init());
    if (p->flag_rk_order != 4)
        return fail(SOD_ERR_UNSUPPORTED, "flag_rk_order=%d: only RK4 is implemented", p->flag_rk_order);
    if (p->flag_les != 0)
        return fail(SOD_ERR_UNSUPPORTED, "flag_les=%d: LES models are outside this library", p->flag_les);
    rt().params = *p;
    return SOD_OK;
}

int sod_get_numerical_params(sod_numerical_params* p) {
    SOD_TRY(ensure_init());
    *p = rt().params;
    return SOD_OK;
}

int sod_enter_data(const void* host, size_t bytes, int copyin) {
    SOD_TRY(ensure_init());
    if (host == nullptr || bytes == 0) return fail(SOD_ERR_ARG, "sod_enter_data: empty block");
    const uintptr_t a = reinterpret_cast<uintptr_t>(host);
    auto it = g_present.lower_bound(a);
    if (it != g_present.end() && it->first < a + bytes)
        return fail(SOD_ERR_ARG, "sod_enter_data: block overlaps a present block");
    if (it != g_present.begin()) {
        auto pr = std::prev(it);
        if (pr->first + pr->second.bytes > a)
            return fail(SOD_ERR_ARG, "sod_enter_data: block overlaps a present block");
    }
    void* d = nullptr;
    SOD_HIP(hipMalloc(&d, bytes));
    if (copyin) SOD_HIP(hipMemcpy(d, host, bytes, hipMemcpyHostToDevice));
    g_present[a] = Block{bytes, d};
    return SOD_OK;
}

int sod_exit_data(const void* host, int copyout) {
    SOD_TRY(ensure_init());
    auto it = g_present.find(reinterpret_cast<uintptr_t>(host));
    if (it == g_present.end()) return fail(SOD_ERR_NOT_PRESENT, "sod_exit_data: %p is not present", host);
    if (copyout) {
        SOD_HIP(hipStreamSynchronize(rt().stream));
        SOD_HIP(hipMemcpy(const_cast<void*>(host), it->second.dev, it->second.bytes,
                          hipMemcpyDeviceToHost));
    }
    SOD_HIP(hipStreamSynchronize(rt().stream));
    // mesh data cached from this block must not outlive it (its device
    // address may be handed out again)
    SOD_TRY(fused_invalidate(it->second.dev, it->second.bytes));
    SOD_TRY(dt_cache_invalidate(it->second.dev, it->second.bytes));
    SOD_HIP(hipFree(it->second.dev));
    g_present.erase(it);
    return SOD_OK;
}

static int find_range(const void* host, size_t bytes, void** dev) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(host);
    auto it = g_present.upper_bound(a);
    if (it == g_present.begin()) return fail(SOD_ERR_NOT_PRESENT, "%p is not present", host);
    --it;
    if (a + bytes > it->first + it->second.bytes)
        return fail(SOD_ERR_NOT_PRESENT, "range %p+%zu is not inside a present block", host, bytes);
    *dev = static_cast<char*>(it->second.dev) + (a - it->first);
    return SOD_OK;
}

int sod_update_device(const void* host, size_t bytes) {
    SOD_TRY(ensure_init());
    void* d = nullptr;
    SOD_TRY(find_range(host, bytes, &d));
    SOD_HIP(hipStreamSynchronize(rt().stream));
    SOD_HIP(hipMemcpy(d, host, bytes, hipMemcpyHostToDevice));
    // new contents of a mesh array: its prepared copies are rebuilt
    SOD_TRY(fused_invalidate(d, bytes));
    SOD_TRY(dt_cache_invalidate(d, bytes));
    return SOD_OK;
}

int sod_invalidate_mesh_cache(void) {
    SOD_TRY(ensure_init());
    SOD_TRY(fused_invalidate(nullptr, 0));
    SOD_TRY(dt_cache_invalidate(nullptr, 0));
    return SOD_OK;
}

int sod_update_host(void* host, size_t bytes) {
    SOD_TRY(ensure_init());
    void* d = nullptr;
    SOD_TRY(find_range(host, bytes, &d));
    SOD_HIP(hipStreamSynchronize(rt().stream));
    SOD_HIP(hipMemcpy(host, d, bytes, hipMemcpyDeviceToHost));
    return SOD_OK;
}

void* sod_deviceptr(const void* host) {
    void* d = nullptr;
    if (find_range(host, 1, &d) != SOD_OK) return nullptr;
    return d;
}

}  // extern "C"
