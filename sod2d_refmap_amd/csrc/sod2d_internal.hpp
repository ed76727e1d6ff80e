// Internal declarations shared by the translation units of libsod2d_hip.so.
//
// Conventions (identical to the reference's Fortran data model):
//   * every array argument is column-major with the reference's extents;
//   * index arrays (connec, lpoin_w, invAtoIJK, gmshAto*) hold 1-based values;
//   * p = 3, nnode = ngaus = 64, ndime = 3 are compile-time constants.
// Element-local work is done in "lattice lane" order: lane l = i + 4*j + 16*k
// where (i,j,k) are the reference's sod2d 1-D indices minus one (0 = xi=-1,
// 1 = xi=+1, 2,3 = interior nodes in increasing order), so lane l holds the
// gmsh-ordered local node a = invAtoIJK(i+1,j+1,k+1).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "sod2d_hip.h"

namespace sod {

constexpr int kP = 3;
constexpr int kN1 = 4;     // nodes per direction
constexpr int kNN = 64;    // nnode = ngaus
constexpr int kND = 3;     // ndime

// ----------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);   // sets message, returns code

#define SOD_HIP(call)                                                              \
    do {                                                                           \
        hipError_t _e = (call);                                                    \
        if (_e != hipSuccess)                                                      \
            return ::sod::fail(SOD_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call, \
                               hipGetErrorString(_e));                             \
    } while (0)

#define SOD_TRY(expr)              \
    do {                           \
        int _rc = (expr);          \
        if (_rc != SOD_OK) return _rc; \
    } while (0)

// ------------------------------------------------------------ host tables
struct HostTables {
    int atoIJK[64];
    double xgp[64 * 3];
    double wgp[64];
    double Ngp[64 * 64];
    double dNgp[3 * 64 * 64];
    double dlxigp_ip[64 * 3 * 4];
    int invAtoIJK[64];
    int gmshAtoI[64], gmshAtoJ[64], gmshAtoK[64];
    double roots[4], weights[4];   // increasing order, FP32-rounded
    double D[4][4];                // D[m][ii] = l'_ii(xi_m), sod2d order
    int lane2node[64];             // lattice lane -> gmsh local node (0-based)
    int node2lane[64];
};
const HostTables& host_tables();

// Device-side constant tables (one copy per device, uploaded at init).
struct DevTables {
    double D[4][4];          // derivative matrix, sod2d order
    double wgp[64];          // by gmsh node
    double w1d[4];           // sod2d order
    int lane2node[64];
    int node2lane[64];
};
// Without relocatable device code every translation unit holds its own copy
// of the constant tables; each copy registers an uploader that sod_init()
// runs once per device.
static __constant__ DevTables c_tab;
using TableUploader = int (*)(const DevTables&);
using TableDumper = int (*)(DevTables* dev_out, hipStream_t s);
int register_table_uploader(TableUploader fn);
int register_table_dumper(TableDumper fn);
namespace {
inline int upload_tu_tables(const DevTables& d) {
    return hipMemcpyToSymbol(HIP_SYMBOL(c_tab), &d, sizeof(DevTables)) == hipSuccess ? 0 : -2;
}
// copies this translation unit's device copy back out (init-time self-check)
__global__ void k_dump_tu_tables(DevTables* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = c_tab;
}
inline int dump_tu_tables(DevTables* out, hipStream_t s) {
    hipLaunchKernelGGL(k_dump_tu_tables, dim3(1), dim3(64), 0, s, out);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}
[[maybe_unused]] const int tu_tables_registered = register_table_uploader(upload_tu_tables);
[[maybe_unused]] const int tu_tables_dumper = register_table_dumper(dump_tu_tables);
}  // namespace

// ------------------------------------------------------------- runtime
struct Runtime {
    bool initialised = false;
    int device = 0;
    hipStream_t stream = nullptr;
    bool sync_on_return = true;
    sod_numerical_params params{};
    int autotune = 0;   // stage configuration choice (sod_set_autotune)
};
Runtime& rt();
int ensure_init();
int finish_call();   // synchronises if sync_on_return
// Resolves a caller pointer (device pointer or registered host pointer) to a
// device address.  `what` names the argument for the error message.
int resolve(const void* p, const char* what, const void** out);
template <class T>
inline int resolve_t(const T* p, const char* what, const T** out) {
    const void* o = nullptr;
    int rc = resolve(p, what, &o);
    *out = static_cast<const T*>(o);
    return rc;
}
template <class T>
inline int resolve_w(T* p, const char* what, T** out) {
    const void* o = nullptr;
    int rc = resolve(p, what, &o);
    *out = const_cast<T*>(static_cast<const T*>(o));
    return rc;
}
// Checks caller-supplied p=3 tables against the library's (once per pointer).
int check_tables(const double* dlxigp_ip, const int* invAtoIJK, const int* gmshAtoI,
                 const int* gmshAtoJ, const int* gmshAtoK);
// The node-order tables checked without remembering the result (no host
// allocation; pointers already checked by check_tables are not read again).
// With `capturing` (the library stream is under graph capture) a table not
// checked before is refused instead of copied to the host.
int check_tables_noalloc(const int* invAtoIJK, const int* gmshAtoI, const int* gmshAtoJ, const int* gmshAtoK,
                         bool capturing);

// Device scratch owned by the runtime, grown on demand.
int scratch(size_t bytes, void** out, int slot);

// Named timing ranges (HIP events on the library stream), enabled by
// sod_profile_enable(1).
int prof_id(const char* name);
bool prof_set(bool on);   // enables/disables recording, returns the previous state
struct ProfScope {
    int id, rec;
    explicit ProfScope(int id);
    ~ProfScope();
};
#define SOD_PROF(name)                      \
    static const int _prof_id = prof_id(name); \
    ProfScope _prof_scope(_prof_id)

// ----------------------------------------------------------- device math
__device__ __forceinline__ double sutherland(double mu_factor, double T) {
    // mod_fluid_viscosity.f90:35-39
    return mu_factor * 0.000001458 * pow(T, 1.5) / (T + 110.4);
}

inline int grid_for(long n, int block) { return (int)((n + block - 1) / block); }

}  // namespace sod
