// Allocation-free entry point: sod_smart_visc_spectral_noalloc, the entropy
// viscosity of sod_smart_visc_spectral (mod_entropy_viscosity.f90:14-123)
// with its reduction scratch -- the per-block partial sums and maxima of the
// entropy normalisation, their sum and the normalisation itself -- in a
// file-scope device array instead of library-allocated memory, for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops).  The kernels are the ones sod_smart_visc_spectral runs.
// Captured into a graph (the library stream, sod_set_stream, under capture),
// the call records only kernel launches.
//
// `scratch` below is a device symbol: in host code its name is a host-side
// stand-in, not a device address.  Its address on each device is looked up
// once with hipGetSymbolAddress (at sod_init, or at the first call on another
// device) and only that address is passed to kernels.  Each device holds its
// own copy until the process ends; calls that overlap on one device share it
// and race.
#include <climits>

#include "sod2d_internal.hpp"

// sized for launch_smart_visc's block partials, their sum and the normalisation
constexpr int kNoAllocScratchDoubles = 2048;   // 16 KiB per device
// File scope with external linkage: hipGetSymbolAddress finds only global
// symbols of the code object, and a `static` array is emitted as a local one
// (measured: "invalid device symbol"); nothing outside this file names it.
__device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace sod {

void launch_smart_visc(int nelem, int npoin, int npoin_w, const int* connec, const int* lpoin_w,
                       const double* Reta, const double* gpvol, double gamma_gas, const double* rho,
                       const double* u, const double* csound, const double* Tem, const double* eta,
                       const double* helem_l, const double* helem, double* mu_e, double* mue_l,
                       double* work);   // k_misc.hip
size_t smart_visc_scratch_doubles();       // k_misc.hip

namespace {
constexpr int kMaxDevices = 64;
double* g_scratch_dev[kMaxDevices] = {};
}  // namespace

// Looks up the current device's address of `scratch` (once per device).
// sod_init calls it with its device current and goes on without it when it
// fails: only the allocation-free entry points need it, and they report the
// failure.
int noalloc_setup() {
    int dev = 0;
    SOD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(SOD_ERR_HIP, "allocation-free entry points: device %d is beyond the %d supported", dev,
                    kMaxDevices);
    if (g_scratch_dev[dev] != nullptr) return SOD_OK;
    void* p = nullptr;
    // ::scratch: inside namespace sod the name is also the runtime's scratch allocator
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    if (e != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        return fail(SOD_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    }
    g_scratch_dev[dev] = static_cast<double*>(p);
    return SOD_OK;
}

static int noalloc_scratch(size_t doubles, double** out) {
    SOD_TRY(noalloc_setup());
    int dev = 0;
    SOD_HIP(hipGetDevice(&dev));
    if (doubles > (size_t)kNoAllocScratchDoubles)
        return fail(SOD_ERR_ARG, "allocation-free entry point needs %zu doubles of scratch, at most %d are built in",
                    doubles, kNoAllocScratchDoubles);
    *out = g_scratch_dev[dev];
    return SOD_OK;
}

}  // namespace sod

using namespace sod;

extern "C" int sod_smart_visc_spectral_noalloc(int nelem, int npoin, int npoin_w, const int* connec,
                                               const int* lpoin_w, const double* Reta, const double* Ngp,
                                               const double* coord, const double* dNgp, const double* gpvol,
                                               const double* wgp, double gamma_gas, const double* rho,
                                               const double* u, const double* csound, const double* Tem,
                                               const double* eta, const double* helem, const double* helem_k,
                                               const double* Ml, double* mu_e, const int* invAtoIJK,
                                               const int* gmshAtoI, const int* gmshAtoJ, const int* gmshAtoK,
                                               double* mue_l) {
    (void)Ngp;
    (void)coord;
    (void)dNgp;
    (void)Ml;
    (void)wgp;
    SOD_TRY(ensure_init());
    if (nelem < 0 || npoin < 0 || npoin_w < 0 || npoin_w > npoin)
        return fail(SOD_ERR_ARG, "smart_visc_spectral_noalloc: sizes");
    if (npoin_w <= 0 && rt().params.flag_normalise_entropy == 1)
        return fail(SOD_ERR_ARG, "smart_visc_spectral_noalloc: npoin_w must be positive");
    // Under graph capture of the library stream (sod_set_stream) the call
    // neither copies nor waits: the node-order tables must have passed an
    // earlier uncaptured check, and the return does not synchronise.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    SOD_HIP(hipStreamIsCapturing(rt().stream, &cap));
    const bool capturing = cap != hipStreamCaptureStatusNone;
    // the element kernel uses the library's GLL weights and node order; the
    // caller's tables are compared with them (without remembering the result)
    SOD_TRY(check_tables_noalloc(invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, capturing));
    const int *dc, *dl;
    const double *dR, *dgv, *drho, *du, *dcs, *dT, *deta, *dhl, *dh;
    double *dmue, *dmuel;
    SOD_TRY(resolve_t(connec, "connec", &dc));
    SOD_TRY(resolve_t(lpoin_w, "lpoin_w", &dl));
    SOD_TRY(resolve_t(Reta, "Reta", &dR));
    SOD_TRY(resolve_t(gpvol, "gpvol", &dgv));
    SOD_TRY(resolve_t(rho, "rho", &drho));
    SOD_TRY(resolve_t(u, "u", &du));
    SOD_TRY(resolve_t(csound, "csound", &dcs));
    SOD_TRY(resolve_t(Tem, "Tem", &dT));
    SOD_TRY(resolve_t(eta, "eta", &deta));
    SOD_TRY(resolve_t(helem, "helem", &dhl));
    SOD_TRY(resolve_t(helem_k, "helem_k", &dh));
    SOD_TRY(resolve_w(mu_e, "mu_e", &dmue));
    SOD_TRY(resolve_w(mue_l, "mue_l", &dmuel));
    double* work = nullptr;
    SOD_TRY(noalloc_scratch(smart_visc_scratch_doubles(), &work));
    // every scratch value is written before it is read (block partials, then
    // their sum, then the normalisation): nothing needs clearing
    launch_smart_visc(nelem, npoin, npoin_w, dc, dl, dR, dgv, gamma_gas, drho, du, dcs, dT, deta, dhl, dh, dmue,
                      dmuel, work);
    if (capturing) {
        SOD_HIP(hipGetLastError());
        return SOD_OK;
    }
    return finish_call();
}
