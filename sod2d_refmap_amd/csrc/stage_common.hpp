// Pieces shared by the fused stage kernels (k_fused.hip) and their adjoint
// (k_adjoint.hip): lattice-lane helpers, stage constants, field indices and
// the lane-layout view of a prepared mesh.
#pragma once

#include "sod2d_internal.hpp"

namespace sod {

// LDS slot of lattice point (i,j,k): conflict-free for the 4-point line
// reads along each of the three directions
__device__ __forceinline__ int swz(int i, int j, int k) {
    return 16 * k + ((i + k) & 3) + 4 * ((j + k) & 3);
}

// LDS slot layouts of an element's nodal values in the stage kernel
enum { LDS_SWIZZLED = 0, LDS_LINEAR = 1 };
template <int LAY>
__device__ __forceinline__ int lds_slot(int i, int j, int k) {
    return LAY == LDS_SWIZZLED ? swz(i, j, k) : i + 4 * j + 16 * k;
}

// v of another lane by a DPP lane move (dpp_ctrl CTRL) in the rows that
// ROW_MASK selects; lanes of the other rows keep v.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_lanes(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), CTRL, ROW_MASK, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Maximum (MAX) or minimum over the 64 lanes of a wavefront, returned to
// every lane.  The lane moves are DPP operands of the VALU: two quad
// permutations ([1,0,3,2], [2,3,0,1]), the half-row and row mirrors (each
// row of 16 then holds its extremum), the broadcast of lane 15 of rows 0
// and 2 into rows 1 and 3, then of lane 31 into rows 2 and 3; lane 63 then
// holds the result, which one readlane makes uniform.  Shuffles through the
// LDS unit (ds_bpermute) would make each of the six steps a dependent LDS
// round trip.  Every lane of the wavefront must be active.  min and max are
// exact, so the order of the comparisons does not change the value.
template <bool MAX>
__device__ __forceinline__ double wave_extremum(double v) {
    auto pick = [](double a, double b) { return MAX ? fmax(a, b) : fmin(a, b); };
    v = pick(v, dpp_lanes<0xB1, 0xF>(v));    // quad_perm [1,0,3,2]
    v = pick(v, dpp_lanes<0x4E, 0xF>(v));    // quad_perm [2,3,0,1]
    v = pick(v, dpp_lanes<0x141, 0xF>(v));   // row_half_mirror
    v = pick(v, dpp_lanes<0x140, 0xF>(v));   // row_mirror
    v = pick(v, dpp_lanes<0x142, 0xA>(v));   // row_bcast:15 into rows 1 and 3
    v = pick(v, dpp_lanes<0x143, 0xC>(v));   // row_bcast:31 into rows 2 and 3
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_max(double v) { return wave_extremum<true>(v); }
__device__ __forceinline__ double wave_min(double v) { return wave_extremum<false>(v); }

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8).  Returns the logical block index that gives XCD x the
// contiguous range [x*q + min(x, r), ...) of the nb blocks (a bijection), so
// that blocks working on neighbouring elements share an L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int q = nb >> 3, r8 = nb & 7, x = b & 7;
    return x * q + min(x, r8) + (b >> 3);
}

// CFL time step candidate of one node (adapt_dt_cfl, mod_time_ops.f90:37-52,
// with mu_sgs = 0).  The reference forms, per element e,
//   cfl_c*(h_e/7)/max_a(|u_a|_inf + c_a)  and  cfl_d*(h_e/7)^2/max_a(mu_a/rho_a)
// and takes the minimum over elements.  Every rounded operation there is
// monotone (non-decreasing in h_e, non-increasing in the maxima), so that
// minimum equals, bit for bit, the minimum over working nodes a of the same
// expressions evaluated with h = min over the elements holding a of h_e and
// the node's own values.  The arithmetic below is the per-element one.
__device__ __forceinline__ double node_dt(double hmin, double u0, double u1, double u2, double c, double mu,
                                          double r, double cfl_c, int has_diff, double cfl_d) {
    double umag = fabs(u0);
    umag = fmax(umag, fabs(u1));
    umag = fmax(umag, fabs(u2));
    const double L3 = fmax(0.0, umag + c);
    const double hh = hmin / 7.0;
    double dt = cfl_c * hh / L3;
    if (has_diff) {
        const double mx = fmax(0.0, (r * 0.0 + mu) / r);
        dt = fmin(dt, cfl_d * (hh * hh) / mx);
    }
    return dt;
}

// orders LDS accesses of the lanes of one wavefront
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-stage constants; ratios of the gas constants are formed once on the
// host so that the element kernels divide only by node values.
// suth: Sutherland's law is evaluated at each stage from the stage
// temperature (flag_real_diff == 1 and flag_diff_suth == 1,
// time_integ.f90:312-316); otherwise the stages read the caller's mu_fluid
// unchanged, and the per-node viscosity input of the element kernels is
// mu_fluid itself instead of mu_factor.
struct StageParams {
    double gamma, gm1, ig1, gog1, iRgas, CpPr, Cp09, c_rho, c_ener;
    int suth;
};

inline bool sutherland_gate(const sod_numerical_params& prm) {
    return prm.flag_real_diff == 1 && prm.flag_diff_suth == 1;
}

inline StageParams make_stage_params(double gamma, double Rgas, double Cp, double Pr,
                                     const sod_numerical_params& prm) {
    const double c_rho = prm.c_rho, c_ener = prm.c_ener;
    StageParams P;
    P.suth = sutherland_gate(prm) ? 1 : 0;
    P.gamma = gamma;
    P.gm1 = gamma - 1.0;
    P.ig1 = 1.0 / (gamma - 1.0);
    P.gog1 = gamma / (gamma - 1.0);
    P.iRgas = 1.0 / Rgas;
    P.CpPr = Cp / Pr;
    P.Cp09 = Cp / 0.9;   // turbulent Prandtl number 0.9 (elem_diffu.f90)
    P.c_rho = c_rho;
    P.c_ener = c_ener;
    return P;
}

enum { P_RHO, P_U0, P_U1, P_U2, P_P, P_H, P_K, P_ETA, P_T, NPRIM };
// assembly of the stage kernel's residuals (k_stage_rhs)
enum { ASM_PULL = 0, ASM_ATOMIC = 1, ASM_COLOR = 2 };
enum { R_MASS, R_MOM0, R_MOM1, R_MOM2, R_ENER, R_ETA, NRES };

// Lane-layout view of the mesh prepared for the drop-in entry points
// (k_fused.hip): conn[e][lane] (0-based), He[e][d+3c][lane], gpv/mue/musgs
// [e][lane]; mue holds the caller's mu_e of the last call, musgs its mu_sgs
// when has_sgs (and is not read otherwise).
struct LaneMesh {
    const int* conn;
    const double *He, *gpv, *mue, *musgs;
    bool has_sgs;
    // node -> (element, lane) positions and the element-node buffer their
    // contributions are summed from
    const int *inv_off, *inv_pos;
    double* relem;
};
int prepare_lane_mesh(int nelem, int npoin, int npoin_w, const int* connec, const int* lpoin_w,
                      const double* He, const double* gpvol, const double* Ml, const double* mu_e,
                      const double* mu_sgs, LaneMesh* out);

}  // namespace sod
