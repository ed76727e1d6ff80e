// rk_4_main (time_integ.f90:137-508), periodic explicit branch: the RK4
// stage loop, the stage equation of state, residual accumulation, the final
// update, the final equation of state, the entropy residual, Sutherland
// viscosity and the entropy viscosity for the next step.
#include "sod2d_internal.hpp"

namespace sod {

void launch_full_convec(bool total_h, int nelem, int npoin, const int* connec, const double* He,
                        const double* gpvol, const double* u, const double* q, const double* rho,
                        const double* pr, const double* E, double* Rmass, double* Rmom,
                        double* Rener);
void launch_scalar_convec(int nelem, int npoin, const int* connec, const double* He,
                          const double* gpvol, const double* q, const double* eta, const double* u,
                          double* R);
void launch_full_diffusion(int nelem, int npoin, const int* connec, const double* He,
                           const double* gpvol, double Cp, double Pr, const double* rho_n,
                           const double* rho, const double* u, const double* Tem,
                           const double* mu_fluid, const double* mu_e, const double* mu_sgs,
                           double* Rmass, double* Rmom, double* Rener);
void launch_sutherland(int npoin, const double* Tem, const double* mu_factor, double* mu_fluid);
void launch_lumped(int npoin, int npoin_w, int ncomp, const int* lpoin_w, const double* Ml,
                   double* R);
void launch_smart_visc(int nelem, int npoin, int npoin_w, const int* connec, const int* lpoin_w,
                       const double* Reta, const double* gpvol, double gamma_gas, const double* rho,
                       const double* u, const double* csound, const double* Tem, const double* eta,
                       const double* helem_l, const double* helem, double* mu_e, double* mue_l,
                       double* work);
int red_blocks();

namespace {

constexpr int BLK = 256;
// RK4 tableau, time_integ.f90:81-84
__constant__ double c_a[4] = {0.0, 0.5, 0.5, 1.0};
__constant__ double c_b[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

// Work arrays of rk_4_main (time_integ.f90:18-23), npoin each.
enum Field {
    F_RMASS, F_RENER, F_RETA, F_RMOM0, F_RMOM1, F_RMOM2,
    F_AUX_RHO, F_AUX_PR, F_AUX_E, F_AUX_TEM, F_AUX_EINT, F_AUX_ETA, F_AUX_H,
    F_AUX_U0, F_AUX_U1, F_AUX_U2, F_AUX_Q0, F_AUX_Q1, F_AUX_Q2,
    F_RMASS_SUM, F_RENER_SUM, F_RETA_SUM, F_RDIFF_MASS, F_RDIFF_ENER,
    F_RMOM_SUM0, F_RMOM_SUM1, F_RMOM_SUM2, F_RDIFF_MOM0, F_RDIFF_MOM1, F_RDIFF_MOM2,
    F_FETA0, F_FETA1, F_FETA2, F_COUNT
};

// aux = y_n - dt a_i R   (time_integ.f90:239-248)
__global__ __launch_bounds__(BLK) void k_stage_state(int npoin, int istep, double dt,
                                                     const double* __restrict__ rho,
                                                     const double* __restrict__ E,
                                                     const double* __restrict__ q,
                                                     double* __restrict__ W) {
    const int n = blockIdx.x * BLK + threadIdx.x;
    if (n >= npoin) return;
    const size_t np = (size_t)npoin;
    const double f = dt * c_a[istep];
    W[F_AUX_RHO * np + n] = rho[n] - f * W[F_RMASS * np + n];
    W[F_AUX_E * np + n] = E[n] - f * W[F_RENER * np + n];
#pragma unroll
    for (int d = 0; d < 3; ++d)
        W[(F_AUX_Q0 + d) * np + n] = q[n + np * d] - f * W[(F_RMOM0 + d) * np + n];
}

// stage EOS on working nodes (time_integ.f90:266-285)
__global__ __launch_bounds__(BLK) void k_stage_eos(int npoin, int npoin_w, const int* __restrict__ lpoin_w,
                                                   double gamma_gas, double Rgas,
                                                   double* __restrict__ W) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= npoin_w) return;
    const size_t np = (size_t)npoin;
    const int n = lpoin_w[i] - 1;
    const double r = W[F_AUX_RHO * np + n];
    double umag = 0.0, uu[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        uu[d] = W[(F_AUX_Q0 + d) * np + n] / r;
        W[(F_AUX_U0 + d) * np + n] = uu[d];
        umag = umag + (uu[d] * uu[d]);
    }
    const double ei = (W[F_AUX_E * np + n] / r) - 0.5 * umag;
    const double p = r * (gamma_gas - 1.0) * ei;
    W[F_AUX_EINT * np + n] = ei;
    W[F_AUX_PR * np + n] = p;
    W[F_AUX_H * np + n] = (gamma_gas / (gamma_gas - 1.0)) * p / r;
    W[F_AUX_TEM * np + n] = p / (r * Rgas);
    const double eta = (r / (gamma_gas - 1.0)) * log(p / pow(r, gamma_gas));
    W[F_AUX_ETA * np + n] = eta;
#pragma unroll
    for (int d = 0; d < 3; ++d) W[(F_FETA0 + d) * np + n] = uu[d] * eta;
}

// R += Rdiff (time_integ.f90:369-373)
__global__ __launch_bounds__(BLK) void k_add_diff(int npoin, double* __restrict__ W) {
    const int n = blockIdx.x * BLK + threadIdx.x;
    if (n >= npoin) return;
    const size_t np = (size_t)npoin;
    W[F_RMASS * np + n] = W[F_RMASS * np + n] + W[F_RDIFF_MASS * np + n];
    W[F_RENER * np + n] = W[F_RENER * np + n] + W[F_RDIFF_ENER * np + n];
#pragma unroll
    for (int d = 0; d < 3; ++d)
        W[(F_RMOM0 + d) * np + n] = W[(F_RMOM0 + d) * np + n] + W[(F_RDIFF_MOM0 + d) * np + n];
}

// R_sum += b_i R (time_integ.f90:399-409)
__global__ __launch_bounds__(BLK) void k_accumulate(int npoin, int istep, double* __restrict__ W) {
    const int n = blockIdx.x * BLK + threadIdx.x;
    if (n >= npoin) return;
    const size_t np = (size_t)npoin;
    const double b = c_b[istep];
    W[F_RMASS_SUM * np + n] = W[F_RMASS_SUM * np + n] + b * W[F_RMASS * np + n];
    W[F_RENER_SUM * np + n] = W[F_RENER_SUM * np + n] + b * W[F_RENER * np + n];
    W[F_RETA_SUM * np + n] = W[F_RETA_SUM * np + n] + b * W[F_RETA * np + n];
#pragma unroll
    for (int d = 0; d < 3; ++d)
        W[(F_RMOM_SUM0 + d) * np + n] = W[(F_RMOM_SUM0 + d) * np + n] + b * W[(F_RMOM0 + d) * np + n];
}

// y = y_n - dt R_sum (time_integ.f90:417-426)
__global__ __launch_bounds__(BLK) void k_rk_update(int npoin, double dt, double* __restrict__ rho,
                                                   double* __restrict__ E, double* __restrict__ q,
                                                   const double* __restrict__ W) {
    const int n = blockIdx.x * BLK + threadIdx.x;
    if (n >= npoin) return;
    const size_t np = (size_t)npoin;
    rho[n] = rho[n] - dt * W[F_RMASS_SUM * np + n];
    E[n] = E[n] - dt * W[F_RENER_SUM * np + n];
#pragma unroll
    for (int d = 0; d < 3; ++d) q[n + np * d] = q[n + np * d] - dt * W[(F_RMOM_SUM0 + d) * np + n];
}

// final EOS + entropy residual (time_integ.f90:445-471)
__global__ __launch_bounds__(BLK) void k_final_eos(int npoin, int npoin_w, const int* __restrict__ lpoin_w,
                                                   double gamma_gas, double Rgas,
                                                   const double* __restrict__ rho,
                                                   const double* __restrict__ q,
                                                   const double* __restrict__ E, double* __restrict__ u,
                                                   double* __restrict__ e_int, double* __restrict__ pr,
                                                   double* __restrict__ csound,
                                                   double* __restrict__ machno,
                                                   double* __restrict__ Tem, double* __restrict__ eta,
                                                   double* __restrict__ W) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= npoin_w) return;
    const size_t np = (size_t)npoin;
    const int n = lpoin_w[i] - 1;
    const double r = rho[n];
    double umag = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double ud = q[n + np * d] / r;
        u[n + np * d] = ud;
        umag = umag + ud * ud;
    }
    const double ei = (E[n] / r) - 0.5 * umag;
    const double p = r * (gamma_gas - 1.0) * ei;
    e_int[n] = ei;
    pr[n] = p;
    const double c = sqrt(gamma_gas * p / r);
    csound[n] = c;
    machno[n] = sqrt(umag) / c;
    Tem[n] = p / (r * Rgas);
    eta[n] = (r / (gamma_gas - 1.0)) * log(p / pow(r, gamma_gas));
    W[F_RETA * np + n] = -W[F_RETA_SUM * np + n];
}

}  // namespace

// Runs one RK4 step with the reference's kernel sequence on device pointers.
// level-2 state pointers; W = F_COUNT*npoin work doubles; vwork = smart_visc
// reduction scratch.
int rk4_step_reference_sequence(int nelem, int npoin, int npoin_w, const int* connec,
                                const int* lpoin_w, const double* He, const double* gpvol,
                                const double* Ml, const double* helem, const double* helem_l,
                                double dt, double Rgas, double gamma_gas, double Cp, double Prt,
                                const double* mu_factor, double* rho, double* u, double* q,
                                double* pr, double* E, double* Tem, double* e_int, double* eta,
                                double* csound, double* machno, double* mu_fluid, double* mu_e,
                                const double* mu_sgs, double* mue_l, double* W, double* vwork) {
    SOD_PROF("rk4_step");
    hipStream_t s = rt().stream;
    const sod_numerical_params& p = rt().params;
    const size_t np = (size_t)npoin;
    const dim3 gN(grid_for(npoin, BLK)), gW(grid_for(npoin_w, BLK)), b(BLK);
    // time_integ.f90:204-225: every work array starts at zero
    SOD_HIP(hipMemsetAsync(W, 0, sizeof(double) * np * F_COUNT, s));
    for (int istep = 0; istep < 4; ++istep) {
        hipLaunchKernelGGL(k_stage_state, gN, b, 0, s, npoin, istep, dt, rho, E, q, W);
        hipLaunchKernelGGL(k_stage_eos, gW, b, 0, s, npoin, npoin_w, lpoin_w, gamma_gas, Rgas, W);
        // generic_scalar_convec_ijk zeroes its output (elem_convec.f90:484-486)
        SOD_HIP(hipMemsetAsync(W + F_RETA * np, 0, sizeof(double) * np, s));
        launch_scalar_convec(nelem, npoin, connec, He, gpvol, W + F_FETA0 * np, W + F_AUX_ETA * np,
                             W + F_AUX_U0 * np, W + F_RETA * np);
        launch_lumped(npoin, npoin_w, 1, lpoin_w, Ml, W + F_RETA * np);
        if (p.flag_real_diff == 1 && p.flag_diff_suth == 1)
            launch_sutherland(npoin, W + F_AUX_TEM * np, mu_factor, mu_fluid);
        SOD_HIP(hipMemsetAsync(W + F_RDIFF_MASS * np, 0, sizeof(double) * np, s));
        SOD_HIP(hipMemsetAsync(W + F_RDIFF_ENER * np, 0, sizeof(double) * np, s));
        SOD_HIP(hipMemsetAsync(W + F_RDIFF_MOM0 * np, 0, sizeof(double) * np * 3, s));
        launch_full_diffusion(nelem, npoin, connec, He, gpvol, Cp, Prt, W + F_AUX_RHO * np,
                              W + F_AUX_RHO * np, W + F_AUX_U0 * np, W + F_AUX_TEM * np, mu_fluid,
                              mu_e, mu_sgs, W + F_RDIFF_MASS * np, W + F_RDIFF_MOM0 * np,
                              W + F_RDIFF_ENER * np);
        SOD_HIP(hipMemsetAsync(W + F_RMASS * np, 0, sizeof(double) * np * 2, s));
        SOD_HIP(hipMemsetAsync(W + F_RMOM0 * np, 0, sizeof(double) * np * 3, s));
        if (p.flag_total_enthalpy)
            launch_full_convec(true, nelem, npoin, connec, He, gpvol, W + F_AUX_U0 * np,
                               W + F_AUX_Q0 * np, W + F_AUX_RHO * np, W + F_AUX_PR * np,
                               W + F_AUX_E * np, W + F_RMASS * np, W + F_RMOM0 * np,
                               W + F_RENER * np);
        else
            launch_full_convec(false, nelem, npoin, connec, He, gpvol, W + F_AUX_U0 * np,
                               W + F_AUX_Q0 * np, W + F_AUX_RHO * np, W + F_AUX_PR * np,
                               W + F_AUX_H * np, W + F_RMASS * np, W + F_RMOM0 * np,
                               W + F_RENER * np);
        hipLaunchKernelGGL(k_add_diff, gN, b, 0, s, npoin, W);
        launch_lumped(npoin, npoin_w, 1, lpoin_w, Ml, W + F_RMASS * np);
        launch_lumped(npoin, npoin_w, 1, lpoin_w, Ml, W + F_RENER * np);
        launch_lumped(npoin, npoin_w, 3, lpoin_w, Ml, W + F_RMOM0 * np);
        hipLaunchKernelGGL(k_accumulate, gN, b, 0, s, npoin, istep, W);
    }
    hipLaunchKernelGGL(k_rk_update, gN, b, 0, s, npoin, dt, rho, E, q, W);
    hipLaunchKernelGGL(k_final_eos, gW, b, 0, s, npoin, npoin_w, lpoin_w, gamma_gas, Rgas, rho, q, E,
                       u, e_int, pr, csound, machno, Tem, eta, W);
    if (p.flag_real_diff == 1 && p.flag_diff_suth == 1)
        launch_sutherland(npoin, Tem, mu_factor, mu_fluid);
    launch_smart_visc(nelem, npoin, npoin_w, connec, lpoin_w, W + F_RETA * np, gpvol, gamma_gas, rho, u,
                      csound, Tem, eta, helem_l, helem, mu_e, mue_l, vwork);
    SOD_HIP(hipGetLastError());
    return SOD_OK;
}

int rk4_work_fields() { return F_COUNT; }

int rk4_step_fused(int nelem, int npoin, int npoin_w, const int* connec, const int* lpoin_w,
                   const double* He, const double* gpvol, const double* Ml, const double* helem,
                   const double* helem_l, double dt, double Rgas, double gamma_gas, double Cp, double Prt,
                   const double* mu_factor, double* rho, double* u, double* q, double* pr, double* E,
                   double* Tem, double* e_int, double* eta, double* csound, double* machno,
                   double* mu_fluid, double* mu_e, const double* mu_sgs, double* mue_l);

// 0: fused stage kernels (default); 1: the reference's kernel sequence
static int g_rk4_path = 0;

}  // namespace sod

using namespace sod;

extern "C" int sod_set_rk4_path(int path) {
    if (path != 0 && path != 1) return fail(SOD_ERR_ARG, "sod_set_rk4_path: 0 (fused) or 1 (reference sequence)");
    g_rk4_path = path;
    return SOD_OK;
}

extern "C" int sod_rk_4_main(
    int noBoundaries, int isWallModelOn, int nelem, int nboun, int npoin, int npoin_w,
    int numBoundsWM, const int* point2elem, const int* lnbn_nodes, const double* dlxigp_ip,
    const double* xgp, const int* atoIJK, const int* invAtoIJK, const int* gmshAtoI,
    const int* gmshAtoJ, const int* gmshAtoK, int ppow, const int* connec, const double* Ngp,
    const double* dNgp, const double* coord, const double* wgp, const double* He, const double* Ml,
    const double* gpvol, double dt, const double* helem, const double* helem_l, double Rgas,
    double gamma_gas, double Cp, double Prt, double* rho, double* u, double* q, double* pr,
    double* E, double* Tem, double* csound, double* machno, double* e_int, double* eta,
    double* mu_e, double* mu_sgs, double* kres, double* etot, double* au, double* ax1, double* ax2,
    double* ax3, const int* lpoin_w, double* mu_fluid, const double* mu_factor, double* mue_l) {
    (void)nboun; (void)point2elem; (void)lnbn_nodes; (void)xgp; (void)atoIJK; (void)ppow;
    (void)Ngp; (void)dNgp; (void)coord; (void)wgp; (void)kres; (void)etot; (void)au;
    (void)ax1; (void)ax2; (void)ax3;
    SOD_TRY(ensure_init());
    if (!noBoundaries)
        return fail(SOD_ERR_UNSUPPORTED, "rk_4_main: boundary conditions are outside this library");
    if (isWallModelOn && numBoundsWM != 0)
        return fail(SOD_ERR_UNSUPPORTED, "rk_4_main: wall models are outside this library");
    if (nelem < 0 || npoin < 0 || npoin_w < 0 || npoin_w > npoin)
        return fail(SOD_ERR_ARG, "rk_4_main: inconsistent sizes");
    SOD_TRY(check_tables(dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK));
    const int *dc, *dl;
    const double *dHe, *dMl, *dgv, *dh, *dhl, *dmf;
    double *drho, *du, *dq, *dpr, *dE, *dT, *dcs, *dma, *dei, *deta, *dmue, *dsgs, *dmuf, *dmuel;
    SOD_TRY(resolve_t(connec, "connec", &dc));
    SOD_TRY(resolve_t(lpoin_w, "lpoin_w", &dl));
    SOD_TRY(resolve_t(He, "He", &dHe));
    SOD_TRY(resolve_t(Ml, "Ml", &dMl));
    SOD_TRY(resolve_t(gpvol, "gpvol", &dgv));
    SOD_TRY(resolve_t(helem, "helem", &dh));
    SOD_TRY(resolve_t(helem_l, "helem_l", &dhl));
    SOD_TRY(resolve_t(mu_factor, "mu_factor", &dmf));
    SOD_TRY(resolve_w(rho, "rho", &drho));
    SOD_TRY(resolve_w(u, "u", &du));
    SOD_TRY(resolve_w(q, "q", &dq));
    SOD_TRY(resolve_w(pr, "pr", &dpr));
    SOD_TRY(resolve_w(E, "E", &dE));
    SOD_TRY(resolve_w(Tem, "Tem", &dT));
    SOD_TRY(resolve_w(csound, "csound", &dcs));
    SOD_TRY(resolve_w(machno, "machno", &dma));
    SOD_TRY(resolve_w(e_int, "e_int", &dei));
    SOD_TRY(resolve_w(eta, "eta", &deta));
    SOD_TRY(resolve_w(mu_e, "mu_e", &dmue));
    SOD_TRY(resolve_w(mu_sgs, "mu_sgs", &dsgs));
    SOD_TRY(resolve_w(mu_fluid, "mu_fluid", &dmuf));
    SOD_TRY(resolve_w(mue_l, "mue_l", &dmuel));
    const size_t np = (size_t)npoin;
    // level pos=2 of each time-levelled array (time_integ.f90:200)
    double* rho2 = drho + np;
    double* u2 = du + 3 * np;
    double* q2 = dq + 3 * np;
    double* pr2 = dpr + np;
    double* E2 = dE + np;
    double* T2 = dT + np;
    double* ei2 = dei + np;
    double* eta2 = deta + np;
    if (g_rk4_path == 0) {
        SOD_TRY(rk4_step_fused(nelem, npoin, npoin_w, dc, dl, dHe, dgv, dMl, dh, dhl, dt, Rgas, gamma_gas, Cp, Prt,
                               dmf, rho2, u2, q2, pr2, E2, T2, ei2, eta2, dcs, dma, dmuf, dmue, dsgs, dmuel));
        return finish_call();
    }
    void *W = nullptr, *vw = nullptr;
    SOD_TRY(scratch(sizeof(double) * np * rk4_work_fields(), &W, 3));
    SOD_TRY(scratch(sizeof(double) * (red_blocks() + 8), &vw, 4));
    SOD_TRY(rk4_step_reference_sequence(nelem, npoin, npoin_w, dc, dl, dHe, dgv, dMl, dh, dhl, dt, Rgas,
                                        gamma_gas, Cp, Prt, dmf, rho2, u2, q2, pr2, E2, T2, ei2, eta2,
                                        dcs, dma, dmuf, dmue, dsgs, dmuel, static_cast<double*>(W),
                                        static_cast<double*>(vw)));
    return finish_call();
}
