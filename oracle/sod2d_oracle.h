/*
 * sod2d_oracle.h -- CPU restatement of SOD2D's explicit compressible
 * spectral-element right-hand side (HEX64, p = 3, FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is what tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py compare the
 * HIP results with.  It is never
 * linked into, called by, or shipped with the product library
 * (sod2d_refmap_amd/, libsod2d_hip.so).
 *
 * Every function restates one routine of the reference and cites the
 * reference file:line it follows (paths relative to the reference root).
 * Array layouts are the reference's Fortran (column-major) layouts and every
 * index array holds the reference's 1-based values:
 *   connec(nelem,64)          -> connec[e + nelem*a]
 *   He(3,3,64,nelem)          -> He[i + 3*j + 9*g + 576*e]
 *   gpvol(1,64,nelem)         -> gpvol[g + 64*e]
 *   dlxigp_ip(64,3,4)         -> dlxigp_ip[g + 64*d + 192*ii]
 *   invAtoIJK(4,4,4)          -> invAtoIJK[i + 4*j + 16*k]
 *   u(npoin,3)                -> u[n + npoin*d]
 *   mu_e(nelem,64)            -> mu_e[e + nelem*g]
 *
 * Parity status of each function (how it is pinned) is listed in the header
 * comment of sod2d_oracle.c.
 */
#ifndef SOD2D_ORACLE_H
#define SOD2D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_PORDER 3
#define ORC_NNODE 64
#define ORC_NGAUS 64

/* Module globals of mod_numerical_params.f90 that the path reads. */
typedef struct {
    double ce;                    /* mod_numerical_params.f90:24 */
    double cmax;                  /* :25 */
    double cglob;                 /* :26 */
    double c_rho;                 /* :27 */
    double c_ener;                /* :28 */
    double Rgas_inf;              /* nscbc_Rgas_inf, :45 (set by the case) */
    int flag_normalise_entropy;   /* :23 */
    int flag_total_enthalpy;      /* :20 */
    int flag_real_diff;           /* :93 */
    int flag_diff_suth;           /* :94 */
} orc_params;

void orc_default_params(orc_params *p);

/* A1: GLL roots and weights (increasing order), rounded through FP32. */
void orc_gll_roots_weights(double roots[4], double weights[4]);

/* A1-A3: every p=3 table the path consumes. */
void orc_hex64_tables(int atoIJK[64], double xgp[64 * 3], double wgp[64],
                      double Ngp[64 * 64], double dNgp[3 * 64 * 64],
                      double dlxigp_ip[64 * 3 * 4], int invAtoIJK[64],
                      int gmshAtoI[64], int gmshAtoJ[64], int gmshAtoK[64]);

/* A4 */
void orc_elem_jacobian(int nelem, int npoin, const int *connec, const double *coord,
                       const double *dNgp, const double *wgp, double *gpvol, double *He);
/* Cofactor inverse of one 3x3 Je(i,j) (column-major), returns det. */
double orc_invert_jacobian(const double Je[9], double He[9]);
/* A6 */
void orc_lumped_mass_spectral(int nelem, int npoin, const int *connec,
                              const double *gpvol, double *Ml);
/* A7 */
void orc_char_length(int nelem, int npoin, const int *connec, const double *coord,
                     double *helem);
void orc_char_length_spectral(int nelem, int npoin, const int *connec,
                              const double *coord, double *helem_l);

/* A13 / A13b */
void orc_full_convec_ijk(int nelem, int npoin, const int *connec, const double *He,
                         const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK,
                         const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                         const double *u, const double *q, const double *rho,
                         const double *pr, const double *E, double *Rmass, double *Rmom,
                         double *Rener);
void orc_full_convec_ijk_H(int nelem, int npoin, const int *connec, const double *He,
                           const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK,
                           const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                           const double *u, const double *q, const double *rho,
                           const double *pr, const double *E, double *Rmass, double *Rmom,
                           double *Rener);
/* A10 */
void orc_generic_scalar_convec_ijk(int nelem, int npoin, const int *connec, const double *He,
                                   const double *gpvol, const double *dlxigp_ip,
                                   const int *invAtoIJK, const int *gmshAtoI,
                                   const int *gmshAtoJ, const int *gmshAtoK, const double *q,
                                   const double *eta, const double *u, double *Rconvec);
/* A12 */
void orc_full_diffusion_ijk(int nelem, int npoin, const int *connec, const double *He,
                            const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK,
                            const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                            double Cp, double Pr, const double *rho_n, const double *rho,
                            const double *u, const double *Tem, const double *mu_fluid,
                            const double *mu_e, const double *mu_sgs, const orc_params *prm,
                            double *Rmass, double *Rmom, double *Rener);
/* A11 */
void orc_sutherland_viscosity(int npoin, const double *Tem, const double *mu_factor,
                              double *mu_fluid);
/* A14 */
void orc_lumped_solver_scal(int npoin, int npoin_w, const int *lpoin_w, const double *Ml,
                            double *R);
void orc_lumped_solver_vect(int npoin, int npoin_w, const int *lpoin_w, const double *Ml,
                            double *R);
/* A17 */
void orc_smart_visc_spectral(int nelem, int npoin, int npoin_w, const int *connec,
                             const int *lpoin_w, const double *Reta, const double *gpvol,
                             const double *wgp, double gamma_gas, const double *rho,
                             const double *u, const double *csound, const double *Tem,
                             const double *eta, const double *helem_l, const double *helem,
                             const int *invAtoIJK, const orc_params *prm, double *mu_e,
                             double *mue_l);
/* A18 */
double orc_adapt_dt_cfl(int nelem, int npoin, const int *connec, const double *helem,
                        const double *u, const double *csound, double cfl_conv,
                        double cfl_diff, const double *mu_fluid, const double *mu_sgs,
                        const double *rho);
/* A21 */
double orc_volAvg_EK(int nelem, int npoin, const int *connec, const double *gpvol,
                     const double *Ngp, double rho0, const double *rho, const double *u);

/* A8-A17: one call of rk_4_main for a periodic case (noBoundaries, no LES,
 * no buffer, no source, no wall model).  State arrays are the time level
 * pos=2 of the reference (rho(:,2), u(:,:,2), ...), updated in place. */
void orc_rk_4_main(int nelem, int npoin, int npoin_w, const int *connec, const int *lpoin_w,
                   const double *He, const double *gpvol, const double *dlxigp_ip,
                   const double *wgp, const int *invAtoIJK, const int *gmshAtoI,
                   const int *gmshAtoJ, const int *gmshAtoK, const double *Ml,
                   const double *helem, const double *helem_l, double dt, double Rgas,
                   double gamma_gas, double Cp, double Prt, const double *mu_factor,
                   const orc_params *prm, double *rho, double *u, double *q, double *pr,
                   double *E, double *Tem, double *e_int, double *eta, double *csound,
                   double *machno, double *mu_fluid, double *mu_e, double *mu_sgs,
                   double *mue_l);

#ifdef __cplusplus
}
#endif
#endif
