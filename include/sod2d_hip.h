/*
 * sod2d_hip.h -- C ABI of libsod2d_hip.so, the MI355X (gfx950) implementation
 * of SOD2D's explicit compressible spectral-element RHS + RK4 step
 * (HEX64, p = 3, FP64, one partition on one GPU).
 *
 * Every hot-path entry point replaces one Fortran routine of the reference
 * and keeps its argument list, order and meaning (reference file:line cited
 * per declaration; paths relative to the reference root).  Conventions:
 *   - arrays are the reference's column-major arrays with the same extents;
 *     index arrays hold the reference's 1-based values;
 *   - array arguments are device pointers, or host pointers registered with
 *     sod_enter_data() (the OpenACC "enter data"/"present" model of the
 *     reference, time_integ.f90:37-68, CFDSolverBase.f90:827-842);
 *   - scalars are passed by value (the Fortran interface declares VALUE);
 *     Fortran LOGICAL arguments are int (0 = .false.);
 *   - compile-time constants porder=3, nnode=ngaus=64, ndime=3 are not
 *     arguments (mod_constants.f90:12-20); sod_init() checks porder;
 *   - intent(out) arrays are fully overwritten, as in the reference;
 *   - return 0 on success, a negative SOD_ERR_* code otherwise, with the
 *     message in sod_last_error() (the reference has no error returns: its
 *     routines stop the program; the Fortran shim turns a non-zero return
 *     into ERROR STOP);
 *   - calls are synchronous on return unless sod_set_sync(0) was called;
 *     work is queued on the stream given to sod_set_stream() (default: the
 *     null stream of the current device).
 */
#ifndef SOD2D_HIP_H
#define SOD2D_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SOD_OK 0
#define SOD_ERR_ARG (-1)         /* invalid argument or unsupported value */
#define SOD_ERR_HIP (-2)         /* HIP runtime error */
#define SOD_ERR_NOT_PRESENT (-3) /* host pointer not registered with sod_enter_data */
#define SOD_ERR_PORDER (-4)      /* porder != 3 */
#define SOD_ERR_UNSUPPORTED (-5) /* feature outside the periodic explicit path */
#define SOD_ERR_NOT_INIT (-6)

/* Module globals of mod_numerical_params.f90 read by the path. */
typedef struct {
    double ce;                  /* mod_numerical_params.f90:24, default 0.1 */
    double cmax;                /* :25, default 0.5 */
    double cglob;               /* :26, default 1.0 */
    double c_rho;               /* :27, default 1.0 */
    double c_ener;              /* :28, default 1.0 */
    double Rgas_inf;            /* nscbc_Rgas_inf, :45 (the case sets it, TGVSolver.f90:104) */
    int flag_normalise_entropy; /* :23, default 1 */
    int flag_total_enthalpy;    /* :20, default 0 */
    int flag_rk_order;          /* :8, only 4 is supported */
    int flag_les;               /* :13, only 0 is supported */
    int flag_real_diff;         /* :93, default 1 */
    int flag_diff_suth;         /* :94, default 1 */
} sod_numerical_params;

/* ----------------------------------------------------------- runtime ---- */

/* No counterpart (the reference initialises OpenACC implicitly).  Selects
 * the current HIP device and checks porder == 3 (mod_hdf5.f90:2781-2784 stops
 * on a porder mismatch). */
int sod_init(int porder);
int sod_finalize(void);
const char *sod_last_error(void);
int sod_set_stream(void *hip_stream);
int sod_set_sync(int sync_on_return);
int sod_synchronize(void);
/* mod_numerical_params.f90:1-100 module variables. */
int sod_set_numerical_params(const sod_numerical_params *p);
int sod_get_numerical_params(sod_numerical_params *p);
void sod_default_numerical_params(sod_numerical_params *p);

/* OpenACC data directives of the reference: enter data create/copyin
 * (time_integ.f90:37-68), update device/host (CFDSolverBase.f90:1182-1226),
 * exit data delete (time_integ.f90:97-135).  `host` may be any address inside
 * a registered block for update/deviceptr. */
int sod_enter_data(const void *host, size_t bytes, int copyin);
int sod_exit_data(const void *host, int copyout);
int sod_update_device(const void *host, size_t bytes);
int sod_update_host(void *host, size_t bytes);
void *sod_deviceptr(const void *host);

/* No reference counterpart (the reference re-reads every mesh array on every
 * call).  The drop-in entry points keep derived copies of the static mesh
 * arrays they receive (connec, lpoin_w, He, gpvol, Ml in lane layout with
 * an inverse connectivity, and a per-node minimum of helem), keyed on the
 * arrays' device addresses.  A caller that rewrites a mesh array in place at
 * the same address calls this before the next call: every derived copy is
 * dropped and rebuilt on first use.  sod_update_device and sod_exit_data do
 * it by themselves for the mesh arrays inside the block they touch.  Without
 * it the library still compares a sampled fingerprint of the arrays (4096
 * sampled elements and working nodes) and a hash of all of connec and helem
 * (CFL step) on every call, which catches most but not every in-place
 * change of He or gpvol.  Time loops (sod_time_loop_*) hold their own copy
 * for their lifetime and are not affected. */
int sod_invalidate_mesh_cache(void);

/* Named kernel timing with HIP events on the library stream; the
 * counterpart of the reference's NVTX ranges (mod_nvtx.f90:50-101).
 * sod_profile_get returns the summed duration and count of the ranges
 * recorded under `name` since the last reset. */
int sod_profile_enable(int on);
int sod_profile_reset(void);
int sod_profile_get(const char *name, double *total_ms, int *count);

/* ------------------------------------------------- init-time tables ---- */

/* GaussLobattoLegendre_hex (quadrature_rules.f90:51-78), hex_highorder /
 * TripleTensorProduct (elem_hex.f90:22-43, mod_maths.f90:258-298),
 * atoIJKInverse (mod_geom.f90:123-144) with a2ijk = gmsh2ijk
 * (mod_meshConversorTool.f90:150-152, mod_ijk_indices.f90:61-114).
 * Host-only: needs no GPU.  Outputs are host arrays:
 * atoIJK(64), xgp(64,3), wgp(64), Ngp(64,64), dNgp(3,64,64),
 * dlxigp_ip(64,3,4), invAtoIJK(4,4,4), gmshAtoI/J/K(64). */
int sod_hex64_tables(int *atoIJK, double *xgp, double *wgp, double *Ngp, double *dNgp,
                     double *dlxigp_ip, int *invAtoIJK, int *gmshAtoI, int *gmshAtoJ,
                     int *gmshAtoK);

/* Synthetic input, host-only: the arrays the reference's mesh conversion
 * tool writes for a structured periodic box of nx*ny*nz HEX64 elements
 * (reorder_nodes_in_mshRank, mod_meshConversorTool.f90:2864-3076;
 * create_working_lists_parallel, :3944-4035; GLL node placement, :3921-3942).
 * elem_order (optional, int64[nelem]) gives the lexicographic element of
 * each element id.  Outputs: connec_orig(nelem,64), connec(nelem,64)
 * (= connecParWork), lpoin_w(27*nelem), coord(npoin,3), and optionally
 * node_lattice[3*npoin] (lattice indices, row-major). */
int sod_structured_box_mesh(int nx, int ny, int nz, double Lx, double Ly, double Lz,
                            double curved_amp, const long long *elem_order, int *connec_orig,
                            int *connec, int *lpoin_w, double *coord, int *node_lattice);

/* elem_jacobian, jacobian_oper.f90:9-138 */
int sod_elem_jacobian(int nelem, int npoin, const int *connec, const double *coord,
                      const double *dNgp, const double *wgp, double *gpvol, double *He);
/* lumped_mass_spectral, mass_matrix.f90:11-42 */
int sod_lumped_mass_spectral(int nelem, int npoin, const int *connec, const double *gpvol,
                             double *Ml);
/* char_length over every element (CFDSolverBase.f90:791-794 calling
 * mod_geom.f90:13-37) -> helem(nelem) */
int sod_char_length(int nelem, int npoin, const int *connec, const double *coord,
                    double *helem);
/* char_length_spectral over every element (CFDSolverBase.f90:1378-1380
 * calling mod_geom.f90:39-66) -> helem_l(nelem,64) */
int sod_char_length_spectral(int nelem, int npoin, const int *connec, const double *coord,
                             double *helem_l);

/* --------------------------------------------------------- hot path ---- */

/* full_convec_ijk, elem_convec.f90:211-224 (Ngp, dNgp and xgp are accepted
 * and unused, as in the reference). */
int sod_full_convec_ijk(int nelem, int npoin, const int *connec, const double *Ngp,
                        const double *dNgp, const double *He, const double *gpvol,
                        const double *dlxigp_ip, const double *xgp, const int *invAtoIJK,
                        const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                        const double *u, const double *q, const double *rho, const double *pr,
                        const double *E, double *Rmass, double *Rmom, double *Rener);
/* full_convec_ijk_H, elem_convec.f90:20-33 */
int sod_full_convec_ijk_H(int nelem, int npoin, const int *connec, const double *Ngp,
                          const double *dNgp, const double *He, const double *gpvol,
                          const double *dlxigp_ip, const double *xgp, const int *invAtoIJK,
                          const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                          const double *u, const double *q, const double *rho, const double *pr,
                          const double *E, double *Rmass, double *Rmom, double *Rener);
/* generic_scalar_convec_ijk, elem_convec.f90:462-476 */
int sod_generic_scalar_convec_ijk(int nelem, int npoin, const int *connec, const double *Ngp,
                                  const double *dNgp, const double *He, const double *gpvol,
                                  const double *dlxigp_ip, const double *xgp,
                                  const int *invAtoIJK, const int *gmshAtoI,
                                  const int *gmshAtoJ, const int *gmshAtoK, const double *q,
                                  const double *eta, const double *u, double *Rconvec);
/* full_diffusion_ijk, elem_diffu.f90:12-25 */
int sod_full_diffusion_ijk(int nelem, int npoin, const int *connec, const double *Ngp,
                           const double *He, const double *gpvol, const double *dlxigp_ip,
                           const int *invAtoIJK, const int *gmshAtoI, const int *gmshAtoJ,
                           const int *gmshAtoK, double Cp, double Pr, const double *rho_n,
                           const double *rho, const double *u, const double *Tem,
                           const double *mu_fluid, const double *mu_e, const double *mu_sgs,
                           const double *Ml, double *Rmass, double *Rmom, double *Rener);
/* sutherland_viscosity, mod_fluid_viscosity.f90:21-41 */
int sod_sutherland_viscosity(int npoin, const double *Tem, const double *mu_factor,
                             double *mu_fluid);
/* lumped_solver_scal / lumped_solver_vect, mod_solver.f90:23-58 */
int sod_lumped_solver_scal(int npoin, int npoin_w, const int *lpoin_w, const double *Ml,
                           double *R);
int sod_lumped_solver_vect(int npoin, int npoin_w, const int *lpoin_w, const double *Ml,
                           double *R);
/* smart_visc_spectral, mod_entropy_viscosity.f90:14-27 (the dummy helem is
 * helem_l(nelem,64), helem_k is helem(nelem), as at time_integ.f90:492). */
int sod_smart_visc_spectral(int nelem, int npoin, int npoin_w, const int *connec,
                            const int *lpoin_w, const double *Reta, const double *Ngp,
                            const double *coord, const double *dNgp, const double *gpvol,
                            const double *wgp, double gamma_gas, const double *rho,
                            const double *u, const double *csound, const double *Tem,
                            const double *eta, const double *helem, const double *helem_k,
                            const double *Ml, double *mu_e, const int *invAtoIJK,
                            const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                            double *mue_l);
/* No reference counterpart.  sod_smart_visc_spectral for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same arguments, kernels and results;
 * its reduction scratch (the entropy normalisation's block partials, their
 * sum and the normalisation) lives in a built-in device array of 16 KiB per
 * GPU, held until the process ends, so the call allocates nothing.  The
 * array is shared by every call on one GPU: calls that overlap on one GPU
 * race.  Requires sod_init; a workload needing more scratch than the array
 * holds is refused with SOD_ERR_ARG (the entropy viscosity needs a fixed
 * 8.3 KiB at any mesh size).  Called while the library stream
 * (sod_set_stream) is under graph capture, it records kernel launches only:
 * it does not wait on return, and it refuses with SOD_ERR_ARG node-order
 * table pointers that no earlier uncaptured call has checked (the check
 * copies them to the host). */
int sod_smart_visc_spectral_noalloc(int nelem, int npoin, int npoin_w, const int *connec,
                                    const int *lpoin_w, const double *Reta, const double *Ngp,
                                    const double *coord, const double *dNgp, const double *gpvol,
                                    const double *wgp, double gamma_gas, const double *rho,
                                    const double *u, const double *csound, const double *Tem,
                                    const double *eta, const double *helem, const double *helem_k,
                                    const double *Ml, double *mu_e, const int *invAtoIJK,
                                    const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                                    double *mue_l);
/* adapt_dt_cfl, mod_time_ops.f90:11-61.  `dt` is a HOST scalar (inout).  The
 * optional arguments cfl_diff/mu_fluid/mu_sgs/rho are all given or all NULL
 * (cfl_diff is passed by address for that reason). */
int sod_adapt_dt_cfl(int nelem, int npoin, const int *connec, const double *helem,
                     const double *u, const double *csound, double cfl_conv, double *dt,
                     const double *cfl_diff, const double *mu_fluid, const double *mu_sgs,
                     const double *rho);
/* volAvg_EK, mod_analysis.f90:10-44.  `EK` is a HOST scalar (out). */
int sod_volAvg_EK(int nelem, int npoin, const int *connec, const double *gpvol,
                  const double *Ngp, double rho0, const double *rho, const double *u,
                  double *EK);

/* One right-hand-side evaluation at a conserved state (rho, q(npoin,3), E):
 * the stage EOS, Sutherland viscosity, entropy residual, viscous and
 * convective residuals that one RK stage of rk_4_main forms
 * (time_integ.f90:266-373), assembled but not lumped-solved:
 * Rmass/Rmom/Rener = R_conv + R_diff, Reta = generic_scalar_convec_ijk of
 * (u eta, eta, u).  No single reference routine computes this sum; it is the
 * unit the fused stage kernel evaluates, exposed for parity tests and for
 * the adjoint.  The numerical parameters select the stage's branches as in
 * rk_4_main: flag_total_enthalpy != 0 takes full_convec_ijk_H with E := the
 * conserved E (time_integ.f90:361-362); with the Sutherland gate on
 * (flag_real_diff == 1 and flag_diff_suth == 1, time_integ.f90:312-316)
 * mu_factor is Sutherland's factor and the fluid viscosity is formed from the
 * stage temperature, with the gate off the array passed as mu_factor is the
 * fluid viscosity mu_fluid itself, as the stages then read the caller's
 * mu_fluid unchanged. */
int sod_full_rhs(int nelem, int npoin, int npoin_w, const int *connec, const int *lpoin_w,
                 const double *He, const double *gpvol, const double *Ml, const double *rho,
                 const double *q, const double *E, const double *mu_factor, const double *mu_e,
                 const double *mu_sgs, double Rgas, double gamma_gas, double Cp, double Prt,
                 double *Rmass, double *Rmom, double *Rener, double *Reta);

/* Adjoint of sod_full_rhs (no reference counterpart: the reference has no
 * adjoint).  For seeds lam_* on the four residuals (lam_mom(npoin,3)) it
 * returns the vector-Jacobian products
 *   g_rho = lam . dR/drho, g_q(npoin,3) = lam . dR/dq, g_E = lam . dR/dE
 * at the nodes and g_mu_e(nelem,64) = lam . dR/dmu_e at the element nodes,
 * with mu_factor and mu_sgs held fixed (with the Sutherland gate off the
 * mu_factor argument is mu_fluid, as for sod_full_rhs, and is held fixed).
 * Outputs are overwritten.  flag_total_enthalpy != 0 is refused with
 * SOD_ERR_UNSUPPORTED: the adjoint of full_convec_ijk_H is not implemented. */
int sod_full_rhs_vjp(int nelem, int npoin, int npoin_w, const int *connec, const int *lpoin_w,
                     const double *He, const double *gpvol, const double *Ml, const double *rho,
                     const double *q, const double *E, const double *mu_factor, const double *mu_e,
                     const double *mu_sgs, double Rgas, double gamma_gas, double Cp, double Prt,
                     const double *lam_mass, const double *lam_mom, const double *lam_ener,
                     const double *lam_eta, double *g_rho, double *g_q, double *g_E, double *g_mu_e);

/* Adjoint of sod_smart_visc_spectral (no reference counterpart).  For a
 * seed lam_mu_e(nelem,64) on mu_e it returns lam . dmu_e/dx for
 * x = rho, u(npoin,3), csound, Tem, eta, Reta (nodal; eta enters through the
 * normalisation max |eta - mean(eta)| over lpoin_w).  The limiters are
 * differentiated as subgradients: min(Ve, beta_e) follows Ve where
 * Ve < beta_e; each max over nodes follows its first maximiser (lowest gmsh
 * node, or first lpoin_w entry); max(x, ce) passes the gradient only when
 * x > ce.  Outputs are overwritten. */
int sod_smart_visc_spectral_vjp(int nelem, int npoin, int npoin_w, const int *connec, const int *lpoin_w,
                                const double *Reta, const double *gpvol, double gamma_gas, const double *rho,
                                const double *u, const double *csound, const double *Tem, const double *eta,
                                const double *helem, const double *helem_k, const double *lam_mu_e,
                                double *g_rho, double *g_u, double *g_csound, double *g_Tem, double *g_eta,
                                double *g_Reta);

/* No reference counterpart: selects how sod_rk_4_main evaluates a step.
 * 0 (default): fused stage kernels (one element kernel + one node kernel
 * per RK stage); 1: one kernel per reference routine, in the reference's
 * order.  Both give the reference's results to the stated tolerances. */
int sod_set_rk4_path(int path);

/* rk_4_main, time_integ.f90:137-191: one RK4 step of the periodic explicit
 * path.  State arrays carry the reference's time levels (rho(npoin,3),
 * u(npoin,3,3), q(npoin,3,3), pr(npoin,3), E(npoin,3), Tem(npoin,2),
 * e_int(npoin,2), eta(npoin,3)) and level pos=2 is advanced in place.
 * noBoundaries must be true, isWallModelOn false (or numBoundsWM 0) and
 * flag_les 0; the boundary-only arguments (point2elem, lnbn_nodes, nboun)
 * are accepted and unused, as the reference ignores them in this branch. */
int sod_rk_4_main(int noBoundaries, int isWallModelOn, int nelem, int nboun, int npoin,
                  int npoin_w, int numBoundsWM, const int *point2elem, const int *lnbn_nodes,
                  const double *dlxigp_ip, const double *xgp, const int *atoIJK,
                  const int *invAtoIJK, const int *gmshAtoI, const int *gmshAtoJ,
                  const int *gmshAtoK, int ppow, const int *connec, const double *Ngp,
                  const double *dNgp, const double *coord, const double *wgp, const double *He,
                  const double *Ml, const double *gpvol, double dt, const double *helem,
                  const double *helem_l, double Rgas, double gamma_gas, double Cp, double Prt,
                  double *rho, double *u, double *q, double *pr, double *E, double *Tem,
                  double *csound, double *machno, double *e_int, double *eta, double *mu_e,
                  double *mu_sgs, double *kres, double *etot, double *au, double *ax1,
                  double *ax2, double *ax3, const int *lpoin_w, double *mu_fluid,
                  const double *mu_factor, double *mue_l);

/* ------------------------------------------------ whole-step driver ---- */

/* The periodic explicit branch of the reference's time loop,
 * CFDSolverBase_evalTimeIteration (CFDSolverBase.f90:1501-1694, called
 * through CFDSolverPeriodic.f90:38-55), run on the device: per step
 *   [level 1 <- level 2 (:1532-1539)], rk_4_main, [level 3 <- level 1
 *   (:1552-1557)], time += dt, volAvg_EK on log steps (:1579-1581),
 *   evalDt = adapt_dt_cfl (:1132-1141, with the diffusive limit when
 *   flag_real_diff == 1),
 * with dt, time, the step counter and the EK history kept in device memory.
 * One step is captured as a HIP graph on the first run (and again when the
 * run's settings change) and replayed; no host round trip happens between
 * steps.  The arrays are the rk_4_main arrays with their time levels
 * (rho(npoin,3), u(npoin,3,3), ...); the loop keeps their addresses, which
 * must stay valid until sod_time_loop_destroy.  mu_e and mu_sgs are re-read
 * from the caller's arrays at the start of every run.
 * No single reference routine corresponds to the create/destroy pair; it
 * holds what the reference keeps in module variables and OpenACC data. */
typedef struct sod_time_loop sod_time_loop;
int sod_time_loop_create(sod_time_loop **loop, int nelem, int npoin, int npoin_w, const int *connec,
                         const int *lpoin_w, const double *He, const double *gpvol, const double *Ml,
                         const double *helem, const double *helem_l, const double *mu_factor, double Rgas,
                         double gamma_gas, double Cp, double Prt, double *rho, double *u, double *q,
                         double *pr, double *E, double *Tem, double *csound, double *machno,
                         double *e_int, double *eta, double *mu_e, const double *mu_sgs,
                         double *mu_fluid, double *mue_l);
/* Steps istep0 .. istep0+nsteps-1.  dt and time are HOST scalars (inout):
 * dt is the step size of the first step (the reference's this%dt) and, on
 * return, the one adapt_dt_cfl computed after the last step (fixed_dt != 0
 * keeps dt fixed and skips adapt_dt_cfl).  A step istep is a log step when
 * log_every > 0, istep >= log_first and (istep - log_first) % log_every == 0
 * (save_logFile_next, CFDSolverBase.f90:989-1010,1655-1657); EK (HOST,
 * nEK >= number of log steps) receives volAvg_EK(rho0, rho(:,2), u(:,:,2))
 * of each.  keep_levels != 0 performs the level copies. */
int sod_time_loop_run(sod_time_loop *loop, int istep0, int nsteps, double cfl_conv, double cfl_diff,
                      int fixed_dt, double *dt, double *time, int log_first, int log_every, double rho0,
                      double *EK, int nEK, int keep_levels);
int sod_time_loop_destroy(sod_time_loop *loop);

/* A flattened ensemble of ninst independent periodic meshes stepped together
 * (no reference counterpart: the reference runs one mesh per solver).  The
 * arrays are those of sod_time_loop_create for the concatenated meshes:
 * instance i owns elements elem_off[i] .. elem_off[i+1]-1, nodes
 * node_off[i] .. node_off[i+1]-1 and the lpoin_w entries work_off[i] ..
 * work_off[i+1]-1 (host arrays of ninst+1 values from 0), and its connec /
 * lpoin_w hold global node ids (1-based) inside its own node range.  Each
 * instance has its own dt (CFL condition over its elements) and its own
 * entropy normalisation; gas constants are shared, mu_factor is per node. */
int sod_time_loop_create_batched(sod_time_loop **loop, int ninst, const int *elem_off, const int *node_off,
                                 const int *work_off, int nelem, int npoin, int npoin_w, const int *connec,
                                 const int *lpoin_w, const double *He, const double *gpvol, const double *Ml,
                                 const double *helem, const double *helem_l, const double *mu_factor,
                                 double Rgas, double gamma_gas, double Cp, double Prt, double *rho, double *u,
                                 double *q, double *pr, double *E, double *Tem, double *csound, double *machno,
                                 double *e_int, double *eta, double *mu_e, const double *mu_sgs,
                                 double *mu_fluid, double *mue_l);
/* nsteps steps of every instance; dt and time are HOST arrays of ninst
 * values (inout), as in sod_time_loop_run. */
int sod_time_loop_run_batched(sod_time_loop *loop, int nsteps, double cfl_conv, double cfl_diff, int fixed_dt,
                              double *dt, double *time, int keep_levels);

/* No reference counterpart: the configuration of the fused stage kernel used
 * by sod_rk_4_main and the time loops.  Assembly strategies:
 *   0 element-node stores: each element's residuals are stored per element
 *     node and the node update sums each node's contributions in a fixed
 *     order (bit-reproducible);
 *   1 atomic scatter into node arrays (the reference's assembly,
 *     elem_convec.f90:442-455; its summation order varies from run to run);
 *   2 colour-ordered scatter: elements are coloured from the connectivity so
 *     that no two of one colour share a node, one launch per colour adds
 *     into the node arrays without atomics (a fixed order; meshes where an
 *     element holds a node twice, or that need more than 64 colours, cannot
 *     use it).
 * Modes:
 *   mode 0 (default): 1 element per workgroup with element-node stores;
 *   mode 1: at the first step of each mesh the library times the
 *     configurations with a fixed summation order (1, 2 or 4 elements per
 *     workgroup or persistent wavefronts with element-node stores, and
 *     colour-ordered scatter) on that mesh and state, and keeps the fastest;
 *   mode 2: as 1, with atomic scatter among the candidates.
 * Every configuration gives the reference's results to the stated
 * tolerances.  sod_full_rhs always uses the default. */
int sod_set_autotune(int mode);
/* No reference counterpart: pins one stage configuration (mode 3) instead of
 * a tuned one: elems_per_workgroup 1, 2 or 4 with assembly 0, 4 with
 * assembly 1, or 1 with assembly 2; elems_per_workgroup 0 with assembly 0
 * selects persistent one-wavefront workgroups (three per SIMD, each taking
 * every G-th element and loading the next element's node ids while it
 * works on the current one).  sod_set_autotune returns to modes 0-2.
 * A pinned colour-ordered configuration on a mesh that cannot be coloured
 * makes the stepping call return SOD_ERR_UNSUPPORTED. */
int sod_set_stage_config(int elems_per_workgroup, int assembly);
/* No reference counterpart: sod_set_stage_config with the LDS layout of the
 * element's nodal values as a third dimension: 0 swizzled slots (the
 * default), 1 the lane index itself (1 element per workgroup with assembly 0
 * only).  The autotuner's modes 1 and 2 time both layouts. */
int sod_set_stage_config_lds(int elems_per_workgroup, int assembly, int lds_layout);
/* No reference counterpart: the LDS layout (0 swizzled, 1 linear) of the
 * stage configuration of loop's mesh (NULL: the mesh most recently stepped
 * through sod_rk_4_main). */
int sod_stage_lds_layout(const sod_time_loop *loop, int *lds_layout);
/* No reference counterpart: 1 when the stage kernel of loop's mesh (NULL:
 * the mesh most recently stepped through sod_rk_4_main) takes its elements
 * in a locality order of its own -- breadth first over elements sharing
 * nodes, built at mesh preparation when fewer than half of the caller's
 * consecutive elements share a node -- else 0.  The order does not change
 * the results. */
int sod_stage_element_order(const sod_time_loop *loop, int *reordered);
/* The stage configuration of loop's mesh (loop NULL: the mesh most recently
 * stepped through sod_rk_4_main): elements per workgroup, assembly (0-2 as
 * above), the mode it was chosen under (0-2, or 3 when pinned) and the
 * measured time of one stage in ms (0 when not tuned).  Any output pointer
 * may be NULL. */
int sod_stage_config(const sod_time_loop *loop, int *elems_per_workgroup, int *assembly, int *tuned_mode,
                     double *stage_ms);
/* No reference counterpart: the element colouring of the colour-ordered
 * assembly of loop's mesh (loop NULL: the mesh most recently stepped through
 * sod_rk_4_main), once it has been built.  colour is a HOST array of nelem
 * ints that receives each element's colour (0-based), ncolours the number
 * of colours. */
int sod_element_colouring(const sod_time_loop *loop, int *ncolours, int *colour, int nelem);

#ifdef __cplusplus
}
#endif
#endif
