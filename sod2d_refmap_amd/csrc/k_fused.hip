// Fused RK4 stage for rk_4_main (time_integ.f90:234-411): one kernel per
// stage evaluates, element by element, the stage equation of state
// (:266-285), Sutherland viscosity (:312-316), the entropy residual
// (generic_scalar_convec_ijk), the viscous residual (full_diffusion_ijk) and
// the split-form convective residual (full_convec_ijk) and scatters the six
// residuals; one node kernel per stage does the lumped solve (:391-393), the
// accumulation (:399-409) and the next stage state (:239-248) or, after the
// last stage, the update and final equation of state (:417-471).
//
// Data layout (built once per mesh, "prepared mesh"):
//   conn  [e][64]    int32 node id (0-based) in lattice-lane order
//   He    [e][9][64] He(i,j) at row i+3j, lane order
//   gpvol [e][64], mu_e [e][64], mu_sgs [e][64]  lane order
// Nodal arrays keep the reference's SoA layout.
//
// Work decomposition: one wavefront per element, lane = lattice node
// (i + 4j + 16k).  Nodal values are gathered once into LDS with a swizzled
// slot (conflict-free 4-point line reads along xi, eta and zeta); each lane
// then forms its derivatives from the 12 points of the three lines through
// its node.  Divergences of products a*u are evaluated with the contravariant
// velocity U^c_b = sum_j He(j,c)(node) u_j(b), and advective terms u.grad(phi)
// with the node's own contravariant velocity, which is the same discrete
// operator as the reference's (He applied to the xi-derivative of each product)
// rearranged by linearity.
#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "stage_common.hpp"

namespace sod {

namespace {

constexpr int BLK = 256;

// ------------------------------------------------------------ element kernel
// Along each direction the four points of a lane's line are visited as the
// lane's own point (offset 0, from registers) and the three others (offsets
// 1..3, from LDS); the derivative coefficients are tabulated by (own index,
// offset): sDn[o][d] = D[o][(o+d)&3] for the strong-form derivative and
// sDt[o][d] = D[(o+d)&3][o] for the weak-form divergence.
// Three wavefronts per SIMD (amdgpu_waves_per_eu(3)) measured faster than
// the two the unconstrained register allocation gives.
// NW elements (wavefronts) per workgroup.  ASM selects the assembly:
//   ASM_PULL    the residuals are stored per element node into relem, and the
//               node update sums each node's contributions in a fixed order;
//   ASM_ATOMIC  they are added into six node arrays (stride rstride) with
//               device-scope atomics, as the reference scatters them
//               (elem_convec.f90:442-455, elem_diffu.f90:183-193);
//   ASM_COLOR   one launch per element colour (no two elements of a colour
//               share a node, no element holds a node twice): plain
//               read-add-write into the six node arrays, in colour order.
//               The workgroup's element is elist[blockIdx.x].
// TOTAL_H: the convective energy residual of full_convec_ijk_H
// (elem_convec.f90:20-208, selected by flag_total_enthalpy,
// time_integ.f90:361-362) with the total enthalpy per volume H = E + p
// instead of the h + k split of full_convec_ijk.
// mufac is mu_factor with the Sutherland gate on (P.suth) and the caller's
// mu_fluid otherwise.
// PERSIST (one wavefront per workgroup, element-node stores): the wavefront
// of workgroup b takes traversal positions b, b + G, b + 2G, ... (G = grid size) and
// loads the next element's node ids after the current element's own loads,
// so that an element waits for one round of loads (its gathers, issued with
// its geometry) instead of two (the node ids, then the gathers).
// LAY: the LDS slot of a lattice point, swizzled (LDS_SWIZZLED, the default)
// or the lane index itself (LDS_LINEAR); an autotuner candidate.
template <bool HAS_SGS, int NW, int ASM, bool TOTAL_H, bool PERSIST = false, int LAY = LDS_SWIZZLED>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_stage_rhs(
    int nelem, const int* __restrict__ elist, const int* __restrict__ conn, const double* __restrict__ He,
    const double* __restrict__ gpv, const double* __restrict__ mue, const double* __restrict__ musgs,
    const double* __restrict__ rho, const double* __restrict__ q0, const double* __restrict__ q1,
    const double* __restrict__ q2, const double* __restrict__ Ein, const double* __restrict__ mufac,
    double* __restrict__ relem, long rstride, StageParams P) {
    __shared__ double sh[NW][18][64];
    // The derivative matrix is copied to LDS with a flat index first:
    // indexing the __constant__ 2-D table with a lane-dependent row gave
    // wrong addresses for this kernel with the current compiler.
    __shared__ double sD[16], sDn[16], sDt[16];
    if (threadIdx.x < 16) sD[threadIdx.x] = (&c_tab.D[0][0])[threadIdx.x];
    __syncthreads();
    if (threadIdx.x < 16) {
        const int o = threadIdx.x >> 2, b = (o + (threadIdx.x & 3)) & 3;
        sDn[threadIdx.x] = sD[o * 4 + b];
        sDt[threadIdx.x] = sD[b * 4 + o];
    }
    __syncthreads();
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    // Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs
    // on XCD b % 8).  Multi-element workgroups give XCD x the contiguous
    // range of element blocks [x*q + min(x, r8), ...) so that neighbouring
    // elements, which share nodes, meet in the same L2; with one element per
    // workgroup the plain order measured 1.6% faster at 64^3.
    const int lb = NW == 1 ? (int)blockIdx.x : xcd_block(blockIdx.x, gridDim.x);
    if (lb * NW + w >= nelem) return;   // uniform per wavefront; only wave-level sync below
    const int i = l & 3, j = (l >> 2) & 3, k = l >> 4;
    // tp: the traversal position, where the element-node residuals are
    // stored; e: the element, elist[tp] when the prepared mesh has a
    // traversal order (elist non-null), else tp.  ASM_COLOR: this colour's
    // element list.
    int tp = lb * NW + w;
    int e = tp;
    if (ASM == ASM_COLOR)
        e = elist[blockIdx.x];
    else if (elist != nullptr)
        e = elist[tp];
    int n = conn[(size_t)e * 64 + l];
    for (;;) {   // one element, or (PERSIST) every G-th element
    const size_t eb = (size_t)e * 64;

    // ---- gather + equation of state (time_integ.f90:266-285) -----------
    const double r = rho[n];
    const double m0 = q0[n], m1 = q1[n], m2 = q2[n], en = Ein[n], mf = mufac[n];
    const double ri = 1.0 / r;
    const double u0 = m0 * ri, u1 = m1 * ri, u2 = m2 * ri;
    const double umag = u0 * u0 + u1 * u1 + u2 * u2;
    const double eint = en * ri - 0.5 * umag;
    const double p = r * P.gm1 * eint;
    const double h = P.gog1 * p * ri;
    const double T = p * ri * P.iRgas;
    // ln(p / rho^gamma) as ln p - gamma ln rho: two logarithms instead of a
    // logarithm and a power (same value to a few ulps; pow costs ~260 instructions)
    const double eta = (r * P.ig1) * (log(p) - P.gamma * log(r));
    const double kk = 0.5 * umag;
    // Sutherland's law (mod_fluid_viscosity.f90:35-39), T^1.5 as T*sqrt(T),
    // or the caller's mu_fluid with the gate off (time_integ.f90:312-316)
    const double muf = P.suth ? mf * 0.000001458 * (T * sqrt(T)) / (T + 110.4) : mf;
    double hm[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) hm[d][c] = He[eb * 9 + (d + 3 * c) * 64 + l];
    const double gv = gpv[eb + l];
    const double mu_e = mue[eb + l];
    const double mu_s = HAS_SGS ? musgs[eb + l] : 0.0;
    // the next element's node ids, issued after this element's loads so
    // that waiting for those does not wait for them
    const int tp_next = tp + (int)gridDim.x;
    int e_next = 0, n_next = 0;
    if (PERSIST && tp_next < nelem) {
        e_next = elist != nullptr ? elist[tp_next] : tp_next;
        n_next = conn[(size_t)e_next * 64 + l];
    }

    double(*S)[64] = sh[w];
    const int me = lds_slot<LAY>(i, j, k);
    // the convected energy variable: h + k (full_convec_ijk), or H = E + p
    // (full_convec_ijk_H: fe = (E+p)u, gradE = grad(p+E))
    const double hk = TOTAL_H ? en + p : h + kk;
    S[P_RHO][me] = r;
    S[P_U0][me] = u0;
    S[P_U1][me] = u1;
    S[P_U2][me] = u2;
    S[P_P][me] = p;
    S[P_H][me] = hk;
    S[P_ETA][me] = eta;
    S[P_T][me] = T;
    wave_sync();

    // ---- directional passes ---------------------------------------------
    // With DU_b = D * U^c_b (U^c the contravariant velocity formed with this
    // node's He), t_b = rho_b DU_b, DA_b = D * U^c(own) and sA_b = rho_b DA_b,
    // the split-form sums of elem_convec.f90:414-434 factor per line point b:
    //   energy   (h+k)_b (t_b + rho DU_b + rho DA_b + sA_b)   [div(rho h u), h div(rho u)...,
    //                                                        the same with k, and the advective terms]
    //   momentum u_a,b (t_b + rho DU_b + sA_b)
    //   entropy  eta_b (DU_b + DA_b)
    // (E := h in the energy flux, time_integ.f90:364; only h + k enters).
    // With TOTAL_H the energy residual has the entropy's form in H:
    //   energy   H_b (DU_b + DA_b)   [div((E+p)u) and u.grad(E+p)]
    double divQ = 0, accE = 0, accEta = 0, accQ0 = 0, accQ1 = 0, accQ2 = 0;
    double gP[3] = {0, 0, 0}, gT[3] = {0, 0, 0}, gR[3] = {0, 0, 0};
    double gU[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};   // gU[a][b] = du_a/dx_b
    const int own[3] = {i, j, k};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double h0 = hm[0][c], h1 = hm[1][c], h2 = hm[2][c];
        const double Ua = h0 * u0 + h1 * u1 + h2 * u2;
        const int o = own[c];
        double dp = 0, du0 = 0, du1 = 0, du2 = 0, dT = 0, dr = 0;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const double Dc = sDn[o * 4 + d];
            double rb, u0b, u1b, u2b, pb, hkb, etab, Tb;
            if (d == 0) {
                rb = r; u0b = u0; u1b = u1; u2b = u2; pb = p; hkb = hk; etab = eta; Tb = T;
            } else {
                const int bb = (o + d) & 3;
                const int sl = c == 0 ? lds_slot<LAY>(bb, j, k)
                                      : (c == 1 ? lds_slot<LAY>(i, bb, k) : lds_slot<LAY>(i, j, bb));
                rb = S[P_RHO][sl]; u0b = S[P_U0][sl]; u1b = S[P_U1][sl]; u2b = S[P_U2][sl];
                pb = S[P_P][sl]; hkb = S[P_H][sl]; etab = S[P_ETA][sl]; Tb = S[P_T][sl];
            }
            dp += Dc * pb;
            du0 += Dc * u0b;
            du1 += Dc * u1b;
            du2 += Dc * u2b;
            dT += Dc * Tb;
            dr += Dc * rb;
            const double DU = Dc * (h0 * u0b + h1 * u1b + h2 * u2b);
            const double DA = Dc * Ua;
            const double t = rb * DU;
            const double Y = t + r * DU + rb * DA;        // t + rho DU + sA
            accQ0 += u0b * Y;
            accQ1 += u1b * Y;
            accQ2 += u2b * Y;
            accE += TOTAL_H ? hkb * (DU + DA) : hkb * (Y + r * DA);
            accEta += etab * (DU + DA);
            divQ += t;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double hd = hm[d][c];
            gP[d] += hd * dp;
            gT[d] += hd * dT;
            gR[d] += hd * dr;
            gU[0][d] += hd * du0;
            gU[1][d] += hd * du1;
            gU[2][d] += hd * du2;
        }
    }

    // ---- convective residual (elem_convec.f90:414-434) -------------------
    const double uu[3] = {u0, u1, u2};
    const double divU = gU[0][0] + gU[1][1] + gU[2][2];
    const double Arho = u0 * gR[0] + u1 * gR[1] + u2 * gR[2];
    const double accQ[3] = {accQ0, accQ1, accQ2};
    double Rm = 0.5 * (divQ + r * divU) + 0.5 * Arho;
    // full_convec_ijk_H: 1/2 (div(fe) + (E+p) div u + u.grad(E+p))
    double Re = TOTAL_H ? 0.5 * (accE + hk * divU) : 0.25 * (accE + hk * (divQ + r * divU + Arho));
    double Rq[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double Au = u0 * gU[a][0] + u1 * gU[a][1] + u2 * gU[a][2];
        Rq[a] = 0.25 * (accQ[a] + uu[a] * (divQ + Arho) + (r * uu[a]) * divU + r * Au) + gP[a];
    }
    const double Reta = 0.5 * (accEta + eta * divU);

    // ---- viscous fluxes at this node (elem_diffu.f90:70-140) -------------
    const double mu_fgp = muf + r * mu_s;
    const double mu_t = mu_fgp + mu_e;
    const double twoThirds = 2.0 / 3.0;
    double tau[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) tau[a][b] = mu_t * (gU[a][b] + gU[b][a]);
#pragma unroll
    for (int a = 0; a < 3; ++a) tau[a][a] = tau[a][a] - mu_fgp * twoThirds * divU;
    double tauU[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) tauU[a] = tau[a][0] * u0 + tau[a][1] * u1 + tau[a][2] * u2;
    // contravariant fluxes G_c = gpvol * sum_d He(d,c) F_d for the six vector
    // fluxes grad(rho), tau rows 0..2, tau.u, grad(T)
    double hg[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) hg[d][c] = gv * hm[d][c];
    double G[6][3];
    wave_sync();   // all lanes are done reading the nodal values
    {
        const double* F[6] = {gR, tau[0], tau[1], tau[2], tauU, gT};
#pragma unroll
        for (int f = 0; f < 6; ++f)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                G[f][c] = hg[0][c] * F[f][0] + hg[1][c] * F[f][1] + hg[2][c] * F[f][2];
                S[3 * f + c][me] = G[f][c];
            }
    }
    wave_sync();
    // ---- weak divergence at this test node (elem_diffu.f90:142-194) ------
    double dD[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int o = own[c];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const double Dt = sDt[o * 4 + d];   // l'_own(xi_q)
            if (d == 0) {
#pragma unroll
                for (int f = 0; f < 6; ++f) dD[f] += Dt * G[f][c];
            } else {
                const int bb = (o + d) & 3;
                const int sl = c == 0 ? lds_slot<LAY>(bb, j, k)
                                      : (c == 1 ? lds_slot<LAY>(i, bb, k) : lds_slot<LAY>(i, j, bb));
#pragma unroll
                for (int f = 0; f < 6; ++f) dD[f] += Dt * S[3 * f + c][sl];
            }
        }
    }
    const double nu_e = P.c_rho * mu_e * ri;
    const double kappa = muf * P.CpPr + P.c_ener * mu_e * P.CpPr + P.Cp09 * r * mu_s;

    const double res[NRES] = {gv * Rm + nu_e * dD[0], gv * Rq[0] + dD[1], gv * Rq[1] + dD[2],
                              gv * Rq[2] + dD[3], gv * Re + dD[4] + kappa * dD[5], gv * Reta};
    if (ASM == ASM_ATOMIC) {
#pragma unroll
        for (int f = 0; f < NRES; ++f) unsafeAtomicAdd(relem + f * rstride + n, res[f]);
    } else if (ASM == ASM_COLOR) {
        // this colour's elements own their nodes exclusively
#pragma unroll
        for (int f = 0; f < NRES; ++f) relem[f * rstride + n] += res[f];
    } else {
        // The residuals of this element at this node, stored at its
        // traversal position's element-node slot; k_stage_update and
        // k_assemble_nodes sum each node's contributions (pull assembly: no
        // atomics, a fixed summation order).
        double2* o = reinterpret_cast<double2*>(relem + ((size_t)tp * 64 + l) * NRES);
        o[0] = make_double2(res[0], res[1]);
        o[1] = make_double2(res[2], res[3]);
        o[2] = make_double2(res[4], res[5]);
    }
    if (!PERSIST || tp_next >= nelem) break;
    tp = tp_next;
    e = e_next;
    n = n_next;
    wave_sync();   // this element's LDS reads before the next element's writes
    }
}

// Sum of the element-node residuals that belong to node n: the positions
// p = e*64 + lane of its occurrences in the lane-layout connectivity,
// ascending (a fixed order, so the assembled residual is reproducible).
// With inv_off indexed by the working-node position instead of the node id
// (woff), the same sum needs no lpoin_w load before its offsets.
__device__ __forceinline__ void node_residuals(int n, const int* __restrict__ inv_off, const int* __restrict__ inv_pos,
                                               const double* __restrict__ relem, double acc[NRES]) {
#pragma unroll
    for (int f = 0; f < NRES; ++f) acc[f] = 0.0;
    const int q1 = inv_off[n + 1];
    for (int q = inv_off[n]; q < q1; ++q) {
        const double2* c = reinterpret_cast<const double2*>(relem + (size_t)inv_pos[q] * NRES);
        const double2 a = c[0], b = c[1], d = c[2];
        acc[0] += a.x;
        acc[1] += a.y;
        acc[2] += b.x;
        acc[3] += b.y;
        acc[4] += d.x;
        acc[5] += d.y;
    }
}

// Assembled residuals of every node into the caller's arrays (sod_full_rhs).
__global__ __launch_bounds__(256) void k_assemble_nodes(int npoin, const int* __restrict__ inv_off,
                                                        const int* __restrict__ inv_pos,
                                                        const double* __restrict__ relem, double* __restrict__ Rmass,
                                                        double* __restrict__ Rmom, double* __restrict__ Rener,
                                                        double* __restrict__ Reta) {
    const size_t np = (size_t)npoin;
    for (int n = blockIdx.x * 256 + threadIdx.x; n < npoin; n += gridDim.x * 256) {
        double acc[NRES];
        node_residuals(n, inv_off, inv_pos, relem, acc);
        Rmass[n] = acc[R_MASS];
        Rmom[n] = acc[R_MOM0];
        Rmom[n + np] = acc[R_MOM1];
        Rmom[n + 2 * np] = acc[R_MOM2];
        Rener[n] = acc[R_ENER];
        Reta[n] = acc[R_ETA];
    }
}

// ------------------------------------------------------------ node kernels
__constant__ double k_a[4] = {0.0, 0.5, 0.5, 1.0};
__constant__ double k_b[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

struct StepState {
    // caller level-2 arrays
    double *rho, *q0, *q1, *q2, *E;
    double *u, *pr, *Tem, *e_int, *eta, *csound, *machno, *mu_fluid;
    const double *Ml, *mu_factor;
    // library arrays
    double *aux;    // 5 x npoin: rho, q0, q1, q2, E of the next stage
    double *Rsum;   // 6 x npoin
    // element-node residuals of the stage kernel and the node -> (element,
    // lane) positions they are summed over
    const double* relem;
    const int *inv_off, *inv_pos;
    // the same positions by working-node position it (lpoin_w order):
    // wpos[woff[it] .. woff[it+1])
    const int *woff, *wpos;
    // node residual arrays of the atomic-scatter stage kernel (6 x npoin),
    // zeroed again by the update that reads them
    double* Rn;
    double *partial;   // reduction partials (RED_BLOCKS)
    // CFL time step of the next step folded into the final update (node_dt;
    // one mesh, mu_sgs = 0): per-node h_min, and the bits of the running
    // minimum.  hmin == nullptr: not folded.
    const double* hmin;
    unsigned long long* dt_bits;
    double cfl_conv, cfl_diff;
    int has_diff;
    // Sutherland gate (time_integ.f90:482-486): the final update writes
    // mu_fluid from the new temperature; otherwise mu_fluid is the caller's
    // and is only read
    int suth;
};

constexpr int RED_BLOCKS = 1024;

// Grid of the node update of stages 1-3: one thread per working node (no
// grid-stride loop), 6% faster than 1024 or 4096 resident blocks at 64^3.
// The final stage keeps RED_BLOCKS blocks: it leaves one partial sum of eta
// per block, and one block sums them.
inline int node_grid(int npoin_w) { return std::max(1, grid_for(npoin_w, BLK)); }

// Stage s in 0..3 after its element kernel: lumped solve, accumulation,
// next stage state (s < 3) or final update + EOS (s == 3).
// With BATCHED the nodes belong to independent instances (a flattened
// ensemble): dt is d_dt[node_inst[n]] and the entropy statistics are formed
// per instance by their own kernels.
template <bool FINAL, bool BATCHED, bool ATOMIC>
__global__ __launch_bounds__(BLK) void k_stage_update(int npoin, int npoin_w, const int* __restrict__ lpoin_w,
                                                      int s, const double* __restrict__ d_dt, double gamma,
                                                      double Rgas, StepState st,
                                                      const int* __restrict__ node_inst) {
    __shared__ double s_red[BLK / 64];
    __shared__ double s_dt[BLK / 64];
    const size_t np = (size_t)npoin;
    double eta_acc = 0.0;
    double dt_min = HUGE_VAL;
    for (int it = blockIdx.x * BLK + threadIdx.x; it < npoin_w; it += gridDim.x * BLK) {
        const int n = lpoin_w[it] - 1;
        const double dt = BATCHED ? d_dt[node_inst[n]] : d_dt[0];
        const double ml = st.Ml[n];
        double Rf[NRES];
        if (ATOMIC) {
#pragma unroll
            for (int f = 0; f < NRES; ++f) {
                Rf[f] = st.Rn[f * np + n];
                st.Rn[f * np + n] = 0.0;
            }
        } else {
            node_residuals(it, st.woff, st.wpos, st.relem, Rf);
        }
#pragma unroll
        for (int f = 0; f < NRES; ++f) Rf[f] = Rf[f] / ml;   // mod_solver.f90:32-36
        const double b = k_b[s];
        double Rs[NRES];
#pragma unroll
        for (int f = 0; f < NRES; ++f) {
            const double prev = s == 0 ? 0.0 : st.Rsum[f * np + n];
            Rs[f] = prev + b * Rf[f];
        }
        const double y[5] = {st.rho[n], st.q0[n], st.q1[n], st.q2[n], st.E[n]};
        // residual order: mass, mom0..2, ener -> state order rho, q0..2, E
        if (!FINAL) {
#pragma unroll
            for (int f = 0; f < NRES; ++f) st.Rsum[f * np + n] = Rs[f];
            const double fa = dt * k_a[s + 1];
#pragma unroll
            for (int f = 0; f < 5; ++f) st.aux[f * np + n] = y[f] - fa * Rf[f];
        } else {
            double yn[5];
#pragma unroll
            for (int f = 0; f < 5; ++f) yn[f] = y[f] - dt * Rs[f];
            st.rho[n] = yn[0];
            st.q0[n] = yn[1];
            st.q1[n] = yn[2];
            st.q2[n] = yn[3];
            st.E[n] = yn[4];
            // final EOS (time_integ.f90:445-463)
            const double r = yn[0];
            const double uv[3] = {yn[1] / r, yn[2] / r, yn[3] / r};
            double umag = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                st.u[n + np * d] = uv[d];
                umag = umag + uv[d] * uv[d];
            }
            const double ei = (yn[4] / r) - 0.5 * umag;
            const double p = r * (gamma - 1.0) * ei;
            st.e_int[n] = ei;
            st.pr[n] = p;
            const double c = sqrt(gamma * p / r);
            st.csound[n] = c;
            st.machno[n] = sqrt(umag) / c;
            const double T = p / (r * Rgas);
            st.Tem[n] = T;
            const double et = (r / (gamma - 1.0)) * (log(p) - gamma * log(r));
            st.eta[n] = et;
            eta_acc += et;
            // entropy residual Reta = -Reta_sum (time_integ.f90:467-471),
            // kept in the (now free) first aux array for the entropy viscosity
            st.aux[n] = -Rs[R_ETA];
            double mu;
            if (st.suth) {
                mu = sutherland(st.mu_factor[n], T);
                st.mu_fluid[n] = mu;
            } else {
                mu = st.mu_fluid[n];
            }
            if (!BATCHED && st.hmin != nullptr) {
                const double h = st.hmin[n];
                if (h < HUGE_VAL)
                    dt_min = fmin(dt_min, node_dt(h, uv[0], uv[1], uv[2], c, mu, r, st.cfl_conv, st.has_diff,
                                                  st.cfl_diff));
            }
        }
    }
    if (FINAL && !BATCHED) {
        // partial sums of eta over working nodes for the normalisation
        double v = eta_acc;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int q = 0; q < BLK / 64; ++q) t += s_red[q];
            st.partial[blockIdx.x] = t;
        }
        if (st.hmin != nullptr) {
            const double v = wave_min(dt_min);
            if ((threadIdx.x & 63) == 0) s_dt[threadIdx.x >> 6] = v;
            __syncthreads();
            if (threadIdx.x == 0) {
                double m = s_dt[0];
                for (int q = 1; q < BLK / 64; ++q) m = fmin(m, s_dt[q]);
                if (m < HUGE_VAL) atomicMin(st.dt_bits, (unsigned long long)__double_as_longlong(m));
            }
        }
    }
}

// mu_fluid at periodic slave nodes: Sutherland of their (unchanged) Tem,
// as the end-of-step call over all npoin does (time_integ.f90:482-486).
__global__ __launch_bounds__(BLK) void k_slave_sutherland(int nslave, const int* __restrict__ slaves,
                                                          const double* __restrict__ Tem,
                                                          const double* __restrict__ mu_factor,
                                                          double* __restrict__ mu_fluid) {
    const int t = blockIdx.x * BLK + threadIdx.x;
    if (t >= nslave) return;
    const int n = slaves[t];
    mu_fluid[n] = sutherland(mu_factor[n], Tem[n]);
}

// mesh preparation: connec (ref layout, 1-based) -> lane layout (0-based),
// He and gpvol -> lane layout
__global__ __launch_bounds__(BLK) void k_prepare_mesh(int nelem, const int* __restrict__ connec,
                                                      const double* __restrict__ He,
                                                      const double* __restrict__ gpvol,
                                                      int* __restrict__ conn_l, double* __restrict__ He_l,
                                                      double* __restrict__ gpv_l) {
    const long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= (long)nelem * 64) return;
    const int e = (int)(t >> 6), l = (int)(t & 63);
    const int a = c_tab.lane2node[l];
    conn_l[t] = connec[(size_t)e + (size_t)nelem * a] - 1;
    for (int c = 0; c < 9; ++c) He_l[(size_t)e * 576 + c * 64 + l] = He[9 * (size_t)a + 576 * (size_t)e + c];
    gpv_l[t] = gpvol[a + 64 * (size_t)e];
}

__global__ void k_zero_int(int* p) {
    if (threadIdx.x == 0) *p = 0;
}

__global__ void k_set_double_f(double* p, double v) {
    if (threadIdx.x == 0) *p = v;
}

// Entropy statistics of a flattened ensemble (mod_entropy_viscosity.f90:43-66
// per instance): the working nodes of instance i are lpoin_w[woff[i] ..
// woff[i+1]); stats[2i] = sum of eta, stats[2i+1] = max |eta - mean| (as the
// bits of a non-negative double, ordered like the value).
__global__ void k_zero_stats(int n, double* __restrict__ stats) {
    for (int t = threadIdx.x; t < n; t += blockDim.x) stats[t] = 0.0;
}

__global__ __launch_bounds__(256) void k_inst_eta_sum(const int* __restrict__ woff, const int* __restrict__ lpoin_w,
                                                      const double* __restrict__ eta, double* __restrict__ stats) {
    __shared__ double s_red[4];
    const int inst = blockIdx.y;
    const int b = woff[inst], end = woff[inst + 1];
    double v = 0.0;
    for (int t = b + blockIdx.x * 256 + threadIdx.x; t < end; t += gridDim.x * 256) v += eta[lpoin_w[t] - 1];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) unsafeAtomicAdd(&stats[2 * inst], (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]));
}

__global__ __launch_bounds__(256) void k_inst_eta_dev(const int* __restrict__ woff, const int* __restrict__ lpoin_w,
                                                      const double* __restrict__ eta, double* __restrict__ stats) {
    __shared__ double s_red[4];
    const int inst = blockIdx.y;
    const int b = woff[inst], end = woff[inst + 1];
    const double mean = stats[2 * inst] / (double)(end - b);
    double v = 0.0;
    for (int t = b + blockIdx.x * 256 + threadIdx.x; t < end; t += gridDim.x * 256)
        v = fmax(v, fabs(eta[lpoin_w[t] - 1] - mean));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double m = fmax(fmax(s_red[0], s_red[1]), fmax(s_red[2], s_red[3]));
        atomicMax(reinterpret_cast<unsigned long long*>(&stats[2 * inst + 1]),
                  (unsigned long long)__double_as_longlong(m));
    }
}

// norms[i] = stats[2i + 1] (one value per instance for the entropy kernel)
__global__ void k_inst_norms(int ninst, const double* __restrict__ stats, double* __restrict__ norms) {
    for (int t = threadIdx.x; t < ninst; t += blockDim.x) norms[t] = stats[2 * t + 1];
}

// Tiled transpose of up to two element-node arrays (reference layout
// (nelem,64), element fastest) into lane layout [e][lane]; 16 elements per
// block so that both the reads and the writes are contiguous.  `nonzero` is
// set when the second array holds a non-zero value; with dstB null the second
// array is only scanned for that flag.
__global__ __launch_bounds__(256) void k_to_lane_tiled(int nelem, const double* __restrict__ srcA,
                                                       double* __restrict__ dstA, const double* __restrict__ srcB,
                                                       double* __restrict__ dstB, int* __restrict__ nonzero) {
    __shared__ double sA[16][65], sB[16][65];
    __shared__ int s_l2n[64];
    if (threadIdx.x < 64) s_l2n[threadIdx.x] = c_tab.lane2node[threadIdx.x];
    const int e0 = xcd_block(blockIdx.x, gridDim.x) * 16;
    const size_t ne = (size_t)nelem;
    int nz = 0;
    for (int idx = threadIdx.x; idx < 1024; idx += 256) {
        const int a = idx >> 4, el = idx & 15, e = e0 + el;
        if (e < nelem) {
            sA[el][a] = srcA[(size_t)e + ne * a];
            if (srcB) {
                const double v = srcB[(size_t)e + ne * a];
                if (dstB) sB[el][a] = v;
                nz |= v != 0.0;
            }
        }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < 1024; idx += 256) {
        const int el = idx >> 6, l = idx & 63, e = e0 + el;
        if (e < nelem) {
            dstA[(size_t)e * 64 + l] = sA[el][s_l2n[l]];
            if (dstB) dstB[(size_t)e * 64 + l] = sB[el][s_l2n[l]];
        }
    }
    if (nz && nonzero) *nonzero = 1;
}

// smart_visc_spectral element loop (mod_entropy_viscosity.f90:68-122) on
// lane-layout mesh data: a block takes 16 consecutive elements, one wavefront
// per element at a time; results are staged in LDS and written to the
// reference-layout mu_e(nelem,64) / mue_l(nelem,64) with contiguous stores.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_smart_visc_lane(
    int nelem, int npoin, const int* __restrict__ conn_l, const double* __restrict__ gpv_l,
    const double* __restrict__ helem_l, const double* __restrict__ helem, const double* __restrict__ Reta,
    double gamma_gas, const double* __restrict__ rho, const double* __restrict__ u,
    const double* __restrict__ csound, const double* __restrict__ Tem, const double* __restrict__ d_norm,
    int normalise, double Rgas, double ce, double cmax, double cglob, double* __restrict__ mu_e,
    double* __restrict__ mue_l, double* __restrict__ mue_lane, const int* __restrict__ elem_inst) {
    // elem_inst (flattened ensembles): d_norm holds one normalisation per
    // instance and element e uses d_norm[elem_inst[e]]
    __shared__ double s_ml[16][65], s_me[16][65];
    __shared__ int s_l2n[64];
    __shared__ double s_w[64];
    if (threadIdx.x < 64) {
        const int a = c_tab.lane2node[threadIdx.x];
        s_l2n[threadIdx.x] = a;
        s_w[threadIdx.x] = c_tab.wgp[a];
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int e0 = blockIdx.x * 16;   // plain dispatch order (2% faster than the XCD remapping)
    const size_t np = (size_t)npoin, ne = (size_t)nelem;
    // helem_l(nelem,64) of the block's elements, read contiguously into s_me
    for (int idx = threadIdx.x; idx < 1024; idx += 256) {
        const int a2 = idx >> 4, el = idx & 15, e = e0 + el;
        if (e < nelem) s_me[el][a2] = helem_l[(size_t)e + ne * a2];
    }
    __syncthreads();
    const int a = s_l2n[l];
    // The wavefront's 4 elements unrolled, so that their gathers are issued
    // together (13% less time than one element after the other): per
    // element the extrema over its 64 lanes (min/max J, max |u|, max c,
    // max beta, uniform) and this lane's R1.  Elements beyond nelem keep
    // placeholders that are never stored.
    double ex[4][4], bet[4], R1v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int el = w + 4 * r, e = e0 + el;
        ex[r][0] = 1.0;
        ex[r][1] = 1.0;
        ex[r][2] = 0.0;
        ex[r][3] = 1.0;
        bet[r] = 0.0;
        R1v[r] = 0.0;
        if (e < nelem) {   // uniform per wavefront
            const double norm = normalise ? d_norm[elem_inst ? elem_inst[e] : 0] : 1.0;
            const size_t eb = (size_t)e * 64 + l;
            const int n = conn_l[eb];
            const double J = gpv_l[eb] / s_w[l];
            const double u0 = u[n], u1 = u[n + np], u2 = u[n + 2 * np];
            const double vmag = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
            ex[r][0] = wave_min(fmin(1000000.0, J));
            ex[r][1] = wave_max(fmax(0.0, J));
            ex[r][2] = wave_max(vmag);
            ex[r][3] = wave_max(csound[n]);
            const double aux3 = sqrt(Rgas * gamma_gas * Tem[n]);
            bet[r] = wave_max((rho[n] * helem[e]) * (cmax / 3.0) * (vmag + aux3));
            R1v[r] = rho[n] * fabs(Reta[n]) / norm;
        }
    }
    // The limiter ced of element r is the same in its 64 lanes: it is formed
    // once for the 4 elements, element r in the lanes l with l % 4 == r (the
    // same operations on the same values), and read back from lane r.
    const int rl = l & 3;
    double xl[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) xl[k] = rl == 0 ? ex[0][k] : rl == 1 ? ex[1][k] : rl == 2 ? ex[2][k] : ex[3][k];
    const double M = xl[2] / xl[3];
    double M15 = M * M;
    M15 = M15 * M15 * M;
    M15 = M15 * M15 * M15;
    const double ceM = fmax(tanh(M15 * 3.14159265358979323846), ce);
    const double rr = xl[0] / xl[1];
    double ced_l = fmax(1.0 - rr * rr, ce);
    ced_l = fmax(ced_l, ceM);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int el = w + 4 * r, e = e0 + el;
        if (e < nelem) {
            const long long cb = __double_as_longlong(ced_l);
            const int lo = __builtin_amdgcn_readlane((int)cb, r), hi = __builtin_amdgcn_readlane((int)(cb >> 32), r);
            const double ced = __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
            const double hl = s_me[el][a];
            const double Ve = ced * R1v[r] * (hl * hl);
            s_ml[el][a] = cglob * fmin(Ve, bet[r]);
        }
    }
    __syncthreads();
    // 27-point (1/4,1/2,1/4)^3 filter over geometric positions; the index
    // table convertIJK(0:5) = {3,1,3,4,2,3} (mod_filters.f90:42-51) maps a
    // geometric position to a sod2d 1-D index, with its end entries as built.
    // Each tap's slot is formed once and applied to the wavefront's 4
    // elements; every element's sum keeps the taps' order.
    const int conv[6] = {3, 1, 3, 4, 2, 3};
    const int gi = (l & 3) + 1, gj = ((l >> 2) & 3) + 1, gk = (l >> 4) + 1;
    const double w3[3] = {0.25, 0.5, 0.25};
    const int so = s_l2n[(conv[gi] - 1) + 4 * (conv[gj] - 1) + 16 * (conv[gk] - 1)];
    int tap[27];
#pragma unroll
    for (int ll = -1; ll <= 1; ++ll)
#pragma unroll
        for (int mm = -1; mm <= 1; ++mm)
#pragma unroll
            for (int nn = -1; nn <= 1; ++nn)
                tap[9 * (ll + 1) + 3 * (mm + 1) + nn + 1] =
                    s_l2n[(conv[gi + ll] - 1) + 4 * (conv[gj + mm] - 1) + 16 * (conv[gk + nn] - 1)];
    for (int r = 0; r < 4; ++r) {
        const int el = w + 4 * r, e = e0 + el;
        if (e >= nelem) break;
        double acc = 0.0;
#pragma unroll
        for (int ll = -1; ll <= 1; ++ll)
#pragma unroll
            for (int mm = -1; mm <= 1; ++mm)
#pragma unroll
                for (int nn = -1; nn <= 1; ++nn)
                    acc = acc + w3[ll + 1] * w3[mm + 1] * w3[nn + 1] * s_ml[el][tap[9 * (ll + 1) + 3 * (mm + 1) + nn + 1]];
        s_me[el][so] = acc;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < 1024; idx += 256) {
        const int a2 = idx >> 4, el = idx & 15, e = e0 + el;
        if (e < nelem) {
            mue_l[(size_t)e + ne * a2] = s_ml[el][a2];
            mu_e[(size_t)e + ne * a2] = s_me[el][a2];
        }
    }
    if (mue_lane) {
        for (int idx = threadIdx.x; idx < 1024; idx += 256) {
            const int el = idx >> 6, ln = idx & 63, e = e0 + el;
            if (e < nelem) mue_lane[(size_t)e * 64 + ln] = s_me[el][s_l2n[ln]];
        }
    }
}

// Position-weighted fingerprint of sampled mesh data (connec, He, gpvol,
// lpoin_w, Ml) so that a prepared mesh is rebuilt when a caller reuses the
// same addresses for different contents.  The 4096 samples k = t + 256 r
// (workgroup r of 16) take 4096 evenly spaced elements, each at a node and
// an He entry that vary with k, and 4096 evenly spaced lpoin_w entries.
// Workgroup r leaves its 4 sums in out[4r .. 4r+3]; the host adds the
// rounds in order (fingerprint()).  This is a safety net for meshes
// rewritten in place, not a proof: callers that change mesh arrays at the
// same addresses say so with sod_invalidate_mesh_cache (sod_update_device
// and sod_exit_data of a registered mesh array do it for them).
__global__ __launch_bounds__(256) void k_fingerprint(int nelem, int npoin_w, const int* __restrict__ connec,
                                                     const double* __restrict__ He,
                                                     const double* __restrict__ gpvol,
                                                     const int* __restrict__ lpoin_w,
                                                     const double* __restrict__ Ml, double* __restrict__ out) {
    __shared__ double s[4][256];
    const int t = threadIdx.x, smp = blockIdx.x;
    const long k = t + 256L * smp;
    const int e = (int)(((long)nelem * k) / 4096);
    const int nd = (int)((k * 37 + (k >> 6)) & 63);
    const int hi = (int)((k * 131 + (k >> 4)) % 576);
    const double wgt = 1.0 + 0.001 * (double)k;
    const int w = (int)(((long)npoin_w * k) / 4096);
    s[0][t] = wgt * connec[(size_t)e + (size_t)nelem * nd];
    s[1][t] = wgt * He[576 * (size_t)e + hi];
    s[2][t] = wgt * gpvol[64 * (size_t)e + nd];
    s[3][t] = wgt * lpoin_w[w] + wgt * Ml[lpoin_w[w] - 1];
    __syncthreads();
    if (t < 4) {
        double v = 0;
        for (int q = 0; q < 256; ++q) v += s[t][q];
        out[4 * smp + t] = v;
    }
}

}  // namespace

// ------------------------------------------------------------ prepared mesh
// Lane-layout copies of the static mesh data, the slave-node list and the
// step's work arrays, built once per mesh.  The drop-in entry points share
// one instance (rebuilt when the arguments' addresses or sampled contents
// change); each whole-step driver owns its own, so that a captured graph
// never refers to memory another call may free.
struct Prepared {
    const void* key[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int nelem = -1, npoin = -1, npoin_w = -1;
    int* conn = nullptr;
    double* He = nullptr;
    double* gpv = nullptr;
    double* mue = nullptr;
    double* musgs = nullptr;
    int* slaves = nullptr;
    int nslave = 0;
    double* work = nullptr;   // aux(5) + Rsum(6) per npoin, + partials
    double* relem = nullptr;  // nelem*64*6 element-node residuals of the stage kernel
    int* inv_off = nullptr;   // npoin+1: node n's positions are inv_pos[inv_off[n] .. inv_off[n+1])
    int* inv_pos = nullptr;   // nelem*64 positions e*64 + lane, ascending per node
    int* woff = nullptr;      // npoin_w+1: the offsets of the working node at lpoin_w position it
    int* wpos = nullptr;      // its positions: inv_pos_r itself when lpoin_w increases, else a copy
    // the positions the stage kernel stores at: r*64 + lane for the element
    // at traversal position r (eorder), in inv_pos's order; inv_pos itself
    // when there is no traversal order
    int* inv_pos_r = nullptr;
    int* flag = nullptr;
    double* dscal = nullptr;  // device scalars: [0] dt of the step
    double* fpbuf = nullptr;
    double fp[4] = {0, 0, 0, 0};
    size_t work_doubles = 0;
    // stage configuration (an entry of kStageCfgs) and how it was chosen:
    // cfg_mode is the autotune mode it was tuned under (0: the default),
    // cfg_ms the measured time of one stage with it
    int cfg = 0, cfg_mode = 0;
    double cfg_ms = 0.0;
    double* rn = nullptr;   // 6 x npoin node residuals of the atomic / colour-ordered configurations
    // element colouring of the colour-ordered configuration (built on first
    // use): elements colour by colour, and where each colour starts
    int* color_elems = nullptr;
    std::vector<int> color_off;
    int color_state = 0;   // 0 not built, 1 built, 2 the mesh cannot be coloured
    // the stage kernel's traversal order (locality_order), or null: the
    // caller's element order
    int* eorder = nullptr;
    void release() {
        (void)hipFree(conn);
        (void)hipFree(He);
        (void)hipFree(gpv);
        (void)hipFree(mue);
        (void)hipFree(musgs);
        (void)hipFree(slaves);
        (void)hipFree(work);
        (void)hipFree(relem);
        (void)hipFree(inv_off);
        if (wpos != inv_pos && wpos != inv_pos_r) (void)hipFree(wpos);
        if (inv_pos_r != inv_pos) (void)hipFree(inv_pos_r);
        (void)hipFree(inv_pos);
        (void)hipFree(woff);
        (void)hipFree(flag);
        (void)hipFree(dscal);
        (void)hipFree(fpbuf);
        (void)hipFree(rn);
        (void)hipFree(color_elems);
        (void)hipFree(eorder);
        *this = Prepared();
    }
};

// Prepared meshes of the drop-in entry points.  A few recent meshes are kept
// (up to kPrepSlots of them and kPrepBytes of device memory), so that a
// caller alternating between meshes -- the members of an ensemble stepped one
// after the other -- does not rebuild them on every call.
constexpr int kPrepSlots = 64;
constexpr size_t kPrepBytes = size_t(32) << 30;
struct PrepSlot {
    Prepared P;
    unsigned long long used = 0;
    size_t bytes = 0;
};
static PrepSlot g_slots[kPrepSlots];
static unsigned long long g_prep_clock = 0;

static size_t prepared_bytes(int nelem, int npoin) {
    // conn, He, gpvol, mu_e, mu_sgs, element-node residuals and positions per
    // element-node; work, inverse offsets per node
    return (size_t)nelem * 64 * (4 + 72 + 8 + 8 + 8 + 48 + 4) + (size_t)npoin * (11 * 8 + 4);
}

void fused_release() {
    for (auto& sl : g_slots) {
        sl.P.release();
        sl = PrepSlot();
    }
}

// Drops the drop-in entry points' prepared meshes built from an array that
// lies in [lo, lo + bytes) (every one of them when lo is null), after the
// work queued on the library stream.
int fused_invalidate(const void* lo, size_t bytes) {
    const char* a = static_cast<const char*>(lo);
    bool synced = false;
    for (auto& sl : g_slots) {
        if (sl.used == 0) continue;
        bool hit = lo == nullptr;
        for (int t = 0; t < 5 && !hit; ++t) {
            const char* k = static_cast<const char*>(sl.P.key[t]);
            hit = k >= a && k < a + bytes;
        }
        if (!hit) continue;
        if (!synced) {
            SOD_HIP(hipStreamSynchronize(rt().stream));
            synced = true;
        }
        sl.P.release();
        sl = PrepSlot();
    }
    return SOD_OK;
}

// connectivity ids must lie in 1..npoin (the kernels gather and scatter
// through them without further checks)
__global__ __launch_bounds__(256) void k_check_conn(long n, const int* __restrict__ conn_l, int npoin,
                                                    int* __restrict__ bad) {
    int b = 0;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
        const int v = conn_l[t];
        b |= (v < 0) | (v >= npoin);
    }
    if (b) *bad = 1;
}

constexpr int kFpRounds = 16;

// The fingerprint of the mesh arrays (k_fingerprint): the rounds' sums added
// in round order on the host.  buf holds 4 * kFpRounds doubles.
static int fingerprint(int nelem, int npoin_w, const int* connec, const double* He, const double* gpvol,
                       const int* lpoin_w, const double* Ml, double* buf, double fp[4]) {
    hipLaunchKernelGGL(k_fingerprint, dim3(kFpRounds), dim3(256), 0, rt().stream, nelem, npoin_w, connec, He, gpvol,
                       lpoin_w, Ml, buf);
    double h[4 * kFpRounds];
    SOD_HIP(hipMemcpyAsync(h, buf, sizeof(h), hipMemcpyDeviceToHost, rt().stream));
    SOD_HIP(hipStreamSynchronize(rt().stream));
    for (int t = 0; t < 4; ++t) {
        double v = 0.0;
        for (int r = 0; r < kFpRounds; ++r) v += h[4 * r + t];
        fp[t] = v;
    }
    return SOD_OK;
}

// The inverse connectivity holds element-node positions e*64 + lane as int32.
static int check_mesh_size(int nelem) {
    if ((long long)nelem * 64 > (long long)INT_MAX)
        return fail(SOD_ERR_ARG, "nelem %d: at most %d elements per mesh (32-bit element-node positions)", nelem,
                    INT_MAX / 64);
    return SOD_OK;
}

// A traversal order for the stage kernel when the caller's element numbering
// keeps no locality (fewer than half of the consecutive elements share a
// node, as after a random permutation): breadth first over the elements that
// share nodes, from element 0 and then from each element not yet reached, so
// that elements processed close in time gather many of the same nodes while
// those are still in the caches.  From the lane-layout connectivity cl and
// its inverse (off, pos) alone.  Returns false, leaving the caller's order,
// when that order keeps its locality.  The element-node residuals are then
// stored at their traversal position, so that the node update also reads
// neighbouring elements' records from nearby memory, and summed through node
// lists kept in element order: the results are the same bits either way.
static bool locality_order(const std::vector<int>& cl, const std::vector<int>& off, const std::vector<int>& pos,
                           int nelem, int npoin, std::vector<int>& order) {
    if (nelem < 2) return false;
    std::vector<int> stamp((size_t)npoin, -1);
    long shared = 0;
    for (int e = 0; e + 1 < nelem; ++e) {
        const int* a = cl.data() + (size_t)e * 64;
        for (int l = 0; l < 64; ++l) stamp[a[l]] = e;
        bool hit = false;
        for (int l = 0; l < 64 && !hit; ++l) hit = stamp[a[64 + l]] == e;
        shared += hit;
    }
    if (2 * shared >= (long)(nelem - 1)) return false;
    order.clear();
    order.reserve((size_t)nelem);
    std::vector<char> seen((size_t)nelem, 0);
    for (int s0 = 0; s0 < nelem; ++s0) {
        if (seen[s0]) continue;
        seen[s0] = 1;
        order.push_back(s0);
        for (size_t h = order.size() - 1; h < order.size(); ++h) {
            const int* a = cl.data() + (size_t)order[h] * 64;
            for (int l = 0; l < 64; ++l)
                for (int q = off[a[l]]; q < off[a[l] + 1]; ++q) {
                    const int f = pos[q] >> 6;
                    if (!seen[f]) {
                        seen[f] = 1;
                        order.push_back(f);
                    }
                }
        }
    }
    return true;
}

static int prepare(Prepared& P, int nelem, int npoin, int npoin_w, const int* connec, const int* lpoin_w,
                   const double* He, const double* gpvol, const double* Ml) {
    SOD_TRY(check_mesh_size(nelem));
    const void* key[6] = {connec, lpoin_w, He, gpvol, Ml, nullptr};
    bool same = P.nelem == nelem && P.npoin == npoin && P.npoin_w == npoin_w;
    for (int t = 0; t < 5 && same; ++t) same = P.key[t] == key[t];
    double fp[4] = {0, 0, 0, 0};
    if (same && nelem > 0 && npoin_w > 0) {
        SOD_TRY(fingerprint(nelem, npoin_w, connec, He, gpvol, lpoin_w, Ml, P.fpbuf, fp));
        if (std::memcmp(fp, P.fp, sizeof(fp)) == 0) return SOD_OK;
    }
    SOD_HIP(hipStreamSynchronize(rt().stream));
    P.release();
    const size_t ne = (size_t)nelem, np = (size_t)npoin;
    SOD_HIP(hipMalloc(&P.conn, sizeof(int) * ne * 64 + 4));
    SOD_HIP(hipMalloc(&P.He, sizeof(double) * ne * 576 + 8));
    SOD_HIP(hipMalloc(&P.gpv, sizeof(double) * ne * 64 + 8));
    SOD_HIP(hipMalloc(&P.mue, sizeof(double) * ne * 64 + 8));
    SOD_HIP(hipMalloc(&P.musgs, sizeof(double) * ne * 64 + 8));
    SOD_HIP(hipMalloc(&P.flag, sizeof(int) * 4));
    SOD_HIP(hipMalloc(&P.dscal, sizeof(double) * 8));
    SOD_HIP(hipMalloc(&P.fpbuf, sizeof(double) * 4 * kFpRounds));
    P.work_doubles = np * 11 + 2 * RED_BLOCKS + 64;
    SOD_HIP(hipMalloc(&P.work, sizeof(double) * P.work_doubles));
    SOD_HIP(hipMemsetAsync(P.work, 0, sizeof(double) * P.work_doubles, rt().stream));
    SOD_HIP(hipMemsetAsync(P.musgs, 0, sizeof(double) * ne * 64, rt().stream));
    SOD_HIP(hipMemsetAsync(P.flag, 0, sizeof(int) * 4, rt().stream));
    int bad = 0;
    if (nelem > 0) {
        hipLaunchKernelGGL(k_prepare_mesh, dim3(grid_for((long)nelem * 64, BLK)), dim3(BLK), 0, rt().stream,
                           nelem, connec, He, gpvol, P.conn, P.He, P.gpv);
        hipLaunchKernelGGL(k_check_conn, dim3(std::min(1024, grid_for((long)nelem * 64, 256))), dim3(256), 0,
                           rt().stream, (long)nelem * 64, P.conn, npoin, P.flag);
        SOD_HIP(hipMemcpyAsync(&bad, P.flag, sizeof(int), hipMemcpyDeviceToHost, rt().stream));
    }
    // slave list: nodes that are not in lpoin_w
    std::vector<int> hw(npoin_w);
    SOD_HIP(hipStreamSynchronize(rt().stream));
    if (bad) {
        P.release();
        return fail(SOD_ERR_ARG, "connec holds an id outside 1..npoin");
    }
    if (npoin_w > 0) SOD_HIP(hipMemcpy(hw.data(), lpoin_w, sizeof(int) * npoin_w, hipMemcpyDeviceToHost));
    for (int t = 0; t < npoin_w; ++t)
        if (hw[t] < 1 || hw[t] > npoin) {
            P.release();
            return fail(SOD_ERR_ARG, "lpoin_w holds an id outside 1..npoin");
        }
    {
        // node -> (element, lane) positions of the lane-layout connectivity,
        // by a counting sort that keeps the positions ascending
        const size_t nl = ne * 64;
        std::vector<int> cl(nl), off(np + 1, 0), pos(nl);
        if (nl > 0) SOD_HIP(hipMemcpy(cl.data(), P.conn, sizeof(int) * nl, hipMemcpyDeviceToHost));
        for (size_t t = 0; t < nl; ++t) ++off[(size_t)cl[t] + 1];
        for (size_t n = 0; n < np; ++n) off[n + 1] += off[n];
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (size_t t = 0; t < nl; ++t) pos[fill[cl[t]]++] = (int)t;
        std::vector<int> order;
        if (locality_order(cl, off, pos, nelem, npoin, order)) {
            SOD_HIP(hipMalloc(&P.eorder, sizeof(int) * ne));
            SOD_HIP(hipMemcpy(P.eorder, order.data(), sizeof(int) * ne, hipMemcpyHostToDevice));
        }
        SOD_HIP(hipMalloc(&P.inv_off, sizeof(int) * (np + 1)));
        SOD_HIP(hipMalloc(&P.inv_pos, sizeof(int) * (nl + 1)));
        SOD_HIP(hipMalloc(&P.relem, sizeof(double) * (nl * NRES + 2)));
        SOD_HIP(hipMemcpy(P.inv_off, off.data(), sizeof(int) * (np + 1), hipMemcpyHostToDevice));
        if (nl > 0) SOD_HIP(hipMemcpy(P.inv_pos, pos.data(), sizeof(int) * nl, hipMemcpyHostToDevice));
        // with a traversal order the stage kernel stores the element at
        // position r at r*64 + lane: the node lists it is summed through keep
        // their element order, so the sums are the same bits
        P.inv_pos_r = P.inv_pos;
        if (P.eorder != nullptr) {
            std::vector<int> rank(ne);
            for (size_t r = 0; r < ne; ++r) rank[(size_t)order[r]] = (int)r;
            for (size_t q = 0; q < nl; ++q) pos[q] = rank[(size_t)(pos[q] >> 6)] * 64 + (pos[q] & 63);
            SOD_HIP(hipMalloc(&P.inv_pos_r, sizeof(int) * (nl + 1)));
            if (nl > 0) SOD_HIP(hipMemcpy(P.inv_pos_r, pos.data(), sizeof(int) * nl, hipMemcpyHostToDevice));
        }
        // the same lists in lpoin_w order
        std::vector<int> woff((size_t)npoin_w + 1, 0);
        bool increasing = true;
        for (int t = 0; t < npoin_w; ++t) {
            const int n = hw[t] - 1;
            woff[t + 1] = woff[t] + (off[n + 1] - off[n]);
            if (t > 0 && hw[t] <= hw[t - 1]) increasing = false;
        }
        SOD_HIP(hipMalloc(&P.woff, sizeof(int) * ((size_t)npoin_w + 1)));
        if (increasing && (size_t)woff[npoin_w] == nl) {
            // the working nodes' lists follow each other in inv_pos_r (nodes
            // outside lpoin_w hold no positions: connec never names them)
            for (int t = 0; t <= npoin_w; ++t) woff[t] += npoin_w > 0 ? off[hw[0] - 1] : 0;
            P.wpos = P.inv_pos_r;
        } else {
            std::vector<int> wp((size_t)woff[npoin_w] + 1);
            for (int t = 0; t < npoin_w; ++t) {
                const int n = hw[t] - 1;
                std::copy(pos.begin() + off[n], pos.begin() + off[n + 1], wp.begin() + woff[t]);
            }
            SOD_HIP(hipMalloc(&P.wpos, sizeof(int) * wp.size()));
            SOD_HIP(hipMemcpy(P.wpos, wp.data(), sizeof(int) * wp.size(), hipMemcpyHostToDevice));
        }
        SOD_HIP(hipMemcpy(P.woff, woff.data(), sizeof(int) * woff.size(), hipMemcpyHostToDevice));
    }
    std::vector<char> isw(npoin, 0);
    for (int t = 0; t < npoin_w; ++t) isw[hw[t] - 1] = 1;
    std::vector<int> sl;
    for (int t = 0; t < npoin; ++t)
        if (!isw[t]) sl.push_back(t);
    P.nslave = (int)sl.size();
    SOD_HIP(hipMalloc(&P.slaves, sizeof(int) * (sl.size() + 1)));
    if (!sl.empty()) SOD_HIP(hipMemcpy(P.slaves, sl.data(), sizeof(int) * sl.size(), hipMemcpyHostToDevice));
    if (nelem > 0 && npoin_w > 0) SOD_TRY(fingerprint(nelem, npoin_w, connec, He, gpvol, lpoin_w, Ml, P.fpbuf, P.fp));
    SOD_HIP(hipGetLastError());
    for (int t = 0; t < 5; ++t) P.key[t] = key[t];
    P.nelem = nelem;
    P.npoin = npoin;
    P.npoin_w = npoin_w;
    return SOD_OK;
}

// The prepared mesh of these arguments from the drop-in cache (built or
// rebuilt by prepare() as needed).
static int prepare_cached(Prepared** out, int nelem, int npoin, int npoin_w, const int* connec,
                          const int* lpoin_w, const double* He, const double* gpvol, const double* Ml) {
    SOD_TRY(check_mesh_size(nelem));
    const void* key[5] = {connec, lpoin_w, He, gpvol, Ml};
    PrepSlot* hit = nullptr;
    for (auto& sl : g_slots) {
        bool same = sl.used != 0 && sl.P.nelem == nelem && sl.P.npoin == npoin && sl.P.npoin_w == npoin_w;
        for (int t = 0; t < 5 && same; ++t) same = sl.P.key[t] == key[t];
        if (same) {
            hit = &sl;
            break;
        }
    }
    const size_t need = prepared_bytes(nelem, npoin);
    if (hit == nullptr) {
        // evict least recently used meshes until the new one fits
        while (true) {
            size_t total = 0;
            PrepSlot* lru = nullptr;
            PrepSlot* empty = nullptr;
            for (auto& sl : g_slots) {
                if (sl.used == 0) {
                    if (!empty) empty = &sl;
                    continue;
                }
                total += sl.bytes;
                if (!lru || sl.used < lru->used) lru = &sl;
            }
            if (empty && (total + need <= kPrepBytes || lru == nullptr)) {
                hit = empty;
                break;
            }
            SOD_HIP(hipStreamSynchronize(rt().stream));   // queued work may still use it
            lru->P.release();
            *lru = PrepSlot();
        }
    }
    hit->used = ++g_prep_clock;
    hit->bytes = need;
    const int rc = prepare(hit->P, nelem, npoin, npoin_w, connec, lpoin_w, He, gpvol, Ml);
    if (rc != SOD_OK) {
        *hit = PrepSlot();
        return rc;
    }
    *out = &hit->P;
    return SOD_OK;
}

// mu_e from the caller's layout into P's lane layout, and mu_sgs too when it
// holds a non-zero value (returned in has_sgs; host synchronisation).  The
// kernels read P.musgs only in that case, so an all-zero mu_sgs, the usual
// one, costs a read and no write.
static int load_viscosities(Prepared& P, int nelem, const double* mu_e, const double* mu_sgs, int* has_sgs) {
    hipStream_t s = rt().stream;
    *has_sgs = 0;
    if (nelem <= 0) return SOD_OK;
    const dim3 grid(grid_for(nelem, 16));
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, P.flag);
    hipLaunchKernelGGL(k_to_lane_tiled, grid, dim3(256), 0, s, nelem, mu_e, P.mue, mu_sgs, (double*)nullptr,
                       P.flag);
    SOD_HIP(hipMemcpyAsync(has_sgs, P.flag, sizeof(int), hipMemcpyDeviceToHost, s));
    SOD_HIP(hipStreamSynchronize(s));
    if (*has_sgs)
        hipLaunchKernelGGL(k_to_lane_tiled, grid, dim3(256), 0, s, nelem, mu_sgs, P.musgs, (const double*)nullptr,
                           (double*)nullptr, (int*)nullptr);
    return SOD_OK;
}

void launch_eta_norm(int npoin_w, const int* lpoin_w, const double* eta, const double* partial, int npartial,
                     double* work);

// Level-2 views of the rk_4_main arguments (device addresses).
struct FusedArgs {
    int nelem, npoin, npoin_w;
    const int *connec, *lpoin_w;
    const double *He, *gpvol, *Ml, *helem, *helem_l, *mu_factor;
    double Rgas, gamma_gas, Cp, Prt;
    double *rho, *u, *q, *pr, *E, *Tem, *e_int, *eta, *csound, *machno, *mu_fluid, *mu_e, *mue_l;
    const double* mu_sgs;
};

// Enqueues one fused RK4 step on the library stream without any host
// synchronisation (so that it can be captured in a graph): mu_e / mu_sgs are
// read from P's lane-layout copies, dt from device memory.  With
// `mue_lane_out` the new entropy viscosity is also left in lane layout in P
// for the next step.
// A flattened ensemble of independent instances (elements and nodes of
// instance i contiguous): per-element / per-node instance ids, the offsets of
// each instance's working nodes in lpoin_w, and device room for its entropy
// statistics.  ninst == 0 means one plain mesh.
struct BatchInfo {
    int ninst = 0;
    const int* elem_inst = nullptr;
    const int* node_inst = nullptr;
    const int* woff = nullptr;
    double* stats = nullptr;   // 2 * ninst
    double* norms = nullptr;   // ninst
    int max_w = 0;             // working nodes of the largest instance
};

// The CFL time step of the next step folded into the final update (single
// mesh without sgs viscosity): per-node h_min and where the minimum goes.
struct DtFold {
    const double* hmin;
    double* dt_bits;
    double cfl_conv, cfl_diff;
    int has_diff;
};

// ---------------------------------------------------- stage configurations
// The stage kernel's tile (elements = wavefronts per workgroup) and assembly
// strategy: element-node stores summed per node in a fixed order (pull),
// device-scope atomic scatter into node arrays (its summation order varies
// from run to run), or colour-ordered scatter (one launch per element
// colour, plain read-add-write, a fixed order).  Entry 0 is the default (the
// fastest at 64^3 TGV: a one-wavefront workgroup releases its LDS as soon as
// its element is done); sod_set_autotune lets the library time the others on
// the caller's mesh and keep the fastest.
// nw == 0: the persistent one-wavefront kernel (k_stage_rhs<..., PERSIST>)
// with element-node stores.
// lds: the LDS slot layout of the element's nodal values (LDS_SWIZZLED or
// LDS_LINEAR).
struct StageCfg {
    int nw;
    int assembly;   // ASM_PULL, ASM_ATOMIC, ASM_COLOR
    int lds;
};
constexpr StageCfg kStageCfgs[] = {{1, ASM_PULL, LDS_SWIZZLED},  {2, ASM_PULL, LDS_SWIZZLED},
                                   {4, ASM_PULL, LDS_SWIZZLED},  {4, ASM_ATOMIC, LDS_SWIZZLED},
                                   {1, ASM_COLOR, LDS_SWIZZLED}, {0, ASM_PULL, LDS_SWIZZLED},
                                   {1, ASM_PULL, LDS_LINEAR}};
constexpr int kNumStageCfgs = sizeof(kStageCfgs) / sizeof(kStageCfgs[0]);
constexpr int kPinned = 16;   // autotune mode kPinned + c: configuration c pinned

// Element colouring for the colour-ordered assembly, from the lane-layout
// connectivity alone: elements in increasing order take the lowest colour
// that no element sharing one of their nodes holds (greedy; at most 64
// colours).  Returns false, leaving P without a colouring, when an element
// holds one node twice (its lanes would race) or 64 colours do not suffice.
// Elements are listed colour by colour, increasing within a colour.
static bool build_colouring(Prepared& P, const std::vector<int>& cl, int nelem, int npoin,
                            std::vector<int>& elist, std::vector<int>& off) {
    std::vector<unsigned long long> used((size_t)npoin, 0ULL);
    std::vector<int> stamp((size_t)npoin, -1);
    std::vector<unsigned char> col((size_t)nelem, 0);
    for (int e = 0; e < nelem; ++e) {
        unsigned long long m = 0;
        const int* c = cl.data() + (size_t)e * 64;
        for (int l = 0; l < 64; ++l) {
            if (stamp[c[l]] == e) return false;   // a node twice in one element
            stamp[c[l]] = e;
            m |= used[c[l]];
        }
        if (m == ~0ULL) return false;
        const int k = __builtin_ctzll(~m);
        for (int l = 0; l < 64; ++l) used[c[l]] |= 1ULL << k;
        col[e] = (unsigned char)k;
    }
    int nc = 0;
    for (int e = 0; e < nelem; ++e) nc = std::max(nc, (int)col[e] + 1);
    off.assign((size_t)nc + 1, 0);
    for (int e = 0; e < nelem; ++e) ++off[(size_t)col[e] + 1];
    for (int k = 0; k < nc; ++k) off[k + 1] += off[k];
    elist.resize((size_t)nelem);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int e = 0; e < nelem; ++e) elist[fill[col[e]]++] = e;
    (void)P;
    return true;
}

// Builds P's colouring on first use (host work over the lane-layout
// connectivity); SOD_ERR_UNSUPPORTED when the mesh cannot be coloured.
static int ensure_colouring(Prepared& P) {
    if (P.color_state == 1) return SOD_OK;
    if (P.color_state == 2)
        return fail(SOD_ERR_UNSUPPORTED, "colour-ordered assembly: an element holds a node twice, or more than "
                                         "64 colours would be needed");
    const size_t nl = (size_t)P.nelem * 64;
    std::vector<int> cl(nl), elist, off;
    SOD_HIP(hipStreamSynchronize(rt().stream));
    if (nl > 0) SOD_HIP(hipMemcpy(cl.data(), P.conn, sizeof(int) * nl, hipMemcpyDeviceToHost));
    if (!build_colouring(P, cl, P.nelem, P.npoin, elist, off)) {
        P.color_state = 2;
        return ensure_colouring(P);
    }
    SOD_HIP(hipMalloc(&P.color_elems, sizeof(int) * (elist.size() + 1)));
    if (!elist.empty())
        SOD_HIP(hipMemcpy(P.color_elems, elist.data(), sizeof(int) * elist.size(), hipMemcpyHostToDevice));
    P.color_off = off;
    P.color_state = 1;
    return SOD_OK;
}

// Workgroups of the persistent stage kernel: three one-wavefront workgroups
// per SIMD, the residency its register budget allows.
static int persist_grid(int nelem) {
    static int cus = 0;
    if (cus == 0) {
        hipDeviceProp_t prop;
        cus = hipGetDeviceProperties(&prop, rt().device) == hipSuccess ? prop.multiProcessorCount : 256;
    }
    return std::max(1, std::min(nelem, cus * 4 * 3));
}

template <bool HS, int NW, int AS, bool TH, int LAY = LDS_SWIZZLED>
static void launch_rhs_t(hipStream_t s, int nelem, const Prepared& P, const double* const in[6], long np,
                         const StageParams& sp) {
    if constexpr (NW == 0) {
        hipLaunchKernelGGL((k_stage_rhs<HS, 1, ASM_PULL, TH, true>), dim3(persist_grid(nelem)), dim3(64), 0, s,
                           nelem, (const int*)P.eorder, P.conn, P.He, P.gpv, P.mue, P.musgs, in[0], in[1], in[2],
                           in[3], in[4], in[5], P.relem, np, sp);
    } else if constexpr (AS == ASM_COLOR) {
        const int nc = (int)P.color_off.size() - 1;
        for (int k = 0; k < nc; ++k) {
            const int b = P.color_off[k], cnt = P.color_off[k + 1] - b;
            if (cnt > 0)
                hipLaunchKernelGGL((k_stage_rhs<HS, NW, AS, TH>), dim3(cnt), dim3(64 * NW), 0, s, cnt,
                                   P.color_elems + b, P.conn, P.He, P.gpv, P.mue, P.musgs, in[0], in[1], in[2],
                                   in[3], in[4], in[5], P.rn, np, sp);
        }
    } else {
        hipLaunchKernelGGL((k_stage_rhs<HS, NW, AS, TH, false, LAY>), dim3(grid_for(nelem, NW)), dim3(64 * NW), 0, s,
                           nelem,
                           (const int*)P.eorder, P.conn, P.He, P.gpv, P.mue, P.musgs, in[0], in[1], in[2], in[3],
                           in[4], in[5], AS == ASM_PULL ? P.relem : P.rn, np, sp);
    }
}

// stage kernel of configuration `cfg` on the state in[] = rho, q0, q1, q2, E
// and the per-node viscosity input (mu_factor, or mu_fluid with the
// Sutherland gate off)
template <bool HS, bool TH>
static void launch_rhs_hs(int cfg, hipStream_t s, int nelem, const Prepared& P, const double* const in[6], long np,
                          const StageParams& sp) {
    switch (cfg) {
        case 1: launch_rhs_t<HS, 2, ASM_PULL, TH>(s, nelem, P, in, np, sp); break;
        case 2: launch_rhs_t<HS, 4, ASM_PULL, TH>(s, nelem, P, in, np, sp); break;
        case 3: launch_rhs_t<HS, 4, ASM_ATOMIC, TH>(s, nelem, P, in, np, sp); break;
        case 4: launch_rhs_t<HS, 1, ASM_COLOR, TH>(s, nelem, P, in, np, sp); break;
        case 5: launch_rhs_t<HS, 0, ASM_PULL, TH>(s, nelem, P, in, np, sp); break;
        case 6: launch_rhs_t<HS, 1, ASM_PULL, TH, LDS_LINEAR>(s, nelem, P, in, np, sp); break;
        default: launch_rhs_t<HS, 1, ASM_PULL, TH>(s, nelem, P, in, np, sp); break;
    }
}
// total_h: full_convec_ijk_H in place of full_convec_ijk (flag_total_enthalpy)
static void launch_stage_rhs(int cfg, bool has_sgs, bool total_h, hipStream_t s, int nelem, const Prepared& P,
                             const double* const in[6], long np, const StageParams& sp) {
    if (has_sgs) {
        if (total_h)
            launch_rhs_hs<true, true>(cfg, s, nelem, P, in, np, sp);
        else
            launch_rhs_hs<true, false>(cfg, s, nelem, P, in, np, sp);
    } else {
        if (total_h)
            launch_rhs_hs<false, true>(cfg, s, nelem, P, in, np, sp);
        else
            launch_rhs_hs<false, false>(cfg, s, nelem, P, in, np, sp);
    }
}

template <bool FINAL, bool BATCHED>
static void launch_update_fb(bool atomic, dim3 grid, hipStream_t s, int npoin, int npoin_w, const int* lpoin_w,
                             int stage, const double* d_dt, double gamma, double Rgas, const StepState& st,
                             const int* ni) {
    if (atomic)
        hipLaunchKernelGGL((k_stage_update<FINAL, BATCHED, true>), grid, dim3(BLK), 0, s, npoin, npoin_w, lpoin_w,
                           stage, d_dt, gamma, Rgas, st, ni);
    else
        hipLaunchKernelGGL((k_stage_update<FINAL, BATCHED, false>), grid, dim3(BLK), 0, s, npoin, npoin_w, lpoin_w,
                           stage, d_dt, gamma, Rgas, st, ni);
}
static void launch_update(bool final_stage, bool batched, bool atomic, dim3 grid, hipStream_t s, int npoin,
                          int npoin_w, const int* lpoin_w, int stage, const double* d_dt, double gamma,
                          double Rgas, const StepState& st, const int* ni) {
    if (final_stage) {
        if (batched)
            launch_update_fb<true, true>(atomic, grid, s, npoin, npoin_w, lpoin_w, stage, d_dt, gamma, Rgas, st, ni);
        else
            launch_update_fb<true, false>(atomic, grid, s, npoin, npoin_w, lpoin_w, stage, d_dt, gamma, Rgas, st, ni);
    } else {
        if (batched)
            launch_update_fb<false, true>(atomic, grid, s, npoin, npoin_w, lpoin_w, stage, d_dt, gamma, Rgas, st, ni);
        else
            launch_update_fb<false, false>(atomic, grid, s, npoin, npoin_w, lpoin_w, stage, d_dt, gamma, Rgas, st,
                                           ni);
    }
}

// The node-kernel view of a step: the caller's level-2 arrays and the
// prepared mesh's work arrays (no folded CFL step).
static StepState step_state(Prepared& P, const FusedArgs& a) {
    const size_t np = (size_t)a.npoin;
    StepState st;
    st.rho = a.rho;
    st.q0 = a.q;
    st.q1 = a.q + np;
    st.q2 = a.q + 2 * np;
    st.E = a.E;
    st.u = a.u;
    st.pr = a.pr;
    st.Tem = a.Tem;
    st.e_int = a.e_int;
    st.eta = a.eta;
    st.csound = a.csound;
    st.machno = a.machno;
    st.mu_fluid = a.mu_fluid;
    st.Ml = a.Ml;
    st.mu_factor = a.mu_factor;
    st.aux = P.work;
    st.Rsum = P.work + 5 * np;
    st.partial = P.work + 11 * np;
    st.relem = P.relem;
    st.inv_off = P.inv_off;
    st.inv_pos = P.inv_pos_r;
    st.woff = P.woff;
    st.wpos = P.wpos;
    st.Rn = P.rn;
    st.hmin = nullptr;
    st.dt_bits = nullptr;
    st.cfl_conv = st.cfl_diff = 0.0;
    st.has_diff = 0;
    st.suth = sutherland_gate(rt().params) ? 1 : 0;
    return st;
}

static void enqueue_fused_step(Prepared& P, const FusedArgs& a, const double* d_dt, bool has_sgs,
                               bool mue_lane_out, const BatchInfo* B = nullptr, const DtFold* fold = nullptr) {
    hipStream_t s = rt().stream;
    const sod_numerical_params& prm = rt().params;
    const int nelem = a.nelem, npoin = a.npoin, npoin_w = a.npoin_w;
    const size_t np = (size_t)npoin;
    StepState st = step_state(P, a);
    if (fold != nullptr && B == nullptr && !has_sgs) {
        st.hmin = fold->hmin;
        st.dt_bits = reinterpret_cast<unsigned long long*>(fold->dt_bits);
        st.cfl_conv = fold->cfl_conv;
        st.cfl_diff = fold->cfl_diff;
        st.has_diff = fold->has_diff;
    }
    const StageParams sp = make_stage_params(a.gamma_gas, a.Rgas, a.Cp, a.Prt, prm);
    const int ugrid = std::min(RED_BLOCKS, std::max(1, grid_for(npoin_w, BLK)));
    for (int stage = 0; stage < 4; ++stage) {
        const double *ir, *i0, *i1, *i2, *iE;
        if (stage == 0) {
            ir = st.rho; i0 = st.q0; i1 = st.q1; i2 = st.q2; iE = st.E;
        } else {
            ir = st.aux; i0 = st.aux + np; i1 = st.aux + 2 * np; i2 = st.aux + 3 * np; iE = st.aux + 4 * np;
        }
        if (nelem > 0) {
            SOD_PROF("stage_rhs");
            const double* const in[6] = {ir, i0, i1, i2, iE, sp.suth ? a.mu_factor : a.mu_fluid};
            launch_stage_rhs(P.cfg, has_sgs, prm.flag_total_enthalpy != 0, s, nelem, P, in, (long)np, sp);
        }
        SOD_PROF("stage_update");
        const dim3 grid(stage == 3 ? ugrid : node_grid(npoin_w));
        launch_update(stage == 3, B != nullptr, kStageCfgs[P.cfg].assembly != ASM_PULL, grid, s, npoin, npoin_w,
                      a.lpoin_w, stage,
                      d_dt, a.gamma_gas, a.Rgas, st, B ? B->node_inst : nullptr);
    }
    if (P.nslave > 0 && st.suth)
        hipLaunchKernelGGL(k_slave_sutherland, dim3(grid_for(P.nslave, BLK)), dim3(BLK), 0, s, P.nslave,
                           P.slaves, a.Tem, a.mu_factor, a.mu_fluid);
    // entropy viscosity for the next step (time_integ.f90:492-493)
    {
        SOD_PROF("smart_visc_fused");
        double* vwork = st.partial + RED_BLOCKS;
        const double* d_norm = vwork + 1;
        if (prm.flag_normalise_entropy == 1) {
            if (B) {
                const dim3 g2(std::max(1, std::min(64, grid_for(B->max_w, 256))), B->ninst);
                hipLaunchKernelGGL(k_zero_stats, dim3(1), dim3(256), 0, s, 2 * B->ninst, B->stats);
                hipLaunchKernelGGL(k_inst_eta_sum, g2, dim3(256), 0, s, B->woff, a.lpoin_w, a.eta, B->stats);
                hipLaunchKernelGGL(k_inst_eta_dev, g2, dim3(256), 0, s, B->woff, a.lpoin_w, a.eta, B->stats);
                hipLaunchKernelGGL(k_inst_norms, dim3(1), dim3(256), 0, s, B->ninst, B->stats, B->norms);
                d_norm = B->norms;
            } else {
                launch_eta_norm(npoin_w, a.lpoin_w, a.eta, st.partial, ugrid, vwork);
            }
        }
        if (nelem > 0)
            hipLaunchKernelGGL(k_smart_visc_lane, dim3(grid_for(nelem, 16)), dim3(256), 0, s, nelem, npoin, P.conn,
                               P.gpv, a.helem_l, a.helem, st.aux, a.gamma_gas, a.rho, a.u, a.csound, a.Tem,
                               d_norm, prm.flag_normalise_entropy == 1, prm.Rgas_inf, prm.ce, prm.cmax,
                               prm.cglob, a.mu_e, a.mue_l, mue_lane_out ? P.mue : nullptr,
                               B ? B->elem_inst : nullptr);
    }
}

static_assert(kNumStageCfgs == 7 && kStageCfgs[0].nw == 1 && kStageCfgs[0].assembly == ASM_PULL &&
                  kStageCfgs[1].nw == 2 && kStageCfgs[1].assembly == ASM_PULL && kStageCfgs[2].nw == 4 &&
                  kStageCfgs[2].assembly == ASM_PULL && kStageCfgs[3].nw == 4 &&
                  kStageCfgs[3].assembly == ASM_ATOMIC && kStageCfgs[4].nw == 1 &&
                  kStageCfgs[4].assembly == ASM_COLOR && kStageCfgs[5].nw == 0 &&
                  kStageCfgs[5].assembly == ASM_PULL && kStageCfgs[6].nw == 1 &&
                  kStageCfgs[6].assembly == ASM_PULL && kStageCfgs[6].lds == LDS_LINEAR,
              "launch_rhs_hs dispatches on these entries");

// Node residual arrays of the scatter configurations (atomic and
// colour-ordered), zeroed once; the node update re-zeroes what it reads.
static int ensure_node_residuals(Prepared& P, size_t np, hipStream_t s) {
    if (P.rn != nullptr) return SOD_OK;
    SOD_HIP(hipMalloc(&P.rn, sizeof(double) * NRES * np + 8));
    SOD_HIP(hipMemsetAsync(P.rn, 0, sizeof(double) * NRES * np, s));
    return SOD_OK;
}

// Sets P.cfg for the autotune mode (sod_set_autotune): 0 keeps the default
// configuration; 1 takes the fastest configuration whose summation order is
// fixed (element-node stores, or colour-ordered scatter when the mesh can be
// coloured); 2 the fastest of all, atomic scatter included; kPinned + c pins
// entry c (sod_set_stage_config).  Each candidate runs the
// first stage of a step on the caller's level-2 state -- the stage kernel and
// the node update that sums its residuals -- once to warm up and three times
// timed with HIP events on the library stream; the smallest time wins.  Only
// library work arrays are written.  Runs once per prepared mesh and mode,
// never inside a graph capture.
static int choose_stage_cfg(Prepared& P, const FusedArgs& a, bool has_sgs, const double* d_dt,
                            const BatchInfo* B) {
    const int mode = rt().autotune;
    if (P.cfg_mode == mode) return SOD_OK;
    hipStream_t s = rt().stream;
    const size_t np = (size_t)a.npoin;
    if (mode == 0 || mode >= kPinned) {
        const int c = mode == 0 ? 0 : mode - kPinned;
        if (kStageCfgs[c].assembly == ASM_COLOR) SOD_TRY(ensure_colouring(P));
        if (kStageCfgs[c].assembly != ASM_PULL) SOD_TRY(ensure_node_residuals(P, np, s));
        P.cfg = c;
        P.cfg_mode = mode;
        P.cfg_ms = 0.0;
        return SOD_OK;
    }
    const bool try_atomic = mode >= 2;
    // a mesh that cannot be coloured just drops that candidate
    const bool try_color = ensure_colouring(P) == SOD_OK;
    SOD_TRY(ensure_node_residuals(P, np, s));
    const StepState st = step_state(P, a);
    const sod_numerical_params& prm = rt().params;
    const StageParams sp = make_stage_params(a.gamma_gas, a.Rgas, a.Cp, a.Prt, prm);
    const double* const in[6] = {a.rho, a.q, a.q + np, a.q + 2 * np, a.E, sp.suth ? a.mu_factor : a.mu_fluid};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SOD_HIP(hipEventCreate(&e0));
    hipError_t err = hipEventCreate(&e1);
    double best = HUGE_VAL;
    int best_cfg = 0;
    for (int c = 0; c < kNumStageCfgs && err == hipSuccess; ++c) {
        if (kStageCfgs[c].assembly == ASM_ATOMIC && !try_atomic) continue;
        if (kStageCfgs[c].assembly == ASM_COLOR && !try_color) continue;
        double tmin = HUGE_VAL;
        for (int rep = 0; rep < 4 && err == hipSuccess; ++rep) {
            err = hipEventRecord(e0, s);
            if (a.nelem > 0)
                launch_stage_rhs(c, has_sgs, prm.flag_total_enthalpy != 0, s, a.nelem, P, in, (long)np, sp);
            launch_update(false, B != nullptr, kStageCfgs[c].assembly != ASM_PULL, dim3(node_grid(a.npoin_w)), s,
                          a.npoin,
                          a.npoin_w, a.lpoin_w,
                          0, d_dt, a.gamma_gas, a.Rgas, st, B ? B->node_inst : nullptr);
            if (err == hipSuccess) err = hipGetLastError();
            if (err == hipSuccess) err = hipEventRecord(e1, s);
            if (err == hipSuccess) err = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0) tmin = std::min(tmin, (double)ms);
        }
        if (tmin < best) {
            best = tmin;
            best_cfg = c;
        }
    }
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (err != hipSuccess) return fail(SOD_ERR_HIP, "stage autotuning: %s", hipGetErrorString(err));
    P.cfg = best_cfg;
    P.cfg_mode = mode;
    P.cfg_ms = best;
    if (kStageCfgs[best_cfg].assembly == ASM_PULL && P.rn != nullptr) {
        SOD_HIP(hipStreamSynchronize(s));
        (void)hipFree(P.rn);
        P.rn = nullptr;
    }
    return SOD_OK;
}

int rk4_step_fused(int nelem, int npoin, int npoin_w, const int* connec, const int* lpoin_w,
                   const double* He, const double* gpvol, const double* Ml, const double* helem,
                   const double* helem_l, double dt, double Rgas, double gamma_gas, double Cp, double Prt,
                   const double* mu_factor, double* rho, double* u, double* q, double* pr, double* E,
                   double* Tem, double* e_int, double* eta, double* csound, double* machno,
                   double* mu_fluid, double* mu_e, const double* mu_sgs, double* mue_l) {
    Prepared* Pp = nullptr;
    SOD_TRY(prepare_cached(&Pp, nelem, npoin, npoin_w, connec, lpoin_w, He, gpvol, Ml));
    Prepared& P = *Pp;
    int has_sgs = 0;
    SOD_TRY(load_viscosities(P, nelem, mu_e, mu_sgs, &has_sgs));
    const FusedArgs a{nelem, npoin, npoin_w, connec, lpoin_w, He, gpvol, Ml, helem, helem_l, mu_factor, Rgas,
                      gamma_gas, Cp, Prt, rho, u, q, pr, E, Tem, e_int, eta, csound, machno, mu_fluid, mu_e,
                      mue_l, mu_sgs};
    hipLaunchKernelGGL(k_set_double_f, dim3(1), dim3(64), 0, rt().stream, P.dscal, dt);
    SOD_TRY(choose_stage_cfg(P, a, has_sgs != 0, P.dscal, nullptr));
    enqueue_fused_step(P, a, P.dscal, has_sgs != 0, false);
    SOD_HIP(hipGetLastError());
    return SOD_OK;
}

int prepare_lane_mesh(int nelem, int npoin, int npoin_w, const int* connec, const int* lpoin_w,
                      const double* He, const double* gpvol, const double* Ml, const double* mu_e,
                      const double* mu_sgs, LaneMesh* out) {
    Prepared* Pp = nullptr;
    SOD_TRY(prepare_cached(&Pp, nelem, npoin, npoin_w, connec, lpoin_w, He, gpvol, Ml));
    Prepared& P = *Pp;
    int has_sgs = 0;
    SOD_TRY(load_viscosities(P, nelem, mu_e, mu_sgs, &has_sgs));
    *out = LaneMesh{P.conn, P.He, P.gpv, P.mue, P.musgs, has_sgs != 0, P.inv_off, P.inv_pos, P.relem};
    return SOD_OK;
}

// One evaluation of the full right-hand side at a conserved state: the
// assembled (not yet lumped-solved) residuals R_conv + R_diff of mass,
// momentum and energy and the entropy residual, as rk_4_main forms them in
// one stage (time_integ.f90:266-373).  Out arrays are overwritten.
int full_rhs(int nelem, int npoin, int npoin_w, const int* connec, const int* lpoin_w, const double* He,
             const double* gpvol, const double* Ml, const double* rho, const double* q, const double* E,
             const double* mu_factor, const double* mu_e, const double* mu_sgs, double Rgas, double gamma_gas,
             double Cp, double Prt, double* Rmass, double* Rmom, double* Rener, double* Reta) {
    Prepared* Pp = nullptr;
    SOD_TRY(prepare_cached(&Pp, nelem, npoin, npoin_w, connec, lpoin_w, He, gpvol, Ml));
    Prepared& Pm = *Pp;
    hipStream_t s = rt().stream;
    const sod_numerical_params& prm = rt().params;
    const size_t np = (size_t)npoin;
    int has_sgs = 0;
    SOD_TRY(load_viscosities(Pm, nelem, mu_e, mu_sgs, &has_sgs));
    SOD_PROF("full_rhs");   // the element kernel and the node sums
    if (nelem > 0) {
        const StageParams sp = make_stage_params(gamma_gas, Rgas, Cp, Prt, prm);
        // the default configuration (element-node stores, summed below)
        const double* const in[6] = {rho, q, q + np, q + 2 * np, E, mu_factor};
        launch_stage_rhs(0, has_sgs != 0, prm.flag_total_enthalpy != 0, s, nelem, Pm, in, (long)np, sp);
    }
    if (npoin > 0)
        hipLaunchKernelGGL(k_assemble_nodes, dim3(std::max(1, grid_for(npoin, 256))), dim3(256), 0, s, npoin,
                           Pm.inv_off, Pm.inv_pos_r, Pm.relem, Rmass, Rmom, Rener, Reta);
    SOD_HIP(hipGetLastError());
    return SOD_OK;
}


// ------------------------------------------------------- whole-step driver
// Device-side pieces of the reference's time loop
// (CFDSolverBase_evalTimeIteration, CFDSolverBase.f90:1501-1694) around
// rk_4_main.  All of them read their scalars from device memory so that a
// step can be captured once in a graph and replayed.

constexpr int kMaxCopy = 12;
struct CopyList {
    const double* src[kMaxCopy];
    double* dst[kMaxCopy];
};

// dst[y][i] = src[y][i], i < n, for the blockIdx.y-th pair
__global__ __launch_bounds__(256) void k_copy_fields(long n, CopyList c) {
    const double* __restrict__ s = c.src[blockIdx.y];
    double* __restrict__ d = c.dst[blockIdx.y];
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) d[i] = s[i];
}

// loop state on the device: [0] dt, [1] time, [2] EK scratch; ints: [0]
// istep of the step being taken, [1] EK values recorded in this run
__global__ void k_loop_begin(int* __restrict__ ist) {
    if (threadIdx.x == 0) ist[0] += 1;
}

__global__ void k_copy_double(const double* __restrict__ src, double* __restrict__ dst) {
    if (threadIdx.x == 0) *dst = *src;
}

// sc[0..n) = dt of each instance, sc[n..2n) = its time
__global__ void k_loop_time(double* __restrict__ sc, int n) {
    for (int t = threadIdx.x; t < n; t += blockDim.x) sc[n + t] = sc[n + t] + sc[t];
}

__device__ __forceinline__ bool is_log_step(int istep, int first, int every) {
    return every > 0 && istep >= first && (istep - first) % every == 0;
}

// volAvg_EK (mod_analysis.f90:10-44) on log steps only; same sums as the
// drop-in entry point
__global__ __launch_bounds__(256) void k_ek_partial_log(int nelem, int npoin, const int* __restrict__ connec,
                                                        const double* __restrict__ gpvol,
                                                        const double* __restrict__ Ngp,
                                                        const double* __restrict__ rho,
                                                        const double* __restrict__ u, const int* __restrict__ ist,
                                                        int first, int every, double* __restrict__ partial) {
    if (!is_log_step(ist[0], first, every)) return;
    __shared__ double s_tmp[4];
    const int le = threadIdx.x >> 6, l = threadIdx.x & 63;
    const size_t np = (size_t)npoin, ne = (size_t)nelem;
    double acc = 0.0;
    for (int e = blockIdx.x * 4 + le; e < nelem; e += gridDim.x * 4) {
        const int n = connec[(size_t)e + ne * l] - 1;
        const double u0 = u[n], u1 = u[n + np], u2 = u[n + 2 * np];
        const double ke = 0.5 * rho[n] * (u0 * u0 + u1 * u1 + u2 * u2);
        const double* gv = gpvol + 64 * (size_t)e;
        for (int g = 0; g < 64; ++g) acc += gv[g] * Ngp[g + 64 * l] * ke;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (l == 0) s_tmp[le] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (s_tmp[0] + s_tmp[1]) + (s_tmp[2] + s_tmp[3]);
}

__global__ __launch_bounds__(256) void k_ek_final_log(int npartial, const double* __restrict__ partial,
                                                      int* __restrict__ ist, int first, int every, double rho0,
                                                      int cap, double* __restrict__ hist) {
    if (!is_log_step(ist[0], first, every)) return;
    __shared__ double s_tmp[4];
    double v = 0.0;
    for (int i = threadIdx.x; i < npartial; i += 256) v += partial[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) s_tmp[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double sum = (s_tmp[0] + s_tmp[1]) + (s_tmp[2] + s_tmp[3]);
        const int k = ist[1];
        // mod_analysis.f90:37 uses 3.14159 for pi
        if (k < cap) hist[k] = sum / (rho0 * pow(2.0 * 3.14159, 3.0));
        ist[1] = k + 1;
    }
}

void launch_adapt_dt(int nelem, int npoin, const int* connec, const double* helem, const double* u,
                     const double* csound, double cfl_conv, int has_diff, double cfl_diff,
                     const double* mu_fluid, const double* mu_sgs, const double* rho, double* d_dt_bits);
void launch_hmin(int nelem, int npoin, const int* connec, const double* helem, double* hmin);
void launch_adapt_dt_nodes_inst(int npoin, const double* hmin, const double* u, const double* csound,
                                double cfl_conv, int has_diff, double cfl_diff, const double* mu_fluid,
                                const double* rho, const int* node_inst, int ninst, double* d_dt_bits);
void launch_adapt_dt_inst(int nelem, int npoin, const int* connec, const double* helem, const double* u,
                          const double* csound, double cfl_conv, int has_diff, double cfl_diff,
                          const double* mu_fluid, const double* mu_sgs, const double* rho, const int* elem_inst,
                          int ninst, double* d_dt_bits);

struct LoopCfg {
    bool valid = false;
    bool has_sgs = false, fixed_dt = false, keep_levels = false;
    double cfl_conv = 0, cfl_diff = 0, rho0 = 0;
    int log_first = 0, log_every = 0, ek_cap = 0, has_diff = 0;
    int stage_cfg = 0;
    sod_numerical_params params{};
    bool operator==(const LoopCfg& o) const {
        return valid == o.valid && has_sgs == o.has_sgs && fixed_dt == o.fixed_dt && stage_cfg == o.stage_cfg &&
               keep_levels == o.keep_levels && cfl_conv == o.cfl_conv && cfl_diff == o.cfl_diff &&
               rho0 == o.rho0 && log_first == o.log_first && log_every == o.log_every && ek_cap == o.ek_cap &&
               has_diff == o.has_diff && std::memcmp(&params, &o.params, sizeof(params)) == 0;
    }
};

}  // namespace sod

struct sod_time_loop {
    sod::Prepared P;
    sod::FusedArgs a;       // level-2 views
    double *rho, *u, *q, *pr, *E, *Tem, *e_int, *eta;   // full level arrays
    double* Ngp = nullptr;  // device copy of the library's Ngp
    double* sc = nullptr;   // dt[ninst], time[ninst], next dt[ninst] (folded CFL)
    double* hmin = nullptr;  // per-node minimum element size
    int* ist = nullptr;     // istep, EK count
    int ninst = 1;          // > 1: a flattened ensemble described by B
    sod::BatchInfo B;
    int* inst_ids = nullptr;     // elem_inst (nelem) then node_inst (npoin)
    int* woff = nullptr;         // ninst + 1
    double* inst_stats = nullptr;   // 3 * ninst
    double* ek_partial = nullptr;
    double* ek_hist = nullptr;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    sod::LoopCfg cfg;
};

namespace sod {

static void loop_free_graph(sod_time_loop* L) {
    if (L->exec) (void)hipGraphExecDestroy(L->exec);
    if (L->graph) (void)hipGraphDestroy(L->graph);
    L->exec = nullptr;
    L->graph = nullptr;
    L->cfg.valid = false;
}

// the kernels of one time-loop step, enqueued on the library stream
static void loop_enqueue_step(sod_time_loop* L, const LoopCfg& c) {
    hipStream_t s = rt().stream;
    const FusedArgs& a = L->a;
    const size_t np = (size_t)a.npoin;
    const dim3 cgrid(std::max(1, std::min(256, grid_for((long)np, 256))), 1);
    hipLaunchKernelGGL(k_loop_begin, dim3(1), dim3(64), 0, s, L->ist);
    if (c.keep_levels) {
        // level 1 <- level 2 (CFDSolverBase.f90:1532-1539)
        CopyList cl{};
        int k = 0;
        auto add = [&](double* base, int ncomp, int nlev) {
            for (int d = 0; d < ncomp; ++d) {
                cl.src[k] = base + np * ((size_t)ncomp * 1 + d);
                cl.dst[k] = base + np * d;
                ++k;
            }
            (void)nlev;
        };
        add(L->rho, 1, 3); add(L->u, 3, 3); add(L->q, 3, 3); add(L->pr, 1, 3); add(L->E, 1, 3);
        add(L->Tem, 1, 2); add(L->e_int, 1, 2); add(L->eta, 1, 3);
        hipLaunchKernelGGL(k_copy_fields, dim3(cgrid.x, k), dim3(256), 0, s, (long)np, cl);
    }
    // the next step's CFL dt folded into the final update (one mesh, no sgs
    // viscosity): its minimum collects in sc[2] and becomes dt after the
    // time advance
    const bool fold = !c.fixed_dt && L->ninst == 1 && !c.has_sgs && L->hmin != nullptr;
    const DtFold f{L->hmin, L->sc + 2, c.cfl_conv, c.cfl_diff, c.has_diff};
    if (fold) hipLaunchKernelGGL(k_set_double_f, dim3(1), dim3(64), 0, s, L->sc + 2, 100000000000000.0);
    enqueue_fused_step(L->P, a, L->sc, c.has_sgs, true, L->ninst > 1 ? &L->B : nullptr, fold ? &f : nullptr);
    if (c.keep_levels) {
        // level 3 <- level 1 (CFDSolverBase.f90:1552-1557)
        CopyList cl{};
        int k = 0;
        auto add = [&](double* base, int ncomp) {
            for (int d = 0; d < ncomp; ++d) {
                cl.src[k] = base + np * d;
                cl.dst[k] = base + np * ((size_t)ncomp * 2 + d);
                ++k;
            }
        };
        add(L->rho, 1); add(L->E, 1); add(L->q, 3); add(L->eta, 1); add(L->u, 3); add(L->pr, 1);
        hipLaunchKernelGGL(k_copy_fields, dim3(cgrid.x, k), dim3(256), 0, s, (long)np, cl);
    }
    hipLaunchKernelGGL(k_loop_time, dim3(1), dim3(256), 0, s, L->sc, L->ninst);
    if (c.log_every > 0) {
        hipLaunchKernelGGL(k_ek_partial_log, dim3(RED_BLOCKS), dim3(256), 0, s, a.nelem, a.npoin, a.connec, a.gpvol,
                           L->Ngp, a.rho, a.u, L->ist, c.log_first, c.log_every, L->ek_partial);
        hipLaunchKernelGGL(k_ek_final_log, dim3(1), dim3(256), 0, s, RED_BLOCKS, L->ek_partial, L->ist,
                           c.log_first, c.log_every, c.rho0, c.ek_cap, L->ek_hist);
    }
    if (!c.fixed_dt) {   // evalDt (CFDSolverBase.f90:1132-1141)
        if (fold)
            hipLaunchKernelGGL(k_copy_double, dim3(1), dim3(64), 0, s, L->sc + 2, L->sc);
        else if (L->ninst > 1 && !c.has_sgs)
            launch_adapt_dt_nodes_inst(a.npoin, L->hmin, a.u, a.csound, c.cfl_conv, c.has_diff, c.cfl_diff,
                                       a.mu_fluid, a.rho, L->B.node_inst, L->ninst, L->sc);
        else if (L->ninst > 1)
            launch_adapt_dt_inst(a.nelem, a.npoin, a.connec, a.helem, a.u, a.csound, c.cfl_conv, c.has_diff,
                                 c.cfl_diff, a.mu_fluid, a.mu_sgs, a.rho, L->B.elem_inst, L->ninst, L->sc);
        else
            launch_adapt_dt(a.nelem, a.npoin, a.connec, a.helem, a.u, a.csound, c.cfl_conv, c.has_diff,
                            c.cfl_diff, a.mu_fluid, a.mu_sgs, a.rho, L->sc);
    }
}

}  // namespace sod

using namespace sod;

extern "C" int sod_full_rhs(int nelem, int npoin, int npoin_w, const int* connec, const int* lpoin_w,
                            const double* He, const double* gpvol, const double* Ml, const double* rho,
                            const double* q, const double* E, const double* mu_factor, const double* mu_e,
                            const double* mu_sgs, double Rgas, double gamma_gas, double Cp, double Prt,
                            double* Rmass, double* Rmom, double* Rener, double* Reta) {
    SOD_TRY(ensure_init());
    if (nelem < 0 || npoin < 0 || npoin_w < 0 || npoin_w > npoin) return fail(SOD_ERR_ARG, "full_rhs: sizes");
    const int *dc, *dl;
    const double *dHe, *dgv, *dMl, *dr, *dq, *dE, *dmf, *dmue, *dsgs;
    double *dRm, *dRq, *dRe, *dRt;
    SOD_TRY(resolve_t(connec, "connec", &dc));
    SOD_TRY(resolve_t(lpoin_w, "lpoin_w", &dl));
    SOD_TRY(resolve_t(He, "He", &dHe));
    SOD_TRY(resolve_t(gpvol, "gpvol", &dgv));
    SOD_TRY(resolve_t(Ml, "Ml", &dMl));
    SOD_TRY(resolve_t(rho, "rho", &dr));
    SOD_TRY(resolve_t(q, "q", &dq));
    SOD_TRY(resolve_t(E, "E", &dE));
    SOD_TRY(resolve_t(mu_factor, "mu_factor", &dmf));
    SOD_TRY(resolve_t(mu_e, "mu_e", &dmue));
    SOD_TRY(resolve_t(mu_sgs, "mu_sgs", &dsgs));
    SOD_TRY(resolve_w(Rmass, "Rmass", &dRm));
    SOD_TRY(resolve_w(Rmom, "Rmom", &dRq));
    SOD_TRY(resolve_w(Rener, "Rener", &dRe));
    SOD_TRY(resolve_w(Reta, "Reta", &dRt));
    SOD_TRY(full_rhs(nelem, npoin, npoin_w, dc, dl, dHe, dgv, dMl, dr, dq, dE, dmf, dmue, dsgs, Rgas, gamma_gas, Cp,
                     Prt, dRm, dRq, dRe, dRt));
    return finish_call();
}

// ------------------------------------------------- whole-step driver (C ABI)
static int loop_create(sod_time_loop** out, int nelem, int npoin, int npoin_w, const int* connec,
                       const int* lpoin_w, const double* He, const double* gpvol, const double* Ml,
                       const double* helem, const double* helem_l, const double* mu_factor, double Rgas,
                       double gamma_gas, double Cp, double Prt, double* rho, double* u, double* q, double* pr,
                       double* E, double* Tem, double* csound, double* machno, double* e_int, double* eta,
                       double* mu_e, const double* mu_sgs, double* mu_fluid, double* mue_l, int ninst) {
    SOD_TRY(ensure_init());
    if (out == nullptr) return fail(SOD_ERR_ARG, "time_loop_create: out is NULL");
    *out = nullptr;
    if (nelem < 1 || npoin < 1 || npoin_w < 1 || npoin_w > npoin)
        return fail(SOD_ERR_ARG, "time_loop_create: inconsistent sizes");
    const int *dc, *dl;
    const double *dHe, *dgv, *dMl, *dh, *dhl, *dmf, *dsgs;
    double *drho, *du, *dq, *dpr, *dE, *dT, *dcs, *dma, *dei, *deta, *dmue, *dmuf, *dmuel;
    SOD_TRY(resolve_t(connec, "connec", &dc));
    SOD_TRY(resolve_t(lpoin_w, "lpoin_w", &dl));
    SOD_TRY(resolve_t(He, "He", &dHe));
    SOD_TRY(resolve_t(gpvol, "gpvol", &dgv));
    SOD_TRY(resolve_t(Ml, "Ml", &dMl));
    SOD_TRY(resolve_t(helem, "helem", &dh));
    SOD_TRY(resolve_t(helem_l, "helem_l", &dhl));
    SOD_TRY(resolve_t(mu_factor, "mu_factor", &dmf));
    SOD_TRY(resolve_t(mu_sgs, "mu_sgs", &dsgs));
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
    SOD_TRY(resolve_w(mu_fluid, "mu_fluid", &dmuf));
    SOD_TRY(resolve_w(mue_l, "mue_l", &dmuel));
    auto* L = new sod_time_loop();
    const size_t np = (size_t)npoin;
    L->ninst = ninst;
    L->rho = drho; L->u = du; L->q = dq; L->pr = dpr; L->E = dE; L->Tem = dT; L->e_int = dei; L->eta = deta;
    L->a = FusedArgs{nelem, npoin, npoin_w, dc, dl, dHe, dgv, dMl, dh, dhl, dmf, Rgas, gamma_gas, Cp, Prt,
                     drho + np, du + 3 * np, dq + 3 * np, dpr + np, dE + np, dT + np, dei + np, deta + np, dcs,
                     dma, dmuf, dmue, dmuel, dsgs};
    int rc = prepare(L->P, nelem, npoin, npoin_w, dc, dl, dHe, dgv, dMl);
    if (rc == SOD_OK) {
        hipError_t e = hipMalloc(&L->Ngp, sizeof(double) * 64 * 64);
        if (e == hipSuccess) e = hipMalloc(&L->sc, sizeof(double) * (3 * (size_t)ninst + 8));
        if (e == hipSuccess) {
            e = hipMalloc(&L->hmin, sizeof(double) * np);
            if (e == hipSuccess) {
                launch_hmin(nelem, npoin, dc, dh, L->hmin);
                e = hipStreamSynchronize(rt().stream);
            }
        }
        if (e == hipSuccess) e = hipMalloc(&L->ist, sizeof(int) * 4);
        if (e == hipSuccess) e = hipMalloc(&L->ek_partial, sizeof(double) * RED_BLOCKS);
        if (e == hipSuccess)
            e = hipMemcpy(L->Ngp, host_tables().Ngp, sizeof(double) * 64 * 64, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking);
        if (e != hipSuccess) rc = fail(SOD_ERR_HIP, "time_loop_create: %s", hipGetErrorString(e));
    }
    if (rc != SOD_OK) {
        sod_time_loop_destroy(L);
        return rc;
    }
    *out = L;
    return SOD_OK;
}

extern "C" int sod_time_loop_create(sod_time_loop** out, int nelem, int npoin, int npoin_w, const int* connec,
                                    const int* lpoin_w, const double* He, const double* gpvol, const double* Ml,
                                    const double* helem, const double* helem_l, const double* mu_factor,
                                    double Rgas, double gamma_gas, double Cp, double Prt, double* rho, double* u,
                                    double* q, double* pr, double* E, double* Tem, double* csound,
                                    double* machno, double* e_int, double* eta, double* mu_e,
                                    const double* mu_sgs, double* mu_fluid, double* mue_l) {
    SOD_TRY(loop_create(out, nelem, npoin, npoin_w, connec, lpoin_w, He, gpvol, Ml, helem, helem_l, mu_factor, Rgas,
                        gamma_gas, Cp, Prt, rho, u, q, pr, E, Tem, csound, machno, e_int, eta, mu_e, mu_sgs,
                        mu_fluid, mue_l, 1));
    return finish_call();
}

extern "C" int sod_time_loop_create_batched(sod_time_loop** out, int ninst, const int* elem_off,
                                            const int* node_off, const int* work_off, int nelem, int npoin,
                                            int npoin_w, const int* connec, const int* lpoin_w, const double* He,
                                            const double* gpvol, const double* Ml, const double* helem,
                                            const double* helem_l, const double* mu_factor, double Rgas,
                                            double gamma_gas, double Cp, double Prt, double* rho, double* u,
                                            double* q, double* pr, double* E, double* Tem, double* csound,
                                            double* machno, double* e_int, double* eta, double* mu_e,
                                            const double* mu_sgs, double* mu_fluid, double* mue_l) {
    if (ninst < 1 || !elem_off || !node_off || !work_off)
        return fail(SOD_ERR_ARG, "time_loop_create_batched: ninst >= 1 and the offset arrays are required");
    if (elem_off[0] != 0 || node_off[0] != 0 || work_off[0] != 0 || elem_off[ninst] != nelem ||
        node_off[ninst] != npoin || work_off[ninst] != npoin_w)
        return fail(SOD_ERR_ARG, "time_loop_create_batched: offsets must run from 0 to nelem/npoin/npoin_w");
    int max_w = 0;
    for (int i = 0; i < ninst; ++i) {
        if (elem_off[i + 1] <= elem_off[i] || node_off[i + 1] <= node_off[i] || work_off[i + 1] <= work_off[i])
            return fail(SOD_ERR_ARG, "time_loop_create_batched: instance %d is empty or the offsets decrease", i);
        max_w = std::max(max_w, work_off[i + 1] - work_off[i]);
    }
    SOD_TRY(loop_create(out, nelem, npoin, npoin_w, connec, lpoin_w, He, gpvol, Ml, helem, helem_l, mu_factor, Rgas,
                        gamma_gas, Cp, Prt, rho, u, q, pr, E, Tem, csound, machno, e_int, eta, mu_e, mu_sgs,
                        mu_fluid, mue_l, ninst));
    sod_time_loop* L = *out;
    std::vector<int> ids((size_t)nelem + (size_t)npoin);
    for (int i = 0; i < ninst; ++i) {
        for (int e = elem_off[i]; e < elem_off[i + 1]; ++e) ids[e] = i;
        for (int n = node_off[i]; n < node_off[i + 1]; ++n) ids[(size_t)nelem + n] = i;
    }
    hipError_t e = hipMalloc(&L->inst_ids, sizeof(int) * ids.size());
    if (e == hipSuccess) e = hipMalloc(&L->woff, sizeof(int) * (ninst + 1));
    if (e == hipSuccess) e = hipMalloc(&L->inst_stats, sizeof(double) * 3 * (size_t)ninst);
    if (e == hipSuccess) e = hipMemcpy(L->inst_ids, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(L->woff, work_off, sizeof(int) * (ninst + 1), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        sod_time_loop_destroy(L);
        *out = nullptr;
        return fail(SOD_ERR_HIP, "time_loop_create_batched: %s", hipGetErrorString(e));
    }
    L->B.ninst = ninst;
    L->B.elem_inst = L->inst_ids;
    L->B.node_inst = L->inst_ids + nelem;
    L->B.woff = L->woff;
    L->B.stats = L->inst_stats;
    L->B.norms = L->inst_stats + 2 * ninst;
    L->B.max_w = max_w;
    return finish_call();
}

static int loop_run(sod_time_loop* L, int istep0, int nsteps, double cfl_conv, double cfl_diff, int fixed_dt,
                    double* dt, double* time, int log_first, int log_every, double rho0, double* EK, int nEK,
                    int keep_levels) {
    SOD_TRY(ensure_init());
    if (L == nullptr || dt == nullptr || time == nullptr) return fail(SOD_ERR_ARG, "time_loop_run: NULL argument");
    if (nsteps < 0) return fail(SOD_ERR_ARG, "time_loop_run: nsteps < 0");
    const int ni = L->ninst;
    if (ni > 1 && log_every > 0) return fail(SOD_ERR_ARG, "time_loop_run: EK logging is per single mesh");
    int nlog = 0;
    if (log_every > 0)
        for (int i = istep0; i < istep0 + nsteps; ++i) nlog += (i >= log_first && (i - log_first) % log_every == 0);
    if (nlog > 0 && (EK == nullptr || nEK < nlog))
        return fail(SOD_ERR_ARG, "time_loop_run: EK must hold %d values", nlog);
    if (nsteps == 0) return SOD_OK;
    int has_sgs = 0;
    SOD_TRY(load_viscosities(L->P, L->a.nelem, L->a.mu_e, L->a.mu_sgs, &has_sgs));
    LoopCfg c;
    c.valid = true;
    c.has_sgs = has_sgs != 0;
    c.fixed_dt = fixed_dt != 0;
    c.keep_levels = keep_levels != 0;
    c.cfl_conv = cfl_conv;
    c.cfl_diff = cfl_diff;
    c.rho0 = rho0;
    c.log_first = log_first;
    c.log_every = log_every > 0 ? log_every : 0;
    c.params = rt().params;
    c.has_diff = rt().params.flag_real_diff == 1;
    c.ek_cap = std::max(L->cfg.ek_cap, std::max(1, nlog));
    if (!L->ek_hist || c.ek_cap > L->cfg.ek_cap) {
        loop_free_graph(L);
        (void)hipFree(L->ek_hist);
        L->ek_hist = nullptr;
        SOD_HIP(hipMalloc(&L->ek_hist, sizeof(double) * c.ek_cap));
    }
    std::vector<double> hsc(2 * (size_t)ni);
    for (int i = 0; i < ni; ++i) {
        hsc[i] = dt[i];
        hsc[ni + i] = time[i];
    }
    const int hist[2] = {istep0 - 1, 0};
    hipStream_t saved = rt().stream;
    SOD_HIP(hipMemcpyAsync(L->sc, hsc.data(), sizeof(double) * hsc.size(), hipMemcpyHostToDevice, saved));
    SOD_HIP(hipMemcpyAsync(L->ist, hist, sizeof(hist), hipMemcpyHostToDevice, saved));
    SOD_HIP(hipStreamSynchronize(saved));
    rt().stream = L->stream;
    const bool prof = prof_set(false);
    int rc = choose_stage_cfg(L->P, L->a, c.has_sgs, L->sc, ni > 1 ? &L->B : nullptr);
    c.stage_cfg = L->P.cfg;
    if (rc == SOD_OK && (!L->exec || !(c == L->cfg))) {
        loop_free_graph(L);
        hipError_t e = hipStreamBeginCapture(L->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            loop_enqueue_step(L, c);
            e = hipStreamEndCapture(L->stream, &L->graph);
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&L->exec, L->graph, nullptr, nullptr, 0);
        if (e == hipSuccess) {
            L->cfg = c;
        } else {
            rc = fail(SOD_ERR_HIP, "time_loop_run: graph capture: %s", hipGetErrorString(e));
            loop_free_graph(L);
        }
    }
    if (rc == SOD_OK) {
        for (int i = 0; i < nsteps && rc == SOD_OK; ++i) {
            const hipError_t e = hipGraphLaunch(L->exec, L->stream);
            if (e != hipSuccess) rc = fail(SOD_ERR_HIP, "time_loop_run: graph launch: %s", hipGetErrorString(e));
        }
        const hipError_t e = hipStreamSynchronize(L->stream);
        if (rc == SOD_OK && e != hipSuccess) rc = fail(SOD_ERR_HIP, "time_loop_run: %s", hipGetErrorString(e));
    }
    rt().stream = saved;
    prof_set(prof);
    SOD_TRY(rc);
    SOD_HIP(hipMemcpy(hsc.data(), L->sc, sizeof(double) * hsc.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < ni; ++i) {
        dt[i] = hsc[i];
        time[i] = hsc[ni + i];
    }
    if (nlog > 0) SOD_HIP(hipMemcpy(EK, L->ek_hist, sizeof(double) * nlog, hipMemcpyDeviceToHost));
    return SOD_OK;
}

extern "C" int sod_time_loop_run(sod_time_loop* L, int istep0, int nsteps, double cfl_conv, double cfl_diff,
                                 int fixed_dt, double* dt, double* time, int log_first, int log_every,
                                 double rho0, double* EK, int nEK, int keep_levels) {
    if (L && L->ninst != 1) return fail(SOD_ERR_ARG, "time_loop_run: use sod_time_loop_run_batched for an ensemble");
    return loop_run(L, istep0, nsteps, cfl_conv, cfl_diff, fixed_dt, dt, time, log_first, log_every, rho0, EK, nEK,
                    keep_levels);
}

extern "C" int sod_time_loop_run_batched(sod_time_loop* L, int nsteps, double cfl_conv, double cfl_diff,
                                         int fixed_dt, double* dt, double* time, int keep_levels) {
    return loop_run(L, 1, nsteps, cfl_conv, cfl_diff, fixed_dt, dt, time, 1, 0, 1.0, nullptr, 0, keep_levels);
}

extern "C" int sod_time_loop_destroy(sod_time_loop* L) {
    if (L == nullptr) return SOD_OK;
    if (L->stream) (void)hipStreamSynchronize(L->stream);
    loop_free_graph(L);
    if (L->stream) (void)hipStreamDestroy(L->stream);
    (void)hipFree(L->Ngp);
    (void)hipFree(L->sc);
    (void)hipFree(L->hmin);
    (void)hipFree(L->ist);
    (void)hipFree(L->ek_partial);
    (void)hipFree(L->ek_hist);
    (void)hipFree(L->inst_ids);
    (void)hipFree(L->woff);
    (void)hipFree(L->inst_stats);
    L->P.release();
    delete L;
    return SOD_OK;
}

// The prepared mesh a query entry point reports on: the time loop's own, or
// (loop == NULL) the one most recently stepped through the drop-in entry
// points; NULL when there is none.
static const Prepared* queried_mesh(const sod_time_loop* loop) {
    if (loop != nullptr) return &loop->P;
    const PrepSlot* last = nullptr;
    for (const auto& sl : g_slots)
        if (sl.P.nelem >= 0 && (last == nullptr || sl.used > last->used)) last = &sl;
    return last != nullptr ? &last->P : nullptr;
}

extern "C" int sod_set_autotune(int mode) {
    if (mode < 0 || mode > 2)
        return fail(SOD_ERR_ARG,
                    "sod_set_autotune: mode %d (0: default, 1: fixed-order configurations, 2: all, atomic "
                    "scatter included)",
                    mode);
    rt().autotune = mode;
    return SOD_OK;
}

extern "C" int sod_set_stage_config_lds(int elems_per_workgroup, int assembly, int lds_layout) {
    static const char* names[3] = {"element-node stores", "atomic scatter", "colour-ordered scatter"};
    for (int c = 0; c < kNumStageCfgs; ++c)
        if (kStageCfgs[c].nw == elems_per_workgroup && kStageCfgs[c].assembly == assembly &&
            kStageCfgs[c].lds == lds_layout) {
            rt().autotune = kPinned + c;
            return SOD_OK;
        }
    return fail(SOD_ERR_ARG,
                "sod_set_stage_config: no configuration with %d elements per workgroup, %s and LDS layout %d",
                elems_per_workgroup, assembly >= 0 && assembly <= 2 ? names[assembly] : "that assembly", lds_layout);
}

extern "C" int sod_set_stage_config(int elems_per_workgroup, int assembly) {
    return sod_set_stage_config_lds(elems_per_workgroup, assembly, LDS_SWIZZLED);
}

extern "C" int sod_stage_config(const sod_time_loop* loop, int* elems_per_workgroup, int* assembly,
                                int* tuned_mode, double* stage_ms) {
    const Prepared* P = queried_mesh(loop);
    if (P == nullptr) return fail(SOD_ERR_ARG, "sod_stage_config: no mesh has been stepped yet");
    const StageCfg& c = kStageCfgs[P->cfg];
    if (elems_per_workgroup) *elems_per_workgroup = c.nw;
    if (assembly) *assembly = c.assembly;
    if (tuned_mode) *tuned_mode = P->cfg_mode >= kPinned ? 3 : P->cfg_mode;
    if (stage_ms) *stage_ms = P->cfg_ms;
    return SOD_OK;
}

extern "C" int sod_stage_lds_layout(const sod_time_loop* loop, int* lds_layout) {
    const Prepared* P = queried_mesh(loop);
    if (P == nullptr) return fail(SOD_ERR_ARG, "sod_stage_lds_layout: no mesh has been stepped yet");
    if (lds_layout == nullptr) return fail(SOD_ERR_ARG, "sod_stage_lds_layout: lds_layout is NULL");
    *lds_layout = kStageCfgs[P->cfg].lds;
    return SOD_OK;
}

extern "C" int sod_stage_element_order(const sod_time_loop* loop, int* reordered) {
    const Prepared* P = queried_mesh(loop);
    if (P == nullptr) return fail(SOD_ERR_ARG, "sod_stage_element_order: no mesh has been stepped yet");
    if (reordered == nullptr) return fail(SOD_ERR_ARG, "sod_stage_element_order: reordered is NULL");
    *reordered = P->eorder != nullptr ? 1 : 0;
    return SOD_OK;
}

extern "C" int sod_element_colouring(const sod_time_loop* loop, int* ncolours, int* colour, int nelem) {
    const Prepared* P = queried_mesh(loop);
    if (P == nullptr || P->color_state != 1)
        return fail(SOD_ERR_ARG, "sod_element_colouring: no coloured mesh (step one with colour-ordered assembly)");
    if (nelem != P->nelem || colour == nullptr || ncolours == nullptr)
        return fail(SOD_ERR_ARG, "sod_element_colouring: nelem %d does not match the mesh (%d) or NULL output",
                    nelem, P->nelem);
    const int nc = (int)P->color_off.size() - 1;
    std::vector<int> elist((size_t)nelem);
    SOD_HIP(hipStreamSynchronize(rt().stream));
    if (nelem > 0) SOD_HIP(hipMemcpy(elist.data(), P->color_elems, sizeof(int) * nelem, hipMemcpyDeviceToHost));
    for (int k = 0; k < nc; ++k)
        for (int t = P->color_off[k]; t < P->color_off[k + 1]; ++t) colour[elist[t]] = k;
    *ncolours = nc;
    return SOD_OK;
}
