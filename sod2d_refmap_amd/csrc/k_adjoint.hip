// Hand-written adjoint (vector-Jacobian product) of the fused stage
// right-hand side of k_fused.hip: for a seed lambda on the assembled
// residuals (Rmass, Rmom, Rener, Reta) it returns lambda^T dR/d(rho, q, E)
// at the nodes and lambda^T dR/d(mu_e) at the element nodes, for the
// operators of elem_convec.f90:211-570 (split-form convection with E := h,
// entropy residual) and elem_diffu.f90:12-198 (viscous fluxes, test-node
// nu_e and kappa) at the EOS of time_integ.f90:266-316.
//
// One wavefront per element, lane = lattice node, as in the forward kernel.
// The kernel recomputes the forward element quantities, back-propagates the
// node-local formulas, and then needs three "transposed" line passes:
//   * the weak divergence of the viscous fluxes transposes into a strong
//     derivative of the seeds (lambda_G = D-derivative of lambda_dD);
//   * the advective derivatives u.grad(a) with this node's contravariant
//     velocity give a derivative of U^c itself (a forward-style line sum);
//   * every line sum over neighbours b of lane l transposes into a sum, at
//     lane b, over the lanes l of b's three lines of D[o_l][o_b] times l's
//     published seeds; lane b evaluates the product-rule factors with its
//     own primitive values.
// Seeds and geometry of each lane are published in LDS (50 rows per
// element), so a workgroup holds two elements.
#include <algorithm>

#include "stage_common.hpp"

namespace sod {

namespace {

constexpr int AWPB = 2;   // wavefronts (elements) per workgroup

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
constexpr int NROW = 50;  // LDS rows per element

// published rows of the transposed pass
enum {
    L_DIVF0 = 0, L_DIVF1, L_DIVF2, L_DIVU0, L_DIVU1, L_DIVU2, L_DIVFE, L_DIVFUE, L_DIVFK, L_DIVFUK, L_DIVFETA,
    L_DIVQ,                                                       // 12 seeds of the DU sums
    L_AH, L_ARH, L_AK, L_ARK, L_AQ0, L_AQ1, L_AQ2, L_AETA,        // 8 seeds of the A sums
    L_DX = 20,                                                    // 18: (p,u0,u1,u2,T,rho) x direction
    L_HM = 38,                                                    // 9: He(d,c) at row d + 3c
    L_UA = 47                                                     // 3: U^c of the lane
};

// SUTH: the Sutherland gate (P.suth) as a compile-time choice; as a run-time
// branch it made the compiler spill ~290 VGPRs to scratch (13x slower).
template <bool HAS_SGS, bool SUTH>
__global__ __launch_bounds__(64 * AWPB) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_stage_rhs_vjp(
    int nelem, int npoin, const int* __restrict__ conn, const double* __restrict__ He,
    const double* __restrict__ gpv, const double* __restrict__ mue, const double* __restrict__ musgs,
    const double* __restrict__ rho, const double* __restrict__ q0, const double* __restrict__ q1,
    const double* __restrict__ q2, const double* __restrict__ Ein, const double* __restrict__ mufac,
    const double* __restrict__ lam, double* __restrict__ relem, double* __restrict__ gmue_lane, StageParams P) {
    __shared__ double sh[AWPB][NROW][64];
    __shared__ double sD[16];
    if (threadIdx.x < 16) sD[threadIdx.x] = (&c_tab.D[0][0])[threadIdx.x];
    __syncthreads();
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int e = blockIdx.x * AWPB + w;
    if (e >= nelem) return;   // uniform per wavefront; only wave-level sync below
    const int i = l & 3, j = (l >> 2) & 3, k = l >> 4;
    const int own[3] = {i, j, k};
    const int me = swz(i, j, k);
    const size_t eb = (size_t)e * 64;
    const size_t np = (size_t)npoin;
    double(*S)[64] = sh[w];

    // ---- forward recompute: EOS ----------------------------------------
    const int n = conn[eb + l];
    const double r = rho[n];
    const double mm[3] = {q0[n], q1[n], q2[n]};
    const double en = Ein[n], mf = mufac[n];
    const double ri = 1.0 / r;
    const double u[3] = {mm[0] * ri, mm[1] * ri, mm[2] * ri};
    const double umag = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double eint = en * ri - 0.5 * umag;
    const double p = r * P.gm1 * eint;
    const double h = P.gog1 * p * ri;
    const double T = p * ri * P.iRgas;
    const double lphi = log(p * pow(r, -P.gamma));
    const double eta = (r * P.ig1) * lphi;
    const double kk = 0.5 * umag;
    const double sT = sqrt(T);
    // Sutherland's law, or the caller's mu_fluid (a constant of the state)
    // with the gate off
    const double muf = SUTH ? mf * 0.000001458 * (T * sT) / (T + 110.4) : mf;
    double hm[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) hm[d][c] = He[eb * 9 + (d + 3 * c) * 64 + l];
    const double gv = gpv[eb + l];
    const double mu_e = mue[eb + l];
    const double mu_s = HAS_SGS ? musgs[eb + l] : 0.0;

    // rows 0..8: primitives; rows 9..26: contravariant viscous fluxes
    S[P_RHO][me] = r;
    S[P_U0][me] = u[0];
    S[P_U1][me] = u[1];
    S[P_U2][me] = u[2];
    S[P_P][me] = p;
    S[P_H][me] = h;
    S[P_K][me] = kk;
    S[P_ETA][me] = eta;
    S[P_T][me] = T;
    wave_sync();

    // ---- forward recompute: line sums ----------------------------------
    double divUa[3] = {0, 0, 0};
    // (the sums that enter the residuals only linearly are not needed here)
    double divFue = 0, divFuk = 0, divQ = 0, Ah = 0, Ak = 0;
    double gP[3] = {0, 0, 0}, gT[3] = {0, 0, 0}, gR[3] = {0, 0, 0};
    double gU[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double Ua[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double h0 = hm[0][c], h1 = hm[1][c], h2 = hm[2][c];
        Ua[c] = h0 * u[0] + h1 * u[1] + h2 * u[2];
        double dp = 0, du0 = 0, du1 = 0, du2 = 0, dT = 0, dr = 0;
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
            const int sl = c == 0 ? swz(bb, j, k) : (c == 1 ? swz(i, bb, k) : swz(i, j, bb));
            const double Dc = sD[own[c] * 4 + bb];
            const double rb = S[P_RHO][sl], u0b = S[P_U0][sl], u1b = S[P_U1][sl], u2b = S[P_U2][sl];
            const double pb = S[P_P][sl], hb = S[P_H][sl], kb = S[P_K][sl], Tb = S[P_T][sl];
            dp += Dc * pb;
            du0 += Dc * u0b;
            du1 += Dc * u1b;
            du2 += Dc * u2b;
            dT += Dc * Tb;
            dr += Dc * rb;
            const double DU = Dc * (h0 * u0b + h1 * u1b + h2 * u2b);
            const double t = rb * DU;
            divUa[0] += u0b * DU;
            divUa[1] += u1b * DU;
            divUa[2] += u2b * DU;
            divFue += hb * DU;
            divFuk += kb * DU;
            divQ += t;
            const double DA = Dc * Ua[c];
            Ah += DA * hb;
            Ak += DA * kb;
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
    const double divU = gU[0][0] + gU[1][1] + gU[2][2];
    const double Arho = u[0] * gR[0] + u[1] * gR[1] + u[2] * gR[2];
    const double mu_fgp = muf + r * mu_s;
    const double mu_t = mu_fgp + mu_e;
    double tau[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) tau[a][b] = mu_t * (gU[a][b] + gU[b][a]);
#pragma unroll
    for (int a = 0; a < 3; ++a) tau[a][a] -= mu_fgp * (2.0 / 3.0) * divU;
    double tauU[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) tauU[a] = tau[a][0] * u[0] + tau[a][1] * u[1] + tau[a][2] * u[2];
    {
        const double* F[6] = {gR, tau[0], tau[1], tau[2], tauU, gT};
#pragma unroll
        for (int f = 0; f < 6; ++f)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                S[9 + 3 * f + c][me] = gv * (hm[0][c] * F[f][0] + hm[1][c] * F[f][1] + hm[2][c] * F[f][2]);
    }
    wave_sync();
    // weak divergences dD0 (grad rho) and dD5 (grad T) are the only ones the
    // adjoint needs (they multiply nu_e and kappa)
    double dD0 = 0, dD5 = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
            const int sl = c == 0 ? swz(bb, j, k) : (c == 1 ? swz(i, bb, k) : swz(i, j, bb));
            const double Dt = sD[bb * 4 + own[c]];
            dD0 += Dt * S[9 + c][sl];
            dD5 += Dt * S[9 + 15 + c][sl];
        }

    // ---- seeds and node-local back-propagation -------------------------
    const double lm = lam[n], lq[3] = {lam[np + n], lam[2 * np + n], lam[3 * np + n]}, le = lam[4 * np + n],
                 lh = lam[5 * np + n];
    const double nu_e = P.c_rho * mu_e * ri;
    const double kappa = muf * P.CpPr + P.c_ener * mu_e * P.CpPr + P.Cp09 * r * mu_s;
    double Lp_r = 0, Lp_u[3] = {0, 0, 0}, Lp_h = 0, Lp_k = 0, Lp_eta = 0;   // own primitives
    double L_muf = 0, L_ri = 0, L_mue = 0;
    {
        const double L_nu = lm * dD0, L_kap = le * dD5;
        L_mue += P.c_rho * ri * L_nu + P.c_ener * P.CpPr * L_kap;
        L_ri += P.c_rho * mu_e * L_nu;
        L_muf += P.CpPr * L_kap;
        Lp_r += P.Cp09 * mu_s * L_kap;
    }
    // convective residual
    const double LRm = gv * lm, LRe = gv * le, LReta = gv * lh;
    double L_divU = 0.5 * r * LRm, L_Arho = 0.5 * LRm, L_Au[3];
    double s_divQ = 0.5 * LRm;
    Lp_r += 0.5 * divU * LRm;
    const double we = 0.25 * LRe;
    const double s_divFe = we, s_divFue = we * r, s_divFk = we, s_divFuk = we * r;
    const double s_Ah = we * r, s_Arh = we, s_Ak = we * r, s_Ark = we;
    Lp_h += we * (divQ + r * divU + Arho);
    Lp_k += we * (divQ + r * divU + Arho);
    s_divQ += we * (h + kk);
    Lp_r += we * (h * divU + divFue + kk * divU + divFuk + Ah + Ak);
    L_divU += we * (r * h + r * kk);
    L_Arho += we * (h + kk);
    double s_divF[3], s_divUa[3], s_Aq[3], L_gP[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double wa = 0.25 * gv * lq[a];
        s_divF[a] = wa;
        Lp_u[a] += wa * (divQ + r * divU + Arho);
        s_divQ += wa * u[a];
        const double Au = u[0] * gU[a][0] + u[1] * gU[a][1] + u[2] * gU[a][2];
        Lp_r += wa * (u[a] * divU + divUa[a] + Au);
        L_divU += wa * r * u[a];
        s_divUa[a] = wa * r;
        L_Arho += wa * u[a];
        L_Au[a] = wa * r;
        s_Aq[a] = wa;
        L_gP[a] = gv * lq[a];
    }
    const double wt = 0.5 * LReta;
    const double s_divFeta = wt, s_Aeta = wt;
    Lp_eta += wt * divU;
    L_divU += wt * eta;
    double L_gU[3][3], L_gR[3], L_gT[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            Lp_u[d] += L_Au[a] * gU[a][d];
            L_gU[a][d] = L_Au[a] * u[d];
        }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        Lp_u[d] += L_Arho * gR[d];
        L_gR[d] = L_Arho * u[d];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) L_gU[a][a] += L_divU;

    // ---- transpose of the weak divergence: lambda_G = D-derivative of
    // the seeds lambda_dD at this lane -----------------------------------
    wave_sync();   // all lanes are done reading the viscous fluxes
    const double L_dD[6] = {nu_e * lm, lq[0], lq[1], lq[2], le, kappa * le};
#pragma unroll
    for (int f = 0; f < 6; ++f) S[9 + f][me] = L_dD[f];
    wave_sync();
    {
        double LG[6][3];
#pragma unroll
        for (int f = 0; f < 6; ++f)
#pragma unroll
            for (int c = 0; c < 3; ++c) LG[f][c] = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int bb = 0; bb < 4; ++bb) {
                const int sl = c == 0 ? swz(bb, j, k) : (c == 1 ? swz(i, bb, k) : swz(i, j, bb));
                const double Dc = sD[own[c] * 4 + bb];
#pragma unroll
                for (int f = 0; f < 6; ++f) LG[f][c] += Dc * S[9 + f][sl];
            }
        double LF[6][3];
#pragma unroll
        for (int f = 0; f < 6; ++f)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                LF[f][d] = gv * (hm[d][0] * LG[f][0] + hm[d][1] * LG[f][1] + hm[d][2] * LG[f][2]);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            L_gR[d] += LF[0][d];
            L_gT[d] += LF[5][d];
        }
        double Ltau[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) Ltau[a][b] = LF[1 + a][b] + LF[4][a] * u[b];
#pragma unroll
        for (int b = 0; b < 3; ++b) Lp_u[b] += LF[4][0] * tau[0][b] + LF[4][1] * tau[1][b] + LF[4][2] * tau[2][b];
        double L_mut = 0.0;
        const double trL = Ltau[0][0] + Ltau[1][1] + Ltau[2][2];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                L_mut += Ltau[a][b] * (gU[a][b] + gU[b][a]);
                L_gU[a][b] += mu_t * (Ltau[a][b] + Ltau[b][a]);
            }
        const double L_mufgp = -(2.0 / 3.0) * divU * trL + L_mut;
        const double L_dv = -(2.0 / 3.0) * mu_fgp * trL;
#pragma unroll
        for (int a = 0; a < 3; ++a) L_gU[a][a] += L_dv;
        L_mue += L_mut;
        L_muf += L_mufgp;
        Lp_r += mu_s * L_mufgp;
    }

    // ---- advective derivatives: the own contravariant velocity ----------
    // (reads the primitives of the line neighbours; rows 0..8 are intact)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double acc = 0.0;
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
            const int sl = c == 0 ? swz(bb, j, k) : (c == 1 ? swz(i, bb, k) : swz(i, j, bb));
            const double Dc = sD[own[c] * 4 + bb];
            const double rb = S[P_RHO][sl], hb = S[P_H][sl], kb = S[P_K][sl], etab = S[P_ETA][sl];
            const double psi = hb * s_Ah + kb * s_Ak + etab * s_Aeta +
                               rb * (hb * s_Arh + kb * s_Ark + S[P_U0][sl] * s_Aq[0] + S[P_U1][sl] * s_Aq[1] +
                                     S[P_U2][sl] * s_Aq[2]);
            acc += Dc * psi;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) Lp_u[d] += hm[d][c] * acc;
    }

    // ---- publish this lane's seeds of the line sums --------------------
    wave_sync();   // all lanes are done reading the primitives
    S[L_DIVF0][me] = s_divF[0];
    S[L_DIVF1][me] = s_divF[1];
    S[L_DIVF2][me] = s_divF[2];
    S[L_DIVU0][me] = s_divUa[0];
    S[L_DIVU1][me] = s_divUa[1];
    S[L_DIVU2][me] = s_divUa[2];
    S[L_DIVFE][me] = s_divFe;
    S[L_DIVFUE][me] = s_divFue;
    S[L_DIVFK][me] = s_divFk;
    S[L_DIVFUK][me] = s_divFuk;
    S[L_DIVFETA][me] = s_divFeta;
    S[L_DIVQ][me] = s_divQ;
    S[L_AH][me] = s_Ah;
    S[L_ARH][me] = s_Arh;
    S[L_AK][me] = s_Ak;
    S[L_ARK][me] = s_Ark;
    S[L_AQ0][me] = s_Aq[0];
    S[L_AQ1][me] = s_Aq[1];
    S[L_AQ2][me] = s_Aq[2];
    S[L_AETA][me] = s_Aeta;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double sP = 0, sTT = 0, sR = 0, sU0 = 0, sU1 = 0, sU2 = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double hd = hm[d][c];
            sP += hd * L_gP[d];
            sTT += hd * L_gT[d];
            sR += hd * L_gR[d];
            sU0 += hd * L_gU[0][d];
            sU1 += hd * L_gU[1][d];
            sU2 += hd * L_gU[2][d];
        }
        S[L_DX + 0 * 3 + c][me] = sP;
        S[L_DX + 1 * 3 + c][me] = sU0;
        S[L_DX + 2 * 3 + c][me] = sU1;
        S[L_DX + 3 * 3 + c][me] = sU2;
        S[L_DX + 4 * 3 + c][me] = sTT;
        S[L_DX + 5 * 3 + c][me] = sR;
#pragma unroll
        for (int d = 0; d < 3; ++d) S[L_HM + d + 3 * c][me] = hm[d][c];
        S[L_UA + c][me] = Ua[c];
    }
    wave_sync();

    // ---- transposed line pass: contributions of the lanes on this lane's
    // three lines (itself included) to this lane's primitives -------------
    double Lp_p = 0, Lp_T = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int ll = 0; ll < 4; ++ll) {
            const int sl = c == 0 ? swz(ll, j, k) : (c == 1 ? swz(i, ll, k) : swz(i, j, ll));
            const double Dc = sD[ll * 4 + own[c]];   // D[o_l][o_b]
            Lp_p += Dc * S[L_DX + 0 + c][sl];
            Lp_u[0] += Dc * S[L_DX + 3 + c][sl];
            Lp_u[1] += Dc * S[L_DX + 6 + c][sl];
            Lp_u[2] += Dc * S[L_DX + 9 + c][sl];
            Lp_T += Dc * S[L_DX + 12 + c][sl];
            Lp_r += Dc * S[L_DX + 15 + c][sl];
            const double hc0 = S[L_HM + 0 + 3 * c][sl], hc1 = S[L_HM + 1 + 3 * c][sl], hc2 = S[L_HM + 2 + 3 * c][sl];
            const double DU = Dc * (hc0 * u[0] + hc1 * u[1] + hc2 * u[2]);
            const double lF0 = S[L_DIVF0][sl], lF1 = S[L_DIVF1][sl], lF2 = S[L_DIVF2][sl];
            const double lU0 = S[L_DIVU0][sl], lU1 = S[L_DIVU1][sl], lU2 = S[L_DIVU2][sl];
            const double lFe = S[L_DIVFE][sl], lFue = S[L_DIVFUE][sl], lFk = S[L_DIVFK][sl],
                         lFuk = S[L_DIVFUK][sl], lFeta = S[L_DIVFETA][sl], lQ = S[L_DIVQ][sl];
            const double X1 = lF0 * u[0] + lF1 * u[1] + lF2 * u[2] + lFe * h + lFk * kk + lQ;
            const double X2 = lU0 * u[0] + lU1 * u[1] + lU2 * u[2] + lFue * h + lFuk * kk + lFeta * eta;
            const double Phi = r * X1 + X2;
            const double DP = Dc * Phi;
            Lp_u[0] += DP * hc0 + DU * (r * lF0 + lU0);
            Lp_u[1] += DP * hc1 + DU * (r * lF1 + lU1);
            Lp_u[2] += DP * hc2 + DU * (r * lF2 + lU2);
            Lp_r += DU * X1;
            Lp_h += DU * (r * lFe + lFue);
            Lp_k += DU * (r * lFk + lFuk);
            Lp_eta += DU * lFeta;
            const double DA = Dc * S[L_UA + c][sl];
            const double lAh = S[L_AH][sl], lArh = S[L_ARH][sl], lAk = S[L_AK][sl], lArk = S[L_ARK][sl];
            const double lAq0 = S[L_AQ0][sl], lAq1 = S[L_AQ1][sl], lAq2 = S[L_AQ2][sl], lAeta = S[L_AETA][sl];
            Lp_h += DA * (lAh + r * lArh);
            Lp_k += DA * (lAk + r * lArk);
            Lp_eta += DA * lAeta;
            Lp_r += DA * (h * lArh + kk * lArk + u[0] * lAq0 + u[1] * lAq1 + u[2] * lAq2);
            const double DAr = DA * r;
            Lp_u[0] += DAr * lAq0;
            Lp_u[1] += DAr * lAq1;
            Lp_u[2] += DAr * lAq2;
        }
    }

    // ---- EOS back-propagation and scatter -------------------------------
    {
        double LT = Lp_T, Lr = Lp_r, Lp = Lp_p, Lri = L_ri, Lk = Lp_k;
        const double den = T + 110.4;
        if (SUTH) LT += L_muf * mf * 0.000001458 * (1.5 * sT * den - T * sT) / (den * den);
        Lr += Lp_eta * P.ig1 * (lphi - P.gamma);
        Lp += Lp_eta * r * P.ig1 / p;
        Lp += LT * ri * P.iRgas + Lp_h * P.gog1 * ri;
        Lri += LT * p * P.iRgas + Lp_h * P.gog1 * p;
        Lr += Lp * P.gm1 * eint;
        const double Leint = Lp * r * P.gm1;
        const double Len = Leint * ri;
        Lri += Leint * en;
        Lk -= Leint;
        double Lm[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double Lu = Lp_u[a] + Lk * u[a];
            Lm[a] = Lu * ri;
            Lri += Lu * mm[a];
        }
        Lr += -ri * ri * Lri;
        // gradients at this element node; k_assemble_grad sums them per node
        double2* o = reinterpret_cast<double2*>(relem + (eb + l) * 6);
        o[0] = make_double2(Lr, Lm[0]);
        o[1] = make_double2(Lm[1], Lm[2]);
        o[2] = make_double2(Len, 0.0);
        gmue_lane[eb + l] = L_mue;
    }
}

// Per-node sums of the element-node gradients of k_stage_rhs_vjp (the
// positions of node n, ascending) into the caller's arrays.
__global__ __launch_bounds__(256) void k_assemble_grad(int npoin, const int* __restrict__ inv_off,
                                                       const int* __restrict__ inv_pos,
                                                       const double* __restrict__ relem, double* __restrict__ g_rho,
                                                       double* __restrict__ g_q, double* __restrict__ g_E) {
    const size_t np = (size_t)npoin;
    for (int n = blockIdx.x * 256 + threadIdx.x; n < npoin; n += gridDim.x * 256) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;
        const int q1 = inv_off[n + 1];
        for (int q = inv_off[n]; q < q1; ++q) {
            const double2* c = reinterpret_cast<const double2*>(relem + (size_t)inv_pos[q] * 6);
            const double2 x = c[0], y = c[1], z = c[2];
            a0 += x.x;
            a1 += x.y;
            a2 += y.x;
            a3 += y.y;
            a4 += z.x;
        }
        g_rho[n] = a0;
        g_q[n] = a1;
        g_q[n + np] = a2;
        g_q[n + 2 * np] = a3;
        g_E[n] = a4;
    }
}

// lane layout [e][lane] -> reference layout (nelem,64), 16 elements per block
__global__ __launch_bounds__(256) void k_lane_to_ref_tiled(int nelem, const double* __restrict__ src,
                                                           double* __restrict__ dst) {
    __shared__ double sA[16][65];
    __shared__ int s_l2n[64];
    if (threadIdx.x < 64) s_l2n[threadIdx.x] = c_tab.lane2node[threadIdx.x];
    const int e0 = blockIdx.x * 16;
    const size_t ne = (size_t)nelem;
    __syncthreads();
    for (int idx = threadIdx.x; idx < 1024; idx += 256) {
        const int el = idx >> 6, ln = idx & 63, e = e0 + el;
        if (e < nelem) sA[el][s_l2n[ln]] = src[(size_t)e * 64 + ln];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < 1024; idx += 256) {
        const int a = idx >> 4, el = idx & 15, e = e0 + el;
        if (e < nelem) dst[(size_t)e + ne * a] = sA[el][a];
    }
}

__global__ __launch_bounds__(256) void k_zero_doubles(long n, double* __restrict__ p) {
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) p[t] = 0.0;
}

// ---------------------------------------------------- entropy viscosity
// Adjoint of smart_visc_spectral (mod_entropy_viscosity.f90:68-122 and the
// filter of mod_filters.f90:15-55): for a seed lambda on mu_e(nelem,64) it
// accumulates lambda . dmu_e/d(rho, u, csound, Tem, Reta) at the nodes and
// lambda . dmu_e/d(norm) into g_norm.  The limiters are differentiated as
// subgradients: min(Ve, beta_e) follows Ve where Ve < beta_e, else beta_e;
// every max over the element's nodes follows its first (lowest-lane)
// maximiser; the max(.., ce) clamps pass the gradient only when the clamped
// argument is strictly larger.  The element's Jacobian extent (minJe/maxJe)
// is geometry and carries no gradient.
__device__ __forceinline__ int first_lane(unsigned long long m) { return __ffsll((long long)m) - 1; }

__global__ __launch_bounds__(256) void k_smart_visc_vjp(
    int nelem, int npoin, const int* __restrict__ connec, const double* __restrict__ Reta,
    const double* __restrict__ gpvol, double gamma_gas, const double* __restrict__ rho,
    const double* __restrict__ u, const double* __restrict__ csound, const double* __restrict__ Tem,
    const double* __restrict__ helem_l, const double* __restrict__ helem, const double* __restrict__ d_stats,
    int normalise, double Rgas, double ce, double cmax, double cglob, const double* __restrict__ lam_mu_e,
    double* __restrict__ g, double* __restrict__ g_norm) {
    __shared__ double s_lam[4][64];   // seed of mue_l, by gmsh node
    const int le = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int e_raw = blockIdx.x * 4 + le;
    const bool active = e_raw < nelem;
    const int e = active ? e_raw : nelem - 1;   // inactive waves mirror a valid element, write nothing
    const size_t np = (size_t)npoin, ne = (size_t)nelem;
    const double norm = normalise ? d_stats[1] : 1.0;
    const int a = l;
    const int n = connec[(size_t)e + ne * a] - 1;
    const double J = gpvol[a + 64 * (size_t)e] / c_tab.wgp[a];
    const double uu[3] = {u[n], u[n + np], u[n + 2 * np]};
    const double vmag = sqrt(uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2]);
    const double cs = csound[n], rr = rho[n], TT = Tem[n], Rt = Reta[n];
    const double minJe = wave_min(fmin(1000000.0, J));
    const double maxJe = wave_max(fmax(0.0, J));
    const double maxV = wave_max(vmag);
    const double maxC = wave_max(cs);
    const double M = maxV / maxC;
    double M15 = M * M;
    M15 = M15 * M15 * M;
    M15 = M15 * M15 * M15;
    const double v_pi = 3.14159265358979323846;
    const double th = tanh(M15 * v_pi);
    const double ceM = fmax(th, ce);
    const double rj = minJe / maxJe;
    const double ced_geo = fmax(1.0 - rj * rj, ce);
    const double ced = fmax(ced_geo, ceM);
    const double aux3 = sqrt(Rgas * gamma_gas * TT);
    const double he = helem[e];
    const double beta_l = (rr * he) * (cmax / 3.0) * (vmag + aux3);
    const double betae = wave_max(beta_l);
    const double R1 = rr * fabs(Rt) / norm;
    const double hl = helem_l[(size_t)e + ne * a];
    const double Ve = ced * R1 * (hl * hl);

    // transpose of the 27-point mirror filter (same position/weight table as
    // the forward): lane = geometric position, scatter into its inputs
    s_lam[le][l] = 0.0;
    __syncthreads();
    {
        const int conv[6] = {3, 1, 3, 4, 2, 3};
        const int gi = (l & 3) + 1, gj = ((l >> 2) & 3) + 1, gk = (l >> 4) + 1;
        const double w3[3] = {0.25, 0.5, 0.25};
        const int so = c_tab.lane2node[(conv[gi] - 1) + 4 * (conv[gj] - 1) + 16 * (conv[gk] - 1)];
        const double lm = active ? lam_mu_e[(size_t)e + ne * so] : 0.0;
#pragma unroll
        for (int ll = -1; ll <= 1; ++ll)
#pragma unroll
            for (int mm = -1; mm <= 1; ++mm)
#pragma unroll
                for (int nn = -1; nn <= 1; ++nn) {
                    const int si = conv[gi + ll] - 1, sj = conv[gj + mm] - 1, sk = conv[gk + nn] - 1;
                    const int b = c_tab.lane2node[si + 4 * sj + 16 * sk];
                    atomicAdd(&s_lam[le][b], w3[ll + 1] * w3[mm + 1] * w3[nn + 1] * lm);
                }
    }
    __syncthreads();
    if (!active) return;
    const double lml = cglob * s_lam[le][a];
    const bool via_ve = Ve < betae;
    const double L_Ve = via_ve ? lml : 0.0;
    const double L_betae = wave_sum_d(via_ve ? 0.0 : lml);
    const double L_ced = wave_sum_d(L_Ve * R1 * (hl * hl));
    const double L_R1 = L_Ve * ced * (hl * hl);
    double Lr = L_R1 * fabs(Rt) / norm;
    const double sgn = Rt > 0.0 ? 1.0 : (Rt < 0.0 ? -1.0 : 0.0);
    double LRt = L_R1 * rr * sgn / norm;
    const double Lnorm = normalise ? wave_sum_d(-L_R1 * R1 / norm) : 0.0;
    double Lvmag = 0.0, Lcs = 0.0, LT = 0.0;
    // ced = max(ced_geo, ceM), ceM = max(tanh(pi M^15), ce), M = maxV / maxC
    if (ceM > ced_geo && th > ce && M > 0.0) {
        double M7 = M * M * M;
        M7 = M7 * M7 * M;
        const double L_M = L_ced * v_pi * (1.0 - th * th) * 15.0 * (M7 * M7);
        const int iv = first_lane(__ballot(vmag == maxV));
        const int ic = first_lane(__ballot(cs == maxC));
        if (l == iv) Lvmag += L_M / maxC;
        if (l == ic) Lcs += -L_M * maxV / (maxC * maxC);
    }
    // beta_e = max over nodes of rho h_e cmax/3 (|u| + sqrt(R gamma T))
    {
        const int ib = first_lane(__ballot(beta_l == betae));
        if (l == ib) {
            const double k3 = he * (cmax / 3.0);
            Lr += L_betae * k3 * (vmag + aux3);
            Lvmag += L_betae * rr * k3;
            const double Laux3 = L_betae * rr * k3;
            if (aux3 > 0.0) LT += Laux3 * Rgas * gamma_gas / (2.0 * aux3);
        }
    }
    double Lu[3] = {0.0, 0.0, 0.0};
    if (vmag > 0.0) {
#pragma unroll
        for (int d = 0; d < 3; ++d) Lu[d] = Lvmag * uu[d] / vmag;
    }
    unsafeAtomicAdd(g + n, Lr);
    unsafeAtomicAdd(g + np + n, Lu[0]);
    unsafeAtomicAdd(g + 2 * np + n, Lu[1]);
    unsafeAtomicAdd(g + 3 * np + n, Lu[2]);
    unsafeAtomicAdd(g + 4 * np + n, Lcs);
    unsafeAtomicAdd(g + 5 * np + n, LT);
    unsafeAtomicAdd(g + 6 * np + n, LRt);
    if (l == 0 && normalise) unsafeAtomicAdd(g_norm, Lnorm);
}

// norm = max over working nodes of |eta - mean|: the first working node
// (in lpoin_w order) attaining it
__global__ __launch_bounds__(256) void k_norm_argmax(int npoin_w, const int* __restrict__ lpoin_w,
                                                     const double* __restrict__ eta,
                                                     const double* __restrict__ d_stats, int* __restrict__ idx) {
    const double mean = d_stats[0] / (double)npoin_w;
    for (int t = blockIdx.x * 256 + threadIdx.x; t < npoin_w; t += gridDim.x * 256)
        if (fabs(eta[lpoin_w[t] - 1] - mean) == d_stats[1]) atomicMin(idx, t);
}

// lambda_eta = g_norm * s * (delta(n, n*) - 1/npoin_w) over working nodes
__global__ __launch_bounds__(256) void k_norm_to_eta(int npoin_w, const int* __restrict__ lpoin_w,
                                                     const double* __restrict__ eta,
                                                     const double* __restrict__ d_stats,
                                                     const int* __restrict__ idx,
                                                     const double* __restrict__ g_norm, double* __restrict__ g_eta) {
    const int istar = *idx;
    if (istar < 0 || istar >= npoin_w) return;
    const double mean = d_stats[0] / (double)npoin_w;
    const double dv = eta[lpoin_w[istar] - 1] - mean;
    const double sgn = dv > 0.0 ? 1.0 : (dv < 0.0 ? -1.0 : 0.0);
    const double Ln = g_norm[0] * sgn;
    for (int t = blockIdx.x * 256 + threadIdx.x; t < npoin_w; t += gridDim.x * 256) {
        const int n = lpoin_w[t] - 1;
        g_eta[n] = (t == istar ? Ln : 0.0) - Ln / (double)npoin_w;
    }
}

__global__ void k_init_vjp_scalars(double* g_norm, int* idx, int big) {
    if (threadIdx.x == 0) {
        *g_norm = 0.0;
        *idx = big;
    }
}

}  // namespace

void launch_eta_stats(int npoin_w, const int* lpoin_w, const double* eta, double* work);

}  // namespace sod

using namespace sod;

extern "C" int sod_smart_visc_spectral_vjp(int nelem, int npoin, int npoin_w, const int* connec,
                                           const int* lpoin_w, const double* Reta, const double* gpvol,
                                           double gamma_gas, const double* rho, const double* u,
                                           const double* csound, const double* Tem, const double* eta,
                                           const double* helem_l, const double* helem, const double* lam_mu_e,
                                           double* g_rho, double* g_u, double* g_csound, double* g_Tem,
                                           double* g_eta, double* g_Reta) {
    SOD_TRY(ensure_init());
    if (nelem < 0 || npoin < 0 || npoin_w < 0 || npoin_w > npoin)
        return fail(SOD_ERR_ARG, "smart_visc_spectral_vjp: inconsistent sizes");
    const int *dc, *dl;
    const double *dRt, *dgv, *dr, *du, *dcs, *dT, *deta, *dhl, *dh, *dlam;
    double *ogr, *ogu, *ogc, *ogT, *oge, *ogR;
    SOD_TRY(resolve_t(connec, "connec", &dc));
    SOD_TRY(resolve_t(lpoin_w, "lpoin_w", &dl));
    SOD_TRY(resolve_t(Reta, "Reta", &dRt));
    SOD_TRY(resolve_t(gpvol, "gpvol", &dgv));
    SOD_TRY(resolve_t(rho, "rho", &dr));
    SOD_TRY(resolve_t(u, "u", &du));
    SOD_TRY(resolve_t(csound, "csound", &dcs));
    SOD_TRY(resolve_t(Tem, "Tem", &dT));
    SOD_TRY(resolve_t(eta, "eta", &deta));
    SOD_TRY(resolve_t(helem_l, "helem_l", &dhl));
    SOD_TRY(resolve_t(helem, "helem", &dh));
    SOD_TRY(resolve_t(lam_mu_e, "lam_mu_e", &dlam));
    SOD_TRY(resolve_w(g_rho, "g_rho", &ogr));
    SOD_TRY(resolve_w(g_u, "g_u", &ogu));
    SOD_TRY(resolve_w(g_csound, "g_csound", &ogc));
    SOD_TRY(resolve_w(g_Tem, "g_Tem", &ogT));
    SOD_TRY(resolve_w(g_eta, "g_eta", &oge));
    SOD_TRY(resolve_w(g_Reta, "g_Reta", &ogR));
    const sod_numerical_params& prm = rt().params;
    hipStream_t s = rt().stream;
    const size_t np = (size_t)npoin;
    constexpr int RB = 1024;
    void* buf = nullptr;
    SOD_TRY(scratch(sizeof(double) * (np * 7 + 2 * RB + 16), &buf, 7));
    double* g = static_cast<double*>(buf);
    double* stats = g + 7 * np;          // [0] sum, [1] norm, partials
    double* g_norm = stats + 2 * RB + 4;
    int* idx = reinterpret_cast<int*>(g_norm + 2);
    const int normalise = prm.flag_normalise_entropy == 1;
    hipLaunchKernelGGL(k_zero_doubles, dim3(std::max(1, std::min(1024, grid_for((long)np * 7, 256)))), dim3(256), 0, s,
                       (long)np * 7, g);
    hipLaunchKernelGGL(k_init_vjp_scalars, dim3(1), dim3(64), 0, s, g_norm, idx, npoin_w);
    if (normalise && npoin_w > 0) launch_eta_stats(npoin_w, dl, deta, stats);
    if (nelem > 0) {
        SOD_PROF("smart_visc_vjp");
        hipLaunchKernelGGL(k_smart_visc_vjp, dim3(grid_for(nelem, 4)), dim3(256), 0, s, nelem, npoin, dc, dRt, dgv,
                           gamma_gas, dr, du, dcs, dT, dhl, dh, stats, normalise, prm.Rgas_inf, prm.ce, prm.cmax,
                           prm.cglob, dlam, g, g_norm);
    }
    // eta enters only through the normalisation
    hipLaunchKernelGGL(k_zero_doubles, dim3(std::max(1, std::min(1024, grid_for((long)np, 256)))), dim3(256), 0, s,
                       (long)np, oge);
    if (normalise && npoin_w > 0) {
        const int gr = std::max(1, std::min(1024, grid_for(npoin_w, 256)));
        hipLaunchKernelGGL(k_norm_argmax, dim3(gr), dim3(256), 0, s, npoin_w, dl, deta, stats, idx);
        hipLaunchKernelGGL(k_norm_to_eta, dim3(gr), dim3(256), 0, s, npoin_w, dl, deta, stats, idx, g_norm, oge);
    }
    SOD_HIP(hipMemcpyAsync(ogr, g, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
    SOD_HIP(hipMemcpyAsync(ogu, g + np, sizeof(double) * np * 3, hipMemcpyDeviceToDevice, s));
    SOD_HIP(hipMemcpyAsync(ogc, g + 4 * np, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
    SOD_HIP(hipMemcpyAsync(ogT, g + 5 * np, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
    SOD_HIP(hipMemcpyAsync(ogR, g + 6 * np, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
    SOD_HIP(hipGetLastError());
    return finish_call();
}

extern "C" int sod_full_rhs_vjp(int nelem, int npoin, int npoin_w, const int* connec, const int* lpoin_w,
                                const double* He, const double* gpvol, const double* Ml, const double* rho,
                                const double* q, const double* E, const double* mu_factor, const double* mu_e,
                                const double* mu_sgs, double Rgas, double gamma_gas, double Cp, double Prt,
                                const double* lam_mass, const double* lam_mom, const double* lam_ener,
                                const double* lam_eta, double* g_rho, double* g_q, double* g_E, double* g_mu_e) {
    SOD_TRY(ensure_init());
    if (nelem < 0 || npoin < 0 || npoin_w < 0 || npoin_w > npoin)
        return fail(SOD_ERR_ARG, "full_rhs_vjp: inconsistent sizes");
    if (rt().params.flag_total_enthalpy != 0)
        return fail(SOD_ERR_UNSUPPORTED,
                    "full_rhs_vjp: the adjoint of full_convec_ijk_H (flag_total_enthalpy) is not implemented");
    const int *dc, *dl;
    const double *dHe, *dgv, *dMl, *dr, *dq, *dE, *dmf, *dmue, *dsgs, *dlm, *dlq, *dle, *dlt;
    double *dgr, *dgq, *dgE, *dgmu;
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
    SOD_TRY(resolve_t(lam_mass, "lam_mass", &dlm));
    SOD_TRY(resolve_t(lam_mom, "lam_mom", &dlq));
    SOD_TRY(resolve_t(lam_ener, "lam_ener", &dle));
    SOD_TRY(resolve_t(lam_eta, "lam_eta", &dlt));
    SOD_TRY(resolve_w(g_rho, "g_rho", &dgr));
    SOD_TRY(resolve_w(g_q, "g_q", &dgq));
    SOD_TRY(resolve_w(g_E, "g_E", &dgE));
    SOD_TRY(resolve_w(g_mu_e, "g_mu_e", &dgmu));
    LaneMesh lm;
    SOD_TRY(prepare_lane_mesh(nelem, npoin, npoin_w, dc, dl, dHe, dgv, dMl, dmue, dsgs, &lm));
    hipStream_t s = rt().stream;
    const size_t np = (size_t)npoin, ne = (size_t)nelem;
    // seeds and gradients as 6 / 5 contiguous fields of npoin
    void* buf = nullptr;
    SOD_TRY(scratch(sizeof(double) * (np * 6 + ne * 64 + 8), &buf, 6));
    double* lam = static_cast<double*>(buf);
    double* gml = lam + 6 * np;
    SOD_HIP(hipMemcpyAsync(lam, dlm, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
    SOD_HIP(hipMemcpyAsync(lam + np, dlq, sizeof(double) * np * 3, hipMemcpyDeviceToDevice, s));
    SOD_HIP(hipMemcpyAsync(lam + 4 * np, dle, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
    SOD_HIP(hipMemcpyAsync(lam + 5 * np, dlt, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
    const sod_numerical_params& prm = rt().params;
    const StageParams sp = make_stage_params(gamma_gas, Rgas, Cp, Prt, prm);
    SOD_PROF("full_rhs_vjp");   // the element kernel, the mu_e transpose and the node sums
    if (nelem > 0) {
        auto* kern = lm.has_sgs ? (sp.suth ? k_stage_rhs_vjp<true, true> : k_stage_rhs_vjp<true, false>)
                                : (sp.suth ? k_stage_rhs_vjp<false, true> : k_stage_rhs_vjp<false, false>);
        hipLaunchKernelGGL(kern, dim3(grid_for(nelem, AWPB)), dim3(64 * AWPB), 0, s, nelem, npoin, lm.conn, lm.He,
                           lm.gpv, lm.mue, lm.musgs, dr, dq, dq + np, dq + 2 * np, dE, dmf, lam, lm.relem, gml, sp);
        hipLaunchKernelGGL(k_lane_to_ref_tiled, dim3(grid_for(nelem, 16)), dim3(256), 0, s, nelem, gml, dgmu);
    }
    if (npoin > 0)
        hipLaunchKernelGGL(k_assemble_grad, dim3(std::max(1, grid_for(npoin, 256))), dim3(256), 0, s, npoin,
                           lm.inv_off, lm.inv_pos, lm.relem, dgr, dgq, dgE);
    SOD_HIP(hipGetLastError());
    return finish_call();
}
