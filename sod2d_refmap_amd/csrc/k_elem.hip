// Element operators with the reference's argument lists and array layouts:
// full_convec_ijk(_H), generic_scalar_convec_ijk and full_diffusion_ijk.
//
// Launch shape: a workgroup of 64*EPB threads takes EPB elements; the 64
// lanes of a wavefront are the 64 nodes of one element in lattice order
// (lane = i + 4j + 16k), so the three 4-point lines through a node are the
// lanes {ii+4j+16k}, {i+4ii+16k}, {i+4j+16ii}.  Nodal fields are gathered
// through connec into LDS once per element; the 4-point derivative sums read
// LDS; the element residual is scattered with FP64 atomics (the reference's
// "!$acc atomic update", elem_convec.f90:439-455).
#include "sod2d_internal.hpp"

namespace sod {

namespace {

constexpr int EPB = 4;

struct Lines {
    int x[4], y[4], z[4];
    double dx[4], dy[4], dz[4];
};

__device__ __forceinline__ void lines_of(int l, Lines& L) {
    const int i = l & 3, j = (l >> 2) & 3, k = l >> 4;
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
        L.x[ii] = ii + 4 * j + 16 * k;
        L.y[ii] = i + 4 * ii + 16 * k;
        L.z[ii] = i + 4 * j + 16 * ii;
        L.dx[ii] = c_tab.D[i][ii];
        L.dy[ii] = c_tab.D[j][ii];
        L.dz[ii] = c_tab.D[k][ii];
    }
}

// Derivatives along xi, eta, zeta of a lane-indexed field f(m).
template <class F>
__device__ __forceinline__ void giso(const Lines& L, F f, double g[3]) {
    g[0] = 0.0;
    g[1] = 0.0;
    g[2] = 0.0;
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
        g[0] += L.dx[ii] * f(L.x[ii]);
        g[1] += L.dy[ii] * f(L.y[ii]);
        g[2] += L.dz[ii] * f(L.z[ii]);
    }
}

// out_i = sum_j He(i,j) g_j
__device__ __forceinline__ void grad_phys(const double h[3][3], const double g[3], double out[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        s += h[i][0] * g[0];
        s += h[i][1] * g[1];
        s += h[i][2] * g[2];
        out[i] = s;
    }
}

__device__ __forceinline__ void load_he(const double* __restrict__ He, int e, int a, double h[3][3]) {
    const double* p = He + 9 * (size_t)a + 576 * (size_t)e;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) h[i][j] = p[i + 3 * j];
}

// ---------------------------------------------------------------- convection
// elem_convec.f90:211-460 (TOTAL_H = false: energy flux split with h and k,
// E argument carries h) and elem_convec.f90:20-208 (TOTAL_H = true).
template <bool TOTAL_H>
__global__ __launch_bounds__(64 * EPB) void k_full_convec(
    int nelem, int npoin, const int* __restrict__ connec, const double* __restrict__ He,
    const double* __restrict__ gpvol, const double* __restrict__ u, const double* __restrict__ q,
    const double* __restrict__ rho, const double* __restrict__ pr, const double* __restrict__ E,
    double* __restrict__ Rmass, double* __restrict__ Rmom, double* __restrict__ Rener) {
    __shared__ double s_rho[EPB][64], s_pr[EPB][64], s_E[EPB][64], s_k[EPB][64];
    __shared__ double s_u[EPB][3][64], s_q[EPB][3][64];
    const int le = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int e = blockIdx.x * EPB + le;
    const bool active = e < nelem;
    const size_t np = (size_t)npoin;
    const int a = c_tab.lane2node[l];
    int n = 0;
    if (active) {
        n = connec[(size_t)e + (size_t)nelem * a] - 1;
        double u0 = u[n], u1 = u[n + np], u2 = u[n + 2 * np];
        s_u[le][0][l] = u0;
        s_u[le][1][l] = u1;
        s_u[le][2][l] = u2;
        s_q[le][0][l] = q[n];
        s_q[le][1][l] = q[n + np];
        s_q[le][2][l] = q[n + 2 * np];
        s_rho[le][l] = rho[n];
        s_pr[le][l] = pr[n];
        s_E[le][l] = E[n];
        double kk = 0.0;   // elem_convec.f90:248-251
        kk = kk + u0 * u0 * 0.5;
        kk = kk + u1 * u1 * 0.5;
        kk = kk + u2 * u2 * 0.5;
        s_k[le][l] = kk;
    }
    __syncthreads();
    if (!active) return;
    const double* R = s_rho[le];
    const double* P = s_pr[le];
    const double* EE = s_E[le];
    const double* K = s_k[le];
    const double* U[3] = {s_u[le][0], s_u[le][1], s_u[le][2]};
    const double* Q[3] = {s_q[le][0], s_q[le][1], s_q[le][2]};
    Lines L;
    lines_of(l, L);
    double h[3][3];
    load_he(He, e, a, h);

    double g[3];
    double gradRho[3], gradP[3], gradE[3];
    giso(L, [&](int m) { return R[m]; }, g);
    grad_phys(h, g, gradRho);
    giso(L, [&](int m) { return P[m]; }, g);
    grad_phys(h, g, gradP);
    if (TOTAL_H)
        giso(L, [&](int m) { return P[m] + EE[m]; }, g);
    else
        giso(L, [&](int m) { return EE[m]; }, g);
    grad_phys(h, g, gradE);
    double gradRE[3] = {0, 0, 0}, gradk[3] = {0, 0, 0}, gradRk[3] = {0, 0, 0};
    if (!TOTAL_H) {
        giso(L, [&](int m) { return R[m] * EE[m]; }, g);
        grad_phys(h, g, gradRE);
        giso(L, [&](int m) { return K[m]; }, g);
        grad_phys(h, g, gradk);
        giso(L, [&](int m) { return R[m] * K[m]; }, g);
        grad_phys(h, g, gradRk);
    }
    // gradU(i,j) = sum_k He(j,k) dU_i/dxi_k ; same for q
    double gradU[3][3], gradQ[3][3];
#pragma unroll
    for (int id = 0; id < 3; ++id) {
        giso(L, [&](int m) { return U[id][m]; }, g);
#pragma unroll
        for (int jd = 0; jd < 3; ++jd) {
            double s = 0.0;
            s += h[jd][0] * g[0];
            s += h[jd][1] * g[1];
            s += h[jd][2] * g[2];
            gradU[id][jd] = s;
        }
        giso(L, [&](int m) { return Q[id][m]; }, g);
#pragma unroll
        for (int jd = 0; jd < 3; ++jd) {
            double s = 0.0;
            s += h[jd][0] * g[0];
            s += h[jd][1] * g[1];
            s += h[jd][2] * g[2];
            gradQ[id][jd] = s;
        }
    }
    // divergences of vector fluxes: sum_i sum_j He(i,j) d(f_i)/dxi_j
    double divFe = 0.0, divFue = 0.0, divFk = 0.0, divFuk = 0.0;
#pragma unroll
    for (int id = 0; id < 3; ++id) {
        if (TOTAL_H) {
            giso(L, [&](int m) { return (EE[m] + P[m]) * U[id][m]; }, g);
            divFe += h[id][0] * g[0];
            divFe += h[id][1] * g[1];
            divFe += h[id][2] * g[2];
        } else {
            giso(L, [&](int m) { return R[m] * EE[m] * U[id][m]; }, g);
            divFe += h[id][0] * g[0];
            divFe += h[id][1] * g[1];
            divFe += h[id][2] * g[2];
            giso(L, [&](int m) { return EE[m] * U[id][m]; }, g);
            divFue += h[id][0] * g[0];
            divFue += h[id][1] * g[1];
            divFue += h[id][2] * g[2];
            giso(L, [&](int m) { return R[m] * K[m] * U[id][m]; }, g);
            divFk += h[id][0] * g[0];
            divFk += h[id][1] * g[1];
            divFk += h[id][2] * g[2];
            giso(L, [&](int m) { return K[m] * U[id][m]; }, g);
            divFuk += h[id][0] * g[0];
            divFuk += h[id][1] * g[1];
            divFuk += h[id][2] * g[2];
        }
    }
    // divF(i) = sum_j sum_k He(j,k) d(q_i u_j)/dxi_k ; divFuu with u_i u_j
    double divF[3], divFuu[3];
#pragma unroll
    for (int id = 0; id < 3; ++id) {
        double sF = 0.0, sU = 0.0;
#pragma unroll
        for (int jd = 0; jd < 3; ++jd) {
            giso(L, [&](int m) { return Q[id][m] * U[jd][m]; }, g);
            sF += h[jd][0] * g[0];
            sF += h[jd][1] * g[1];
            sF += h[jd][2] * g[2];
            giso(L, [&](int m) { return U[id][m] * U[jd][m]; }, g);
            sU += h[jd][0] * g[0];
            sU += h[jd][1] * g[1];
            sU += h[jd][2] * g[2];
        }
        divF[id] = sF;
        divFuu[id] = sU;
    }
    const double divU = gradU[0][0] + gradU[1][1] + gradU[2][2];
    const double divQ = gradQ[0][0] + gradQ[1][1] + gradQ[2][2];
    const double rl = R[l], El = EE[l], kl = K[l], pl = P[l];
    const double ul[3] = {U[0][l], U[1][l], U[2][l]};
    const double ql[3] = {Q[0][l], Q[1][l], Q[2][l]};
    double Re_mass = 0.5 * (divQ + rl * divU);
    double Re_ener;
    if (TOTAL_H)
        Re_ener = 0.5 * (divFe + (El + pl) * divU);
    else
        Re_ener = 0.25 * (divFe + El * divQ + (rl * El) * divU + rl * divFue) +
                  0.25 * (divFk + kl * divQ + (rl * kl) * divU + rl * divFuk);
    double Re_mom[3];
#pragma unroll
    for (int id = 0; id < 3; ++id) {
        double m = 0.25 * (divF[id] + ul[id] * divQ + ql[id] * divU + rl * divFuu[id]) + gradP[id];
        Re_mass = Re_mass + 0.5 * (ul[id] * gradRho[id]);
        if (TOTAL_H)
            Re_ener = Re_ener + 0.5 * (ul[id] * gradE[id]);
        else
            Re_ener = Re_ener +
                      0.25 * (ql[id] * gradE[id] + ul[id] * gradRE[id] + (El * ul[id]) * gradRho[id]) +
                      0.25 * (ql[id] * gradk[id] + ul[id] * gradRk[id] + (kl * ul[id]) * gradRho[id]);
#pragma unroll
        for (int jd = 0; jd < 3; ++jd)
            m = m + 0.25 * ((ul[id] * ul[jd]) * gradRho[jd] + ql[jd] * gradU[id][jd] +
                            ul[jd] * gradQ[id][jd]);
        Re_mom[id] = m;
    }
    const double gv = gpvol[a + 64 * (size_t)e];
    unsafeAtomicAdd(&Rmom[n], gv * Re_mom[0]);
    unsafeAtomicAdd(&Rmom[n + np], gv * Re_mom[1]);
    unsafeAtomicAdd(&Rmom[n + 2 * np], gv * Re_mom[2]);
    unsafeAtomicAdd(&Rmass[n], gv * Re_mass);
    unsafeAtomicAdd(&Rener[n], gv * Re_ener);
}

// ---------------------------------------------------------- entropy convection
// generic_scalar_convec_ijk, elem_convec.f90:462-570
__global__ __launch_bounds__(64 * EPB) void k_scalar_convec(
    int nelem, int npoin, const int* __restrict__ connec, const double* __restrict__ He,
    const double* __restrict__ gpvol, const double* __restrict__ q, const double* __restrict__ eta,
    const double* __restrict__ u, double* __restrict__ Rconvec) {
    __shared__ double s_u[EPB][3][64], s_f[EPB][3][64], s_eta[EPB][64];
    const int le = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int e = blockIdx.x * EPB + le;
    const bool active = e < nelem;
    const size_t np = (size_t)npoin;
    const int a = c_tab.lane2node[l];
    int n = 0;
    if (active) {
        n = connec[(size_t)e + (size_t)nelem * a] - 1;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            s_u[le][d][l] = u[n + np * d];
            s_f[le][d][l] = q[n + np * d];
        }
        s_eta[le][l] = eta[n];
    }
    __syncthreads();
    if (!active) return;
    Lines L;
    lines_of(l, L);
    double h[3][3];
    load_he(He, e, a, h);
    const double* ET = s_eta[le];
    double g[3], gradE[3];
    giso(L, [&](int m) { return ET[m]; }, g);
    grad_phys(h, g, gradE);
    double divU = 0.0, divFe = 0.0;
#pragma unroll
    for (int id = 0; id < 3; ++id) {
        const double* UU = s_u[le][id];
        const double* FF = s_f[le][id];
        double gu[3], gf[3];
        giso(L, [&](int m) { return UU[m]; }, gu);
        giso(L, [&](int m) { return FF[m]; }, gf);
#pragma unroll
        for (int jd = 0; jd < 3; ++jd) {
            divU += h[id][jd] * gu[jd];
            divFe += h[id][jd] * gf[jd];
        }
    }
    double Re = 0.5 * (divFe + ET[l] * divU);
#pragma unroll
    for (int id = 0; id < 3; ++id) Re = Re + 0.5 * (s_u[le][id][l] * gradE[id]);
    unsafeAtomicAdd(&Rconvec[n], gpvol[a + 64 * (size_t)e] * Re);
}

// -------------------------------------------------------------- diffusion
// full_diffusion_ijk, elem_diffu.f90:12-198.  Phase 1 builds tau, tau.u,
// grad T and grad rho at every node; phase 2 forms the weak divergence at
// each test node with kappa and nu_e evaluated at the test node.
__global__ __launch_bounds__(64 * EPB) void k_full_diffusion(
    int nelem, int npoin, const int* __restrict__ connec, const double* __restrict__ He,
    const double* __restrict__ gpvol, double Cp, double Pr, double c_rho, double c_ener,
    const double* __restrict__ rho_n, const double* __restrict__ rho, const double* __restrict__ u,
    const double* __restrict__ Tem, const double* __restrict__ mu_fluid,
    const double* __restrict__ mu_e, const double* __restrict__ mu_sgs, double* __restrict__ Rmass,
    double* __restrict__ Rmom, double* __restrict__ Rener) {
    // phase-1 inputs, then phase-2 fluxes (18 per node) and He*gpvol (9)
    __shared__ double s_in[EPB][5][64];
    __shared__ double s_fl[EPB][18][64];
    __shared__ double s_hg[EPB][9][64];
    const int le = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int e = blockIdx.x * EPB + le;
    const bool active = e < nelem;
    const size_t np = (size_t)npoin, ne = (size_t)nelem;
    const int a = c_tab.lane2node[l];
    int n = 0;
    double rl = 0.0, rnl = 0.0, mufl = 0.0, ul[3] = {0, 0, 0};
    if (active) {
        n = connec[(size_t)e + ne * a] - 1;
        rl = rho[n];
        rnl = rho_n[n];
        mufl = mu_fluid[n];
        s_in[le][0][l] = rl;
        s_in[le][1][l] = Tem[n];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            ul[d] = u[n + np * d];
            s_in[le][2 + d][l] = ul[d];
        }
    }
    __syncthreads();
    double h[3][3];
    double mue = 0.0, musgs = 0.0;
    if (active) {
        Lines L;
        lines_of(l, L);
        load_he(He, e, a, h);
        mue = mu_e[(size_t)e + ne * a];
        musgs = mu_sgs[(size_t)e + ne * a];
        const double mu_fgp = mufl + rnl * musgs;
        const double mu_egp = mue;
        double g[3], gradRho[3], gradT[3], gradU[3][3];
        const double* RR = s_in[le][0];
        const double* TT = s_in[le][1];
        giso(L, [&](int m) { return RR[m]; }, g);
        grad_phys(h, g, gradRho);
        giso(L, [&](int m) { return TT[m]; }, g);
        grad_phys(h, g, gradT);
#pragma unroll
        for (int id = 0; id < 3; ++id) {
            const double* UU = s_in[le][2 + id];
            giso(L, [&](int m) { return UU[m]; }, g);
#pragma unroll
            for (int jd = 0; jd < 3; ++jd) {
                double s = 0.0;
                s += h[jd][0] * g[0];
                s += h[jd][1] * g[1];
                s += h[jd][2] * g[2];
                gradU[id][jd] = s;
            }
        }
        const double divU = gradU[0][0] + gradU[1][1] + gradU[2][2];
        const double twoThirds = 2.0 / 3.0;
        double tau[3][3], tauU[3];
#pragma unroll
        for (int id = 0; id < 3; ++id) {
            double t = 0.0;
#pragma unroll
            for (int jd = 0; jd < 3; ++jd) {
                t = t + (mu_fgp + mu_egp) * (gradU[id][jd] + gradU[jd][id]) * ul[jd];
                tau[id][jd] = (mu_fgp + mu_egp) * (gradU[id][jd] + gradU[jd][id]);
            }
            tauU[id] = t - (mu_fgp)*twoThirds * divU * ul[id];
            tau[id][id] = tau[id][id] - (mu_fgp)*twoThirds * divU;
        }
        const double gv = gpvol[a + 64 * (size_t)e];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            s_fl[le][0 + d][l] = tau[0][d];
            s_fl[le][3 + d][l] = tau[1][d];
            s_fl[le][6 + d][l] = tau[2][d];
            s_fl[le][9 + d][l] = gradT[d];
            s_fl[le][12 + d][l] = gradRho[d];
            s_fl[le][15 + d][l] = tauU[d];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) s_hg[le][i + 3 * j][l] = h[i][j] * gv;
    }
    __syncthreads();
    if (!active) return;
    const int i = l & 3, j = (l >> 2) & 3, k = l >> 4;
    const double nu_e = c_rho * mue / rnl;
    const double kappa_e = mufl * Cp / Pr + c_ener * mue * Cp / Pr + Cp * rnl * musgs / 0.9;
    double divDr = 0.0, divDe = 0.0, divDm[3] = {0.0, 0.0, 0.0};
    const double(*F)[64] = s_fl[le];
    const double(*HG)[64] = s_hg[le];
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
        const int qx = ii + 4 * j + 16 * k, qy = i + 4 * ii + 16 * k, qz = i + 4 * j + 16 * ii;
        // dlxigp_ip(q,dir,iso) = l'_iso(xi_q) = D[ii][iso]
        const double dx = c_tab.D[ii][i], dy = c_tab.D[ii][j], dz = c_tab.D[ii][k];
#pragma unroll
        for (int id = 0; id < 3; ++id) {
            const double wx = HG[id + 0][qx] * dx;
            const double wy = HG[id + 3][qy] * dy;
            const double wz = HG[id + 6][qz] * dz;
            divDr = divDr + wx * F[12 + id][qx];
            divDr = divDr + wy * F[12 + id][qy];
            divDr = divDr + wz * F[12 + id][qz];
            divDe = divDe + wx * (F[15 + id][qx] + kappa_e * F[9 + id][qx]);
            divDe = divDe + wy * (F[15 + id][qy] + kappa_e * F[9 + id][qy]);
            divDe = divDe + wz * (F[15 + id][qz] + kappa_e * F[9 + id][qz]);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                divDm[c] = divDm[c] + wx * F[3 * c + id][qx];
                divDm[c] = divDm[c] + wy * F[3 * c + id][qy];
                divDm[c] = divDm[c] + wz * F[3 * c + id][qz];
            }
        }
    }
    unsafeAtomicAdd(&Rmass[n], nu_e * divDr);
    unsafeAtomicAdd(&Rener[n], divDe);
    unsafeAtomicAdd(&Rmom[n], divDm[0]);
    unsafeAtomicAdd(&Rmom[n + np], divDm[1]);
    unsafeAtomicAdd(&Rmom[n + 2 * np], divDm[2]);
}

int zero(double* p, size_t n) {
    SOD_HIP(hipMemsetAsync(p, 0, n * sizeof(double), rt().stream));
    return SOD_OK;
}

template <bool TOTAL_H>
int convec_entry(int nelem, int npoin, const int* connec, const double* He, const double* gpvol,
                 const double* dlxigp_ip, const int* invAtoIJK, const int* gmshAtoI,
                 const int* gmshAtoJ, const int* gmshAtoK, const double* u, const double* q,
                 const double* rho, const double* pr, const double* E, double* Rmass,
                 double* Rmom, double* Rener) {
    SOD_TRY(ensure_init());
    if (nelem < 0 || npoin < 0) return fail(SOD_ERR_ARG, "negative nelem/npoin");
    SOD_TRY(check_tables(dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK));
    const int *d_connec;
    const double *d_He, *d_gpvol, *d_u, *d_q, *d_rho, *d_pr, *d_E;
    double *d_Rmass, *d_Rmom, *d_Rener;
    SOD_TRY(resolve_t(connec, "connec", &d_connec));
    SOD_TRY(resolve_t(He, "He", &d_He));
    SOD_TRY(resolve_t(gpvol, "gpvol", &d_gpvol));
    SOD_TRY(resolve_t(u, "u", &d_u));
    SOD_TRY(resolve_t(q, "q", &d_q));
    SOD_TRY(resolve_t(rho, "rho", &d_rho));
    SOD_TRY(resolve_t(pr, "pr", &d_pr));
    SOD_TRY(resolve_t(E, "E", &d_E));
    SOD_TRY(resolve_w(Rmass, "Rmass", &d_Rmass));
    SOD_TRY(resolve_w(Rmom, "Rmom", &d_Rmom));
    SOD_TRY(resolve_w(Rener, "Rener", &d_Rener));
    SOD_TRY(zero(d_Rmass, npoin));
    SOD_TRY(zero(d_Rmom, 3 * (size_t)npoin));
    SOD_TRY(zero(d_Rener, npoin));
    if (nelem > 0)
        hipLaunchKernelGGL(k_full_convec<TOTAL_H>, dim3(grid_for(nelem, EPB)), dim3(64 * EPB), 0,
                           rt().stream, nelem, npoin, d_connec, d_He, d_gpvol, d_u, d_q, d_rho, d_pr,
                           d_E, d_Rmass, d_Rmom, d_Rener);
    return finish_call();
}

}  // namespace

// Launchers reused by the step driver (device pointers, no checks).
void launch_full_convec(bool total_h, int nelem, int npoin, const int* connec, const double* He,
                        const double* gpvol, const double* u, const double* q, const double* rho,
                        const double* pr, const double* E, double* Rmass, double* Rmom,
                        double* Rener) {
    if (nelem <= 0) return;
    SOD_PROF("full_convec");
    if (total_h)
        hipLaunchKernelGGL(k_full_convec<true>, dim3(grid_for(nelem, EPB)), dim3(64 * EPB), 0,
                           rt().stream, nelem, npoin, connec, He, gpvol, u, q, rho, pr, E, Rmass,
                           Rmom, Rener);
    else
        hipLaunchKernelGGL(k_full_convec<false>, dim3(grid_for(nelem, EPB)), dim3(64 * EPB), 0,
                           rt().stream, nelem, npoin, connec, He, gpvol, u, q, rho, pr, E, Rmass,
                           Rmom, Rener);
}

void launch_scalar_convec(int nelem, int npoin, const int* connec, const double* He,
                          const double* gpvol, const double* q, const double* eta, const double* u,
                          double* R) {
    if (nelem <= 0) return;
    SOD_PROF("scalar_convec");
    hipLaunchKernelGGL(k_scalar_convec, dim3(grid_for(nelem, EPB)), dim3(64 * EPB), 0, rt().stream,
                       nelem, npoin, connec, He, gpvol, q, eta, u, R);
}

void launch_full_diffusion(int nelem, int npoin, const int* connec, const double* He,
                           const double* gpvol, double Cp, double Pr, const double* rho_n,
                           const double* rho, const double* u, const double* Tem,
                           const double* mu_fluid, const double* mu_e, const double* mu_sgs,
                           double* Rmass, double* Rmom, double* Rener) {
    if (nelem <= 0) return;
    SOD_PROF("full_diffusion");
    const sod_numerical_params& p = rt().params;
    hipLaunchKernelGGL(k_full_diffusion, dim3(grid_for(nelem, EPB)), dim3(64 * EPB), 0, rt().stream,
                       nelem, npoin, connec, He, gpvol, Cp, Pr, p.c_rho, p.c_ener, rho_n, rho, u,
                       Tem, mu_fluid, mu_e, mu_sgs, Rmass, Rmom, Rener);
}

}  // namespace sod

using namespace sod;

extern "C" {

int sod_full_convec_ijk(int nelem, int npoin, const int* connec, const double* Ngp,
                        const double* dNgp, const double* He, const double* gpvol,
                        const double* dlxigp_ip, const double* xgp, const int* invAtoIJK,
                        const int* gmshAtoI, const int* gmshAtoJ, const int* gmshAtoK,
                        const double* u, const double* q, const double* rho, const double* pr,
                        const double* E, double* Rmass, double* Rmom, double* Rener) {
    (void)Ngp;
    (void)dNgp;
    (void)xgp;
    return convec_entry<false>(nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK, gmshAtoI,
                               gmshAtoJ, gmshAtoK, u, q, rho, pr, E, Rmass, Rmom, Rener);
}

int sod_full_convec_ijk_H(int nelem, int npoin, const int* connec, const double* Ngp,
                          const double* dNgp, const double* He, const double* gpvol,
                          const double* dlxigp_ip, const double* xgp, const int* invAtoIJK,
                          const int* gmshAtoI, const int* gmshAtoJ, const int* gmshAtoK,
                          const double* u, const double* q, const double* rho, const double* pr,
                          const double* E, double* Rmass, double* Rmom, double* Rener) {
    (void)Ngp;
    (void)dNgp;
    (void)xgp;
    return convec_entry<true>(nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK, gmshAtoI,
                              gmshAtoJ, gmshAtoK, u, q, rho, pr, E, Rmass, Rmom, Rener);
}

int sod_generic_scalar_convec_ijk(int nelem, int npoin, const int* connec, const double* Ngp,
                                  const double* dNgp, const double* He, const double* gpvol,
                                  const double* dlxigp_ip, const double* xgp,
                                  const int* invAtoIJK, const int* gmshAtoI, const int* gmshAtoJ,
                                  const int* gmshAtoK, const double* q, const double* eta,
                                  const double* u, double* Rconvec) {
    (void)Ngp;
    (void)dNgp;
    (void)xgp;
    SOD_TRY(ensure_init());
    if (nelem < 0 || npoin < 0) return fail(SOD_ERR_ARG, "negative nelem/npoin");
    SOD_TRY(check_tables(dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK));
    const int* d_connec;
    const double *d_He, *d_gpvol, *d_q, *d_eta, *d_u;
    double* d_R;
    SOD_TRY(resolve_t(connec, "connec", &d_connec));
    SOD_TRY(resolve_t(He, "He", &d_He));
    SOD_TRY(resolve_t(gpvol, "gpvol", &d_gpvol));
    SOD_TRY(resolve_t(q, "q", &d_q));
    SOD_TRY(resolve_t(eta, "eta", &d_eta));
    SOD_TRY(resolve_t(u, "u", &d_u));
    SOD_TRY(resolve_w(Rconvec, "Rconvec", &d_R));
    SOD_TRY(zero(d_R, npoin));
    launch_scalar_convec(nelem, npoin, d_connec, d_He, d_gpvol, d_q, d_eta, d_u, d_R);
    return finish_call();
}

int sod_full_diffusion_ijk(int nelem, int npoin, const int* connec, const double* Ngp,
                           const double* He, const double* gpvol, const double* dlxigp_ip,
                           const int* invAtoIJK, const int* gmshAtoI, const int* gmshAtoJ,
                           const int* gmshAtoK, double Cp, double Pr, const double* rho_n,
                           const double* rho, const double* u, const double* Tem,
                           const double* mu_fluid, const double* mu_e, const double* mu_sgs,
                           const double* Ml, double* Rmass, double* Rmom, double* Rener) {
    (void)Ngp;
    (void)Ml;
    SOD_TRY(ensure_init());
    if (nelem < 0 || npoin < 0) return fail(SOD_ERR_ARG, "negative nelem/npoin");
    SOD_TRY(check_tables(dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK));
    const int* d_connec;
    const double *d_He, *d_gpvol, *d_rho_n, *d_rho, *d_u, *d_T, *d_muf, *d_mue, *d_musgs;
    double *d_Rmass, *d_Rmom, *d_Rener;
    SOD_TRY(resolve_t(connec, "connec", &d_connec));
    SOD_TRY(resolve_t(He, "He", &d_He));
    SOD_TRY(resolve_t(gpvol, "gpvol", &d_gpvol));
    SOD_TRY(resolve_t(rho_n, "rho_n", &d_rho_n));
    SOD_TRY(resolve_t(rho, "rho", &d_rho));
    SOD_TRY(resolve_t(u, "u", &d_u));
    SOD_TRY(resolve_t(Tem, "Tem", &d_T));
    SOD_TRY(resolve_t(mu_fluid, "mu_fluid", &d_muf));
    SOD_TRY(resolve_t(mu_e, "mu_e", &d_mue));
    SOD_TRY(resolve_t(mu_sgs, "mu_sgs", &d_musgs));
    SOD_TRY(resolve_w(Rmass, "Rmass", &d_Rmass));
    SOD_TRY(resolve_w(Rmom, "Rmom", &d_Rmom));
    SOD_TRY(resolve_w(Rener, "Rener", &d_Rener));
    SOD_TRY(zero(d_Rmass, npoin));
    SOD_TRY(zero(d_Rmom, 3 * (size_t)npoin));
    SOD_TRY(zero(d_Rener, npoin));
    launch_full_diffusion(nelem, npoin, d_connec, d_He, d_gpvol, Cp, Pr, d_rho_n, d_rho, d_u, d_T,
                          d_muf, d_mue, d_musgs, d_Rmass, d_Rmom, d_Rener);
    return finish_call();
}

}  // extern "C"
