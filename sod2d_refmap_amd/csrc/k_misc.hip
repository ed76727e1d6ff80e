// Pointwise and per-element routines of the path with the reference's
// argument lists: Sutherland viscosity, lumped-mass solve, entropy viscosity
// (smart_visc_spectral), CFL time step, kinetic-energy average, and the
// init-time geometry routines (Jacobians, lumped mass, element lengths).
#include <algorithm>
#include <cstdint>
#include <vector>

#include "stage_common.hpp"

namespace sod {

namespace {

constexpr int BLK = 256;

// ------------------------------------------------------------ reductions
// Deterministic two-pass sum: fixed-size grid of partial sums, then one block.
constexpr int RED_BLOCKS = 1024;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <int OP>   // 0 sum, 1 max
__device__ __forceinline__ double block_reduce(double v, double* s_tmp) {
    v = OP == 0 ? wave_sum(v) : wave_max(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) s_tmp[w] = v;
    __syncthreads();
    double r = OP == 0 ? 0.0 : -HUGE_VAL;
    if (threadIdx.x == 0) {
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r = OP == 0 ? r + s_tmp[i] : fmax(r, s_tmp[i]);
    }
    return r;
}

// partial[b] = sum over i = b, b+G, ... of eta(lpoin_w(i))
__global__ __launch_bounds__(BLK) void k_eta_partial_sum(int npoin_w, const int* __restrict__ lpoin_w,
                                                         const double* __restrict__ eta,
                                                         double* __restrict__ partial) {
    __shared__ double s_tmp[BLK / 64];
    double v = 0.0;
    for (int i = blockIdx.x * BLK + threadIdx.x; i < npoin_w; i += gridDim.x * BLK)
        v += eta[lpoin_w[i] - 1];
    double r = block_reduce<0>(v, s_tmp);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ __launch_bounds__(BLK) void k_final_sum(int n, const double* __restrict__ partial,
                                                   double* __restrict__ out) {
    __shared__ double s_tmp[BLK / 64];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += BLK) v += partial[i];
    double r = block_reduce<0>(v, s_tmp);
    if (threadIdx.x == 0) *out = r;
}

// partial[b] = max |eta - mean| ; mean = sum[0] / npoin_w
__global__ __launch_bounds__(BLK) void k_eta_partial_dev(int npoin_w, const int* __restrict__ lpoin_w,
                                                         const double* __restrict__ eta,
                                                         const double* __restrict__ sum,
                                                         double* __restrict__ partial) {
    __shared__ double s_tmp[BLK / 64];
    const double mean = sum[0] / (double)npoin_w;
    double v = 0.0;
    for (int i = blockIdx.x * BLK + threadIdx.x; i < npoin_w; i += gridDim.x * BLK)
        v = fmax(v, fabs(eta[lpoin_w[i] - 1] - mean));
    double r = block_reduce<1>(v, s_tmp);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ __launch_bounds__(BLK) void k_final_max(int n, const double* __restrict__ partial,
                                                   double* __restrict__ out) {
    __shared__ double s_tmp[BLK / 64];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += BLK) v = fmax(v, partial[i]);
    double r = block_reduce<1>(v, s_tmp);
    if (threadIdx.x == 0) *out = r;
}

// ------------------------------------------------------------ pointwise
__global__ __launch_bounds__(BLK) void k_sutherland(int npoin, const double* __restrict__ Tem,
                                                    const double* __restrict__ mu_factor,
                                                    double* __restrict__ mu_fluid) {
    const int n = blockIdx.x * BLK + threadIdx.x;
    if (n < npoin) mu_fluid[n] = sutherland(mu_factor[n], Tem[n]);
}

__global__ __launch_bounds__(BLK) void k_lumped(int npoin, int npoin_w, int ncomp,
                                                const int* __restrict__ lpoin_w,
                                                const double* __restrict__ Ml, double* __restrict__ R) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= npoin_w) return;
    const int n = lpoin_w[i] - 1;
    const double m = Ml[n];
    for (int c = 0; c < ncomp; ++c) R[n + (size_t)npoin * c] = R[n + (size_t)npoin * c] / m;
}

// ------------------------------------------------------ entropy viscosity
// smart_visc_spectral element loop, mod_entropy_viscosity.f90:68-122.  One
// wavefront per element; lane = lattice node.  The 3x3x3 filter runs over
// "geometric" positions (convertIJK, mod_filters.f90:42-51) whose end
// mirrors are kept exactly as the reference builds them.
__global__ __launch_bounds__(256) void k_smart_visc(
    int nelem, int npoin, const int* __restrict__ connec, const double* __restrict__ Reta,
    const double* __restrict__ gpvol, double gamma_gas, const double* __restrict__ rho,
    const double* __restrict__ u, const double* __restrict__ csound, const double* __restrict__ Tem,
    const double* __restrict__ helem_l, const double* __restrict__ helem,
    const double* __restrict__ d_norm, int normalise, double Rgas, double ce, double cmax,
    double cglob, double* __restrict__ mu_e, double* __restrict__ mue_l) {
    __shared__ double s_mue[4][64];   // by gmsh node
    const int le = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int e_raw = blockIdx.x * 4 + le;
    const bool active = e_raw < nelem;
    const int e = active ? e_raw : nelem - 1;   // inactive waves mirror a valid element
    const size_t np = (size_t)npoin, ne = (size_t)nelem;
    const double norm = normalise ? d_norm[0] : 1.0;
    const int a = l;   // this phase works in gmsh node order
    const int n = connec[(size_t)e + ne * a] - 1;
    const double J = gpvol[a + 64 * (size_t)e] / c_tab.wgp[a];
    const double u0 = u[n], u1 = u[n + np], u2 = u[n + 2 * np];
    const double vmag = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
    const double minJe = wave_min(fmin(1000000.0, J));
    const double maxJe = wave_max(fmax(0.0, J));
    const double maxV = wave_max(vmag);
    const double maxC = wave_max(csound[n]);
    const double M = maxV / maxC;
    double M15 = M * M;   // M**15 by repeated multiplication
    M15 = M15 * M15 * M;  // M^5
    M15 = M15 * M15 * M15;
    const double v_pi = 3.14159265358979323846;
    const double ceM = fmax(tanh(M15 * v_pi), ce);
    const double r = minJe / maxJe;
    double ced = fmax(1.0 - r * r, ce);
    ced = fmax(ced, ceM);
    const double aux3 = sqrt(Rgas * gamma_gas * Tem[n]);
    const double betae = wave_max((rho[n] * helem[e]) * (cmax / 3.0) * (vmag + aux3));
    const double R1 = rho[n] * fabs(Reta[n]) / norm;
    const double hl = helem_l[(size_t)e + ne * a];
    const double Ve = ced * R1 * (hl * hl);
    const double ml = cglob * fmin(Ve, betae);
    if (active) mue_l[(size_t)e + ne * a] = ml;
    s_mue[le][a] = ml;
    __syncthreads();
    if (!active) return;
    // filter: lane -> geometric position (gi,gj,gk) in 1..4
    // convertIJK(0:5) = {3,1,3,4,2,3} maps geometric -> sod2d index
    const int conv[6] = {3, 1, 3, 4, 2, 3};
    const int gi = (l & 3) + 1, gj = ((l >> 2) & 3) + 1, gk = (l >> 4) + 1;
    const double w3[3] = {0.25, 0.5, 0.25};
    double acc = 0.0;
#pragma unroll
    for (int ll = -1; ll <= 1; ++ll)
#pragma unroll
        for (int mm = -1; mm <= 1; ++mm)
#pragma unroll
            for (int nn = -1; nn <= 1; ++nn) {
                const int si = conv[gi + ll] - 1, sj = conv[gj + mm] - 1, sk = conv[gk + nn] - 1;
                const int b = c_tab.lane2node[si + 4 * sj + 16 * sk];
                acc = acc + w3[ll + 1] * w3[mm + 1] * w3[nn + 1] * s_mue[le][b];
            }
    const int so = c_tab.lane2node[(conv[gi] - 1) + 4 * (conv[gj] - 1) + 16 * (conv[gk] - 1)];
    mu_e[(size_t)e + ne * so] = acc;
}

__global__ void k_set_double(double* p, double v) {
    if (threadIdx.x == 0) *p = v;
}

// ------------------------------------------------------------- CFL dt
// adapt_dt_cfl, mod_time_ops.f90:30-57; per-element minima folded into a
// global minimum with an integer atomic on the (positive) double's bits.
__global__ __launch_bounds__(256) void k_adapt_dt(int nelem, int npoin, const int* __restrict__ connec,
                                                  const double* __restrict__ helem,
                                                  const double* __restrict__ u,
                                                  const double* __restrict__ csound, double cfl_conv,
                                                  int has_diff, double cfl_diff,
                                                  const double* __restrict__ mu_fluid,
                                                  const double* __restrict__ mu_sgs,
                                                  const double* __restrict__ rho,
                                                  const int* __restrict__ run_flag,
                                                  unsigned long long* __restrict__ out_bits) {
    if (run_flag != nullptr && *run_flag == 0) return;
    // A block takes tiles of 16 consecutive elements: thread (el, a0) reads
    // connec(e, a0 + 16r), r = 0..3, so each group of 16 lanes reads one
    // contiguous run of connec and every lane has four independent gathers in
    // flight.  Maxima over an element's 64 nodes are exact in any order.
    __shared__ double s_L[4][16], s_M[4][16];
    const size_t np = (size_t)npoin, ne = (size_t)nelem;
    const int el = threadIdx.x & 15, a0 = threadIdx.x >> 4;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // each block takes a contiguous chunk of tiles; blocks of one XCD
    // take neighbouring chunks (shared nodes stay in that XCD's L2)
    const int ntiles = (nelem + 15) / 16;
    const int chunk = (ntiles + gridDim.x - 1) / gridDim.x;
    const int t0 = xcd_block(blockIdx.x, gridDim.x) * chunk, t1 = min(ntiles, t0 + chunk);
    double best = HUGE_VAL;
    for (int t = t0; t < t1; ++t) {
        const int e = t * 16 + el;
        double L3 = 0.0, mx = 0.0;
        if (e < nelem) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const size_t ea = (size_t)e + ne * (a0 + 16 * r);
                const int n = connec[ea] - 1;
                double umag = fabs(u[n]);
                umag = fmax(umag, fabs(u[n + np]));
                umag = fmax(umag, fabs(u[n + 2 * np]));
                L3 = fmax(L3, umag + csound[n]);
                if (has_diff) mx = fmax(mx, (rho[n] * mu_sgs[ea] + mu_fluid[n]) / rho[n]);
            }
        }
        L3 = fmax(L3, __shfl_xor(L3, 16));
        L3 = fmax(L3, __shfl_xor(L3, 32));
        mx = fmax(mx, __shfl_xor(mx, 16));
        mx = fmax(mx, __shfl_xor(mx, 32));
        if (lane < 16) {
            s_L[w][lane] = L3;
            s_M[w][lane] = mx;
        }
        __syncthreads();
        if (threadIdx.x < 16 && e < nelem) {
            L3 = fmax(fmax(s_L[0][el], s_L[1][el]), fmax(s_L[2][el], s_L[3][el]));
            mx = fmax(fmax(s_M[0][el], s_M[1][el]), fmax(s_M[2][el], s_M[3][el]));
            const double hh = helem[e] / 7.0;
            double dt_l = cfl_conv * hh / L3;
            if (has_diff) dt_l = fmin(dt_l, cfl_diff * (hh * hh) / mx);
            best = fmin(best, dt_l);
        }
        __syncthreads();
    }
    if (threadIdx.x < 64) {
        best = wave_min(best);
        if (threadIdx.x == 0 && best < HUGE_VAL)
            atomicMin(out_bits, (unsigned long long)__double_as_longlong(best));
    }
}

// Per-node minimum element size h_min(a) = min over the elements e holding
// node a of helem(e), as the bits of a positive double (ordered like the
// value); nodes no element holds keep +inf.  Skipped when `run_flag` points
// at 0.
constexpr unsigned long long kInfBits = 0x7FF0000000000000ull;

__global__ __launch_bounds__(256) void k_fill_u64(long n, unsigned long long* __restrict__ p,
                                                  unsigned long long v, const int* __restrict__ run_flag) {
    if (run_flag != nullptr && *run_flag == 0) return;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) p[i] = v;
}

__global__ __launch_bounds__(256) void k_hmin_build(int nelem, const int* __restrict__ connec,
                                                    const double* __restrict__ helem,
                                                    unsigned long long* __restrict__ hmin_bits,
                                                    const int* __restrict__ run_flag) {
    if (run_flag != nullptr && *run_flag == 0) return;
    const long n = (long)nelem * 64;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
        const int e = (int)(t % nelem);   // connec(e, a) at e + nelem*a: t runs through it in memory order
        atomicMin(hmin_bits + (connec[t] - 1), (unsigned long long)__double_as_longlong(helem[e]));
    }
}

// adapt_dt_cfl regrouped by node (node_dt): one pass over the nodes with
// their cached h_min, no connectivity gather.  Skipped when `skip_flag`
// points at a non-zero value (mu_sgs non-zero: the element form runs).
__global__ __launch_bounds__(256) void k_adapt_dt_nodes(int npoin, const double* __restrict__ hmin,
                                                        const double* __restrict__ u,
                                                        const double* __restrict__ csound, double cfl_conv,
                                                        int has_diff, double cfl_diff,
                                                        const double* __restrict__ mu_fluid,
                                                        const double* __restrict__ rho,
                                                        const int* __restrict__ skip_flag,
                                                        unsigned long long* __restrict__ out_bits) {
    __shared__ double s_m[4];
    if (skip_flag != nullptr && *skip_flag != 0) return;
    const size_t np = (size_t)npoin;
    double best = HUGE_VAL;
    for (int n = blockIdx.x * 256 + threadIdx.x; n < npoin; n += gridDim.x * 256) {
        const double h = hmin[n];
        if (!(h < HUGE_VAL)) continue;   // a node no element holds
        best = fmin(best, node_dt(h, u[n], u[n + np], u[n + 2 * np], csound[n], has_diff ? mu_fluid[n] : 0.0,
                                  has_diff ? rho[n] : 1.0, cfl_conv, has_diff, cfl_diff));
    }
    best = wave_min(best);
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double v = fmin(fmin(s_m[0], s_m[1]), fmin(s_m[2], s_m[3]));
        if (v < HUGE_VAL) atomicMin(out_bits, (unsigned long long)__double_as_longlong(v));
    }
}

// Whether the cached h_min of the drop-in CFL entry point still belongs to
// (connec, helem): every entry of both arrays enters an exact hash
// (a 64-bit sum, mod 2^64, of a mixing function of each value's bits and
// position, so that the order of the additions does not matter).
// k_dt_hash leaves its block's two sums (connec, helem) in
// part[2b], part[2b + 1]; k_dt_cache_check adds the blocks' sums, compares
// them with the stored fingerprint fp[0..1], sets *stale (or `force`) and
// stores them.  The arrays are read 16 bytes per load when their addresses
// allow it (vec), four independent loads per thread and round.
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
    // splitmix64 finaliser (Steele, Lea and Flood, OOPSLA 2014)
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebULL;
    x ^= x >> 31;
    return x;
}

__device__ __forceinline__ unsigned long long conn_term(int v, long t) {
    return mix64((unsigned long long)(unsigned)v ^ mix64((unsigned long long)t));
}
__device__ __forceinline__ unsigned long long helem_term(double v, long t) {
    return mix64((unsigned long long)__double_as_longlong(v) ^ mix64((unsigned long long)t + 0x9e37ULL));
}
__device__ __forceinline__ unsigned long long conn4_terms(int4 v, long g) {
    const long t = 4 * g;
    return conn_term(v.x, t) + conn_term(v.y, t + 1) + conn_term(v.z, t + 2) + conn_term(v.w, t + 3);
}
__device__ __forceinline__ unsigned long long helem2_terms(double2 v, long g) {
    return helem_term(v.x, 2 * g) + helem_term(v.y, 2 * g + 1);
}

constexpr int kDtHashBlocks = 1024;

__global__ __launch_bounds__(256) void k_dt_hash(long n_conn, const int* __restrict__ connec, int nelem,
                                                 const double* __restrict__ helem, int vec,
                                                 unsigned long long* __restrict__ part) {
    __shared__ unsigned long long s_a[4], s_b[4];
    unsigned long long a = 0, b = 0;
    const long stride = (long)gridDim.x * 256;
    const long t0 = (long)blockIdx.x * 256 + threadIdx.x;
    long done_c = 0, done_h = 0;   // entries covered by the 16-byte loops
    if (vec) {
        const int4* c4 = reinterpret_cast<const int4*>(connec);
        const long n4 = n_conn / 4;
        long g = t0;
        for (; g + 3 * stride < n4; g += 4 * stride) {
            const int4 v0 = c4[g], v1 = c4[g + stride], v2 = c4[g + 2 * stride], v3 = c4[g + 3 * stride];
            a += conn4_terms(v0, g) + conn4_terms(v1, g + stride) + conn4_terms(v2, g + 2 * stride) +
                 conn4_terms(v3, g + 3 * stride);
        }
        for (; g < n4; g += stride) a += conn4_terms(c4[g], g);
        const double2* h2 = reinterpret_cast<const double2*>(helem);
        const long n2 = nelem / 2;
        g = t0;
        for (; g + 3 * stride < n2; g += 4 * stride) {
            const double2 v0 = h2[g], v1 = h2[g + stride], v2 = h2[g + 2 * stride], v3 = h2[g + 3 * stride];
            b += helem2_terms(v0, g) + helem2_terms(v1, g + stride) + helem2_terms(v2, g + 2 * stride) +
                 helem2_terms(v3, g + 3 * stride);
        }
        for (; g < n2; g += stride) b += helem2_terms(h2[g], g);
        done_c = 4 * n4;
        done_h = 2 * n2;
    }
    for (long t = done_c + t0; t < n_conn; t += stride) a += conn_term(connec[t], t);
    for (long t = done_h + t0; t < nelem; t += stride) b += helem_term(helem[t], t);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_a[threadIdx.x >> 6] = a;
        s_b[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s_a[0] + s_a[1] + s_a[2] + s_a[3];
        part[2 * blockIdx.x + 1] = s_b[0] + s_b[1] + s_b[2] + s_b[3];
    }
}

__global__ __launch_bounds__(256) void k_dt_cache_check(int nblk, const unsigned long long* __restrict__ part,
                                                        unsigned long long* __restrict__ fp,
                                                        int* __restrict__ stale, int force) {
    __shared__ unsigned long long s_a[4], s_b[4];
    unsigned long long a = 0, b = 0;
    for (int q = threadIdx.x; q < nblk; q += 256) {
        a += part[2 * q];
        b += part[2 * q + 1];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_a[threadIdx.x >> 6] = a;
        s_b[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    a = s_a[0] + s_a[1] + s_a[2] + s_a[3];
    b = s_b[0] + s_b[1] + s_b[2] + s_b[3];
    const bool same = !force && fp[0] == a && fp[1] == b;
    *stale = same ? 0 : 1;
    fp[0] = a;
    fp[1] = b;
}

// flag = 1 if any of the n values is non-zero
__global__ __launch_bounds__(256) void k_any_nonzero(long n, const double* __restrict__ v, int* __restrict__ flag) {
    int nz = 0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) nz |= v[i] != 0.0;
    if (__any(nz) && (threadIdx.x & 63) == 0) *flag = 1;
}

__global__ void k_zero_ints(int n, int* __restrict__ p) {
    for (int t = threadIdx.x; t < n; t += blockDim.x) p[t] = 0;
}

// adapt_dt_cfl of each instance of a flattened ensemble in node form
// (node_dt, mu_sgs = 0): node n belongs to instance node_inst[n] and the
// nodes of an instance are contiguous.  A wavefront takes 512 consecutive
// nodes (8 per lane); when they all belong to one instance it folds their
// minimum with one atomic, otherwise each lane folds its own.
__global__ __launch_bounds__(256) void k_adapt_dt_nodes_inst(int npoin, const double* __restrict__ hmin,
                                                             const double* __restrict__ u,
                                                             const double* __restrict__ csound, double cfl_conv,
                                                             int has_diff, double cfl_diff,
                                                             const double* __restrict__ mu_fluid,
                                                             const double* __restrict__ rho,
                                                             const int* __restrict__ node_inst,
                                                             unsigned long long* __restrict__ out_bits) {
    const size_t np = (size_t)npoin;
    const int lane = threadIdx.x & 63;
    const long nwaves = ((long)npoin + 511) / 512;
    for (long wv = (long)blockIdx.x * 4 + (threadIdx.x >> 6); wv < nwaves; wv += (long)gridDim.x * 4) {
        const long n0 = wv * 512, n1 = min((long)npoin, n0 + 512);
        const bool one = node_inst[n0] == node_inst[n1 - 1];
        double best = HUGE_VAL;
        for (long n = n0 + lane; n < n1; n += 64) {
            const double h = hmin[n];
            if (!(h < HUGE_VAL)) continue;   // a node no element holds
            const double v = node_dt(h, u[n], u[n + np], u[n + 2 * np], csound[n], has_diff ? mu_fluid[n] : 0.0,
                                     has_diff ? rho[n] : 1.0, cfl_conv, has_diff, cfl_diff);
            if (one)
                best = fmin(best, v);
            else
                atomicMin(out_bits + node_inst[n], (unsigned long long)__double_as_longlong(v));
        }
        if (one) {
            best = wave_min(best);
            if (lane == 0 && best < HUGE_VAL)
                atomicMin(out_bits + node_inst[n0], (unsigned long long)__double_as_longlong(best));
        }
    }
}

// adapt_dt_cfl of each instance of a flattened ensemble: element e belongs to
// instance elem_inst[e] and its dt folds into out_bits[elem_inst[e]].  Same
// per-element arithmetic as k_adapt_dt; a tile whose 16 elements share one
// instance folds its minimum with one atomic.
__global__ __launch_bounds__(256) void k_adapt_dt_inst(int nelem, int npoin, const int* __restrict__ connec,
                                                       const double* __restrict__ helem,
                                                       const double* __restrict__ u,
                                                       const double* __restrict__ csound, double cfl_conv,
                                                       int has_diff, double cfl_diff,
                                                       const double* __restrict__ mu_fluid,
                                                       const double* __restrict__ mu_sgs,
                                                       const double* __restrict__ rho,
                                                       const int* __restrict__ elem_inst,
                                                       unsigned long long* __restrict__ out_bits) {
    __shared__ double s_L[4][16], s_M[4][16];
    const size_t np = (size_t)npoin, ne = (size_t)nelem;
    const int el = threadIdx.x & 15, a0 = threadIdx.x >> 4;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ntiles = (nelem + 15) / 16;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int e = t * 16 + el;
        double L3 = 0.0, mx = 0.0;
        if (e < nelem) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const size_t ea = (size_t)e + ne * (a0 + 16 * r);
                const int n = connec[ea] - 1;
                double umag = fabs(u[n]);
                umag = fmax(umag, fabs(u[n + np]));
                umag = fmax(umag, fabs(u[n + 2 * np]));
                L3 = fmax(L3, umag + csound[n]);
                if (has_diff) mx = fmax(mx, (rho[n] * mu_sgs[ea] + mu_fluid[n]) / rho[n]);
            }
        }
        L3 = fmax(L3, __shfl_xor(L3, 16));
        L3 = fmax(L3, __shfl_xor(L3, 32));
        mx = fmax(mx, __shfl_xor(mx, 16));
        mx = fmax(mx, __shfl_xor(mx, 32));
        if (lane < 16) {
            s_L[w][lane] = L3;
            s_M[w][lane] = mx;
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            double dt_l = HUGE_VAL;
            int inst = -1;
            if (threadIdx.x < 16 && e < nelem) {
                L3 = fmax(fmax(s_L[0][el], s_L[1][el]), fmax(s_L[2][el], s_L[3][el]));
                mx = fmax(fmax(s_M[0][el], s_M[1][el]), fmax(s_M[2][el], s_M[3][el]));
                const double hh = helem[e] / 7.0;
                dt_l = cfl_conv * hh / L3;
                if (has_diff) dt_l = fmin(dt_l, cfl_diff * (hh * hh) / mx);
                inst = elem_inst[e];
            }
            const int last = min(t * 16 + 15, nelem - 1);
            const bool one = elem_inst[t * 16] == elem_inst[last];
            if (one) {
                double v = dt_l;
#pragma unroll
                for (int o = 8; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
                if (threadIdx.x == 0 && v < HUGE_VAL)
                    atomicMin(out_bits + elem_inst[t * 16], (unsigned long long)__double_as_longlong(v));
            } else if (inst >= 0) {
                atomicMin(out_bits + inst, (unsigned long long)__double_as_longlong(dt_l));
            }
        }
        __syncthreads();
    }
}

__global__ void k_set_doubles(int n, double* p, double v) {
    for (int t = threadIdx.x; t < n; t += blockDim.x) p[t] = v;
}

// ------------------------------------------------------------- EK
// volAvg_EK, mod_analysis.f90:10-44 (Ngp is the identity for collocated
// GLL nodes, so only the diagonal of Ngp is used).
__global__ __launch_bounds__(256) void k_ek_partial(int nelem, int npoin, const int* __restrict__ connec,
                                                    const double* __restrict__ gpvol,
                                                    const double* __restrict__ Ngp,
                                                    const double* __restrict__ rho,
                                                    const double* __restrict__ u,
                                                    double* __restrict__ partial) {
    __shared__ double s_tmp[4];
    const int le = threadIdx.x >> 6, l = threadIdx.x & 63;
    const size_t np = (size_t)npoin, ne = (size_t)nelem;
    double acc = 0.0;
    for (int e = blockIdx.x * 4 + le; e < nelem; e += gridDim.x * 4) {
        // lane = inode: sum over igaus of gpvol(igaus)*Ngp(igaus,inode)
        const int n = connec[(size_t)e + ne * l] - 1;
        const double u0 = u[n], u1 = u[n + np], u2 = u[n + 2 * np];
        const double ke = 0.5 * rho[n] * (u0 * u0 + u1 * u1 + u2 * u2);
        const double* gv = gpvol + 64 * (size_t)e;
        for (int g = 0; g < 64; ++g) acc += gv[g] * Ngp[g + 64 * l] * ke;
    }
    double r = block_reduce<0>(acc, s_tmp);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// ------------------------------------------------------- init geometry
// elem_jacobian, jacobian_oper.f90:35-136: one thread per (element, Gauss
// point); Je(i,j) = sum_a dNgp(i,a,g) coord(connec(e,a),j) over all 64 a.
__global__ __launch_bounds__(BLK) void k_elem_jacobian(int nelem, int npoin,
                                                       const int* __restrict__ connec,
                                                       const double* __restrict__ coord,
                                                       const double* __restrict__ dNgp,
                                                       const double* __restrict__ wgp,
                                                       double* __restrict__ gpvol,
                                                       double* __restrict__ He) {
    const long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= (long)nelem * 64) return;
    const int e = (int)(t >> 6), g = (int)(t & 63);
    const size_t np = (size_t)npoin, ne = (size_t)nelem;
    double Je[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < 64; ++a) {
        const int n = connec[(size_t)e + ne * a] - 1;
        const double x0 = coord[n], x1 = coord[n + np], x2 = coord[n + 2 * np];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double dn = dNgp[i + 3 * a + 192 * g];
            Je[i][0] += dn * x0;
            Je[i][1] += dn * x1;
            Je[i][2] += dn * x2;
        }
    }
    // cofactor inverse (jacobian_oper.f90:74-131)
    const double det = Je[0][0] * Je[1][1] * Je[2][2] + Je[0][1] * Je[1][2] * Je[2][0] +
                       Je[0][2] * Je[1][0] * Je[2][1] - Je[2][0] * Je[1][1] * Je[0][2] -
                       Je[2][1] * Je[1][2] * Je[0][0] - Je[2][2] * Je[1][0] * Je[0][1];
    const double inv = 1.0 / det;
    double* h = He + 9 * (size_t)g + 576 * (size_t)e;
    // He = adj(Je)/det, column-major He(i,j) at h[i + 3j]
    h[0 + 3 * 0] = inv * (Je[1][1] * Je[2][2] - Je[2][1] * Je[1][2]);
    h[0 + 3 * 1] = inv * -(Je[0][1] * Je[2][2] - Je[2][1] * Je[0][2]);
    h[0 + 3 * 2] = inv * (Je[0][1] * Je[1][2] - Je[1][1] * Je[0][2]);
    h[1 + 3 * 0] = inv * -(Je[1][0] * Je[2][2] - Je[2][0] * Je[1][2]);
    h[1 + 3 * 1] = inv * (Je[0][0] * Je[2][2] - Je[2][0] * Je[0][2]);
    h[1 + 3 * 2] = inv * -(Je[0][0] * Je[1][2] - Je[1][0] * Je[0][2]);
    h[2 + 3 * 0] = inv * (Je[1][0] * Je[2][1] - Je[2][0] * Je[1][1]);
    h[2 + 3 * 1] = inv * -(Je[0][0] * Je[2][1] - Je[2][0] * Je[0][1]);
    h[2 + 3 * 2] = inv * (Je[0][0] * Je[1][1] - Je[1][0] * Je[0][1]);
    gpvol[g + 64 * (size_t)e] = wgp[g] * det;
}

__global__ __launch_bounds__(BLK) void k_lumped_mass(int nelem, const int* __restrict__ connec,
                                                     const double* __restrict__ gpvol,
                                                     double* __restrict__ Ml) {
    const long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= (long)nelem * 64) return;
    const int e = (int)(t >> 6), a = (int)(t & 63);
    unsafeAtomicAdd(&Ml[connec[(size_t)e + (size_t)nelem * a] - 1], gpvol[a + 64 * (size_t)e]);
}

// char_length: min of the 12 corner-edge lengths (elem_hex.f90:45-75)
__global__ __launch_bounds__(BLK) void k_char_length(int nelem, int npoin, const int* __restrict__ connec,
                                                     const double* __restrict__ coord,
                                                     double* __restrict__ helem) {
    const int e = blockIdx.x * BLK + threadIdx.x;
    if (e >= nelem) return;
    const int ed[12][2] = {{1, 0}, {2, 1}, {3, 2}, {0, 3}, {5, 4}, {6, 5},
                           {7, 6}, {4, 7}, {4, 0}, {5, 1}, {6, 2}, {7, 3}};
    const size_t np = (size_t)npoin, ne = (size_t)nelem;
    double best = 1000000000000.0;
    for (int k = 0; k < 12; ++k) {
        const int n1 = connec[(size_t)e + ne * ed[k][0]] - 1, n0 = connec[(size_t)e + ne * ed[k][1]] - 1;
        const double d0 = coord[n1] - coord[n0], d1 = coord[n1 + np] - coord[n0 + np],
                     d2 = coord[n1 + 2 * np] - coord[n0 + 2 * np];
        best = fmin(best, sqrt(d0 * d0 + d1 * d1 + d2 * d2));
    }
    helem[e] = best;
}

// char_length_spectral: per node, min distance to the other 63 nodes
__global__ __launch_bounds__(256) void k_char_length_spectral(int nelem, int npoin,
                                                              const int* __restrict__ connec,
                                                              const double* __restrict__ coord,
                                                              double* __restrict__ helem_l) {
    __shared__ double s_x[4][3][64];
    const int le = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int e_raw = blockIdx.x * 4 + le;
    const bool active = e_raw < nelem;
    const int e = active ? e_raw : nelem - 1;
    const size_t np = (size_t)npoin, ne = (size_t)nelem;
    const int n = connec[(size_t)e + ne * l] - 1;
    s_x[le][0][l] = coord[n];
    s_x[le][1][l] = coord[n + np];
    s_x[le][2][l] = coord[n + 2 * np];
    __syncthreads();
    if (!active) return;
    double best = 1000000000000.0;
    for (int b = 0; b < 64; ++b) {
        if (b == l) continue;
        const double d0 = s_x[le][0][l] - s_x[le][0][b], d1 = s_x[le][1][l] - s_x[le][1][b],
                     d2 = s_x[le][2][l] - s_x[le][2][b];
        best = fmin(sqrt(d0 * d0 + d1 * d1 + d2 * d2), best);
    }
    helem_l[(size_t)e + ne * l] = best;
}

}  // namespace

// ----------------------------------------------------- launchers (device)
void launch_sutherland(int npoin, const double* Tem, const double* mu_factor, double* mu_fluid) {
    if (npoin <= 0) return;
    hipLaunchKernelGGL(k_sutherland, dim3(grid_for(npoin, BLK)), dim3(BLK), 0, rt().stream, npoin,
                       Tem, mu_factor, mu_fluid);
}

void launch_lumped(int npoin, int npoin_w, int ncomp, const int* lpoin_w, const double* Ml,
                   double* R) {
    if (npoin_w <= 0) return;
    hipLaunchKernelGGL(k_lumped, dim3(grid_for(npoin_w, BLK)), dim3(BLK), 0, rt().stream, npoin,
                       npoin_w, ncomp, lpoin_w, Ml, R);
}

// Entropy normalisation + element loop. `work` holds >= RED_BLOCKS + 2 doubles.
size_t smart_visc_scratch_doubles() { return RED_BLOCKS + 8; }

void launch_smart_visc(int nelem, int npoin, int npoin_w, const int* connec, const int* lpoin_w,
                       const double* Reta, const double* gpvol, double gamma_gas, const double* rho,
                       const double* u, const double* csound, const double* Tem, const double* eta,
                       const double* helem_l, const double* helem, double* mu_e, double* mue_l,
                       double* work) {
    const sod_numerical_params& p = rt().params;
    hipStream_t s = rt().stream;
    double* partial = work;
    double* sum = work + RED_BLOCKS;
    double* norm = work + RED_BLOCKS + 1;
    if (p.flag_normalise_entropy == 1) {
        hipLaunchKernelGGL(k_eta_partial_sum, dim3(RED_BLOCKS), dim3(BLK), 0, s, npoin_w, lpoin_w,
                           eta, partial);
        hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(BLK), 0, s, RED_BLOCKS, partial, sum);
        hipLaunchKernelGGL(k_eta_partial_dev, dim3(RED_BLOCKS), dim3(BLK), 0, s, npoin_w, lpoin_w,
                           eta, sum, partial);
        hipLaunchKernelGGL(k_final_max, dim3(1), dim3(BLK), 0, s, RED_BLOCKS, partial, norm);
    }
    if (nelem > 0)
        hipLaunchKernelGGL(k_smart_visc, dim3(grid_for(nelem, 4)), dim3(256), 0, s, nelem, npoin,
                           connec, Reta, gpvol, gamma_gas, rho, u, csound, Tem, helem_l, helem, norm,
                           p.flag_normalise_entropy == 1, p.Rgas_inf, p.ce, p.cmax, p.cglob, mu_e,
                           mue_l);
}

// Sum and maximum deviation of eta over the working nodes
// (mod_entropy_viscosity.f90:43-66): work[0] = sum, work[1] = max |eta -
// sum/npoin_w|; `work` holds >= 2 * RED_BLOCKS + 2 doubles.
void launch_eta_stats(int npoin_w, const int* lpoin_w, const double* eta, double* work) {
    hipStream_t s = rt().stream;
    double* partial = work + 2;
    hipLaunchKernelGGL(k_eta_partial_sum, dim3(RED_BLOCKS), dim3(BLK), 0, s, npoin_w, lpoin_w, eta, partial);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(BLK), 0, s, RED_BLOCKS, partial, work);
    hipLaunchKernelGGL(k_eta_partial_dev, dim3(RED_BLOCKS), dim3(BLK), 0, s, npoin_w, lpoin_w, eta, work, partial);
    hipLaunchKernelGGL(k_final_max, dim3(1), dim3(BLK), 0, s, RED_BLOCKS, partial, work + 1);
}

// Entropy normalisation after a fused step (mod_entropy_viscosity.f90:43-66):
// work[1] = max |eta - mean(eta)| over working nodes, from `npartial` block
// partial sums of eta.  `work` holds >= RED_BLOCKS + 2 doubles.
void launch_eta_norm(int npoin_w, const int* lpoin_w, const double* eta, const double* partial, int npartial,
                     double* work) {
    hipStream_t s = rt().stream;
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(BLK), 0, s, npartial, partial, work);
    hipLaunchKernelGGL(k_eta_partial_dev, dim3(RED_BLOCKS), dim3(BLK), 0, s, npoin_w, lpoin_w, eta, work, work + 2);
    hipLaunchKernelGGL(k_final_max, dim3(1), dim3(BLK), 0, s, RED_BLOCKS, work + 2, work + 1);
}

// Entropy viscosity after a fused step: the sum of eta over working nodes
// arrives as `npartial` block partials; `work` holds >= RED_BLOCKS + 2 doubles.
void launch_smart_visc_fused(int nelem, int npoin, int npoin_w, const int* connec, const int* lpoin_w,
                             const double* Reta, const double* gpvol, double gamma_gas, const double* rho,
                             const double* u, const double* csound, const double* Tem,
                             const double* helem_l, const double* helem, double* mu_e, double* mue_l,
                             const double* eta, const double* partial, int npartial, double* work) {
    const sod_numerical_params& p = rt().params;
    hipStream_t s = rt().stream;
    double* sum = work;
    double* norm = work + 1;
    double* dev_partial = work + 2;
    if (p.flag_normalise_entropy == 1) {
        hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(BLK), 0, s, npartial, partial, sum);
        hipLaunchKernelGGL(k_eta_partial_dev, dim3(RED_BLOCKS), dim3(BLK), 0, s, npoin_w, lpoin_w, eta, sum,
                           dev_partial);
        hipLaunchKernelGGL(k_final_max, dim3(1), dim3(BLK), 0, s, RED_BLOCKS, dev_partial, norm);
    }
    if (nelem > 0)
        hipLaunchKernelGGL(k_smart_visc, dim3(grid_for(nelem, 4)), dim3(256), 0, s, nelem, npoin, connec, Reta,
                           gpvol, gamma_gas, rho, u, csound, Tem, helem_l, helem, norm,
                           p.flag_normalise_entropy == 1, p.Rgas_inf, p.ce, p.cmax, p.cglob, mu_e, mue_l);
}

void launch_adapt_dt(int nelem, int npoin, const int* connec, const double* helem, const double* u,
                     const double* csound, double cfl_conv, int has_diff, double cfl_diff,
                     const double* mu_fluid, const double* mu_sgs, const double* rho,
                     double* d_dt_bits) {
    hipStream_t s = rt().stream;
    const double init = 100000000000000.0;
    // initial value written by a kernel so that the sequence stays capturable
    hipLaunchKernelGGL(k_set_double, dim3(1), dim3(64), 0, s, d_dt_bits, init);
    if (nelem > 0)
        hipLaunchKernelGGL(k_adapt_dt, dim3(std::min(2048, grid_for(nelem, 16))), dim3(256), 0, s, nelem, npoin, connec,
                           helem, u, csound, cfl_conv, has_diff, cfl_diff, mu_fluid, mu_sgs, rho, nullptr,
                           reinterpret_cast<unsigned long long*>(d_dt_bits));
}

void launch_hmin(int nelem, int npoin, const int* connec, const double* helem, double* hmin) {
    hipStream_t s = rt().stream;
    auto* bits = reinterpret_cast<unsigned long long*>(hmin);
    if (npoin > 0)
        hipLaunchKernelGGL(k_fill_u64, dim3(std::min(1024, grid_for(npoin, 256))), dim3(256), 0, s, (long)npoin,
                           bits, kInfBits, nullptr);
    if (nelem > 0)
        hipLaunchKernelGGL(k_hmin_build, dim3(std::min(2048, grid_for((long)nelem * 64, 256))), dim3(256), 0, s,
                           nelem, connec, helem, bits, nullptr);
}

// Cached per-node h_min of the drop-in CFL entry point, keyed on the device
// addresses of connec and helem and checked on the device against a
// fingerprint of their contents on every call (a stale cache is rebuilt by
// the same launches, without a host synchronisation).
struct DtNodeCache {
    const int* connec = nullptr;
    const double* helem = nullptr;
    int nelem = -1, npoin = -1;
    double* hmin = nullptr;
    unsigned long long* fp = nullptr;   // stored hashes (2), then two sums per k_dt_hash block
    int* flags = nullptr;   // [0] cache stale, [1] mu_sgs holds a non-zero value
    void release() {
        (void)hipFree(hmin);
        (void)hipFree(fp);
        (void)hipFree(flags);
        *this = DtNodeCache();
    }
};

static DtNodeCache g_dtc;

void dt_cache_release() { g_dtc.release(); }

// Drops the cached h_min when connec or helem lies in [lo, lo + bytes)
// (always when lo is null), after the work queued on the library stream.
int dt_cache_invalidate(const void* lo, size_t bytes) {
    DtNodeCache& C = g_dtc;
    if (C.hmin == nullptr) return SOD_OK;
    const char* a = static_cast<const char*>(lo);
    const char* c = reinterpret_cast<const char*>(C.connec);
    const char* h = reinterpret_cast<const char*>(C.helem);
    if (lo == nullptr || (c >= a && c < a + bytes) || (h >= a && h < a + bytes)) {
        SOD_HIP(hipStreamSynchronize(rt().stream));
        C.release();
    }
    return SOD_OK;
}

// adapt_dt_cfl through the node form when mu_sgs is absent or zero (decided
// on the device), through the element form otherwise; result bits in
// d_dt_bits.
int launch_adapt_dt_cached(int nelem, int npoin, const int* connec, const double* helem, const double* u,
                           const double* csound, double cfl_conv, int has_diff, double cfl_diff,
                           const double* mu_fluid, const double* mu_sgs, const double* rho, double* d_dt_bits) {
    hipStream_t s = rt().stream;
    DtNodeCache& C = g_dtc;
    int force = 0;
    if (C.connec != connec || C.helem != helem || C.nelem != nelem || C.npoin != npoin) {
        SOD_HIP(hipStreamSynchronize(s));
        C.release();
        SOD_HIP(hipMalloc(&C.hmin, sizeof(double) * (size_t)std::max(1, npoin)));
        SOD_HIP(hipMalloc(&C.fp, sizeof(unsigned long long) * (2 + 2 * kDtHashBlocks)));
        SOD_HIP(hipMemsetAsync(C.fp, 0, sizeof(unsigned long long) * 2, s));
        SOD_HIP(hipMalloc(&C.flags, sizeof(int) * 2));
        C.connec = connec;
        C.helem = helem;
        C.nelem = nelem;
        C.npoin = npoin;
        force = 1;
    }
    auto* out = reinterpret_cast<unsigned long long*>(d_dt_bits);
    auto* hbits = reinterpret_cast<unsigned long long*>(C.hmin);
    hipLaunchKernelGGL(k_set_double, dim3(1), dim3(64), 0, s, d_dt_bits, 100000000000000.0);
    if (nelem <= 0 || npoin <= 0) return SOD_OK;
    const int vec = (reinterpret_cast<uintptr_t>(connec) % 16 == 0 && reinterpret_cast<uintptr_t>(helem) % 16 == 0);
    const int nblk = std::min(kDtHashBlocks, grid_for((long)nelem * 16, 256));
    hipLaunchKernelGGL(k_dt_hash, dim3(nblk), dim3(256), 0, s, (long)nelem * 64, connec, nelem, helem, vec,
                       C.fp + 2);
    hipLaunchKernelGGL(k_dt_cache_check, dim3(1), dim3(256), 0, s, nblk, C.fp + 2, C.fp, C.flags, force);
    hipLaunchKernelGGL(k_fill_u64, dim3(std::min(1024, grid_for(npoin, 256))), dim3(256), 0, s, (long)npoin, hbits,
                       kInfBits, C.flags);
    hipLaunchKernelGGL(k_hmin_build, dim3(std::min(2048, grid_for((long)nelem * 64, 256))), dim3(256), 0, s, nelem,
                       connec, helem, hbits, C.flags);
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, 1, C.flags + 1);
    if (has_diff)
        hipLaunchKernelGGL(k_any_nonzero, dim3(std::min(1024, grid_for((long)nelem * 64, 256))), dim3(256), 0, s,
                           (long)nelem * 64, mu_sgs, C.flags + 1);
    hipLaunchKernelGGL(k_adapt_dt_nodes, dim3(std::min(1024, grid_for(npoin, 256))), dim3(256), 0, s, npoin,
                       C.hmin, u, csound, cfl_conv, has_diff, cfl_diff, mu_fluid, rho, C.flags + 1, out);
    if (has_diff)
        hipLaunchKernelGGL(k_adapt_dt, dim3(std::min(2048, grid_for(nelem, 16))), dim3(256), 0, s, nelem, npoin,
                           connec, helem, u, csound, cfl_conv, has_diff, cfl_diff, mu_fluid, mu_sgs, rho,
                           C.flags + 1, out);
    return SOD_OK;
}

void launch_adapt_dt_nodes_inst(int npoin, const double* hmin, const double* u, const double* csound,
                                double cfl_conv, int has_diff, double cfl_diff, const double* mu_fluid,
                                const double* rho, const int* node_inst, int ninst, double* d_dt_bits) {
    hipStream_t s = rt().stream;
    hipLaunchKernelGGL(k_set_doubles, dim3(1), dim3(256), 0, s, ninst, d_dt_bits, 100000000000000.0);
    if (npoin > 0)
        hipLaunchKernelGGL(k_adapt_dt_nodes_inst, dim3(std::min(2048, grid_for(npoin, 2048))), dim3(256), 0, s,
                           npoin, hmin, u, csound, cfl_conv, has_diff, cfl_diff, mu_fluid, rho, node_inst,
                           reinterpret_cast<unsigned long long*>(d_dt_bits));
}

void launch_adapt_dt_inst(int nelem, int npoin, const int* connec, const double* helem, const double* u,
                          const double* csound, double cfl_conv, int has_diff, double cfl_diff,
                          const double* mu_fluid, const double* mu_sgs, const double* rho, const int* elem_inst,
                          int ninst, double* d_dt_bits) {
    hipStream_t s = rt().stream;
    hipLaunchKernelGGL(k_set_doubles, dim3(1), dim3(256), 0, s, ninst, d_dt_bits, 100000000000000.0);
    if (nelem > 0)
        hipLaunchKernelGGL(k_adapt_dt_inst, dim3(std::min(2048, grid_for(nelem, 16))), dim3(256), 0, s, nelem, npoin,
                           connec, helem, u, csound, cfl_conv, has_diff, cfl_diff, mu_fluid, mu_sgs, rho, elem_inst,
                           reinterpret_cast<unsigned long long*>(d_dt_bits));
}

int red_blocks() { return RED_BLOCKS; }

}  // namespace sod

using namespace sod;

extern "C" {

int sod_sutherland_viscosity(int npoin, const double* Tem, const double* mu_factor,
                             double* mu_fluid) {
    SOD_TRY(ensure_init());
    const double *dT, *dmf;
    double* dmu;
    SOD_TRY(resolve_t(Tem, "Tem", &dT));
    SOD_TRY(resolve_t(mu_factor, "mu_factor", &dmf));
    SOD_TRY(resolve_w(mu_fluid, "mu_fluid", &dmu));
    launch_sutherland(npoin, dT, dmf, dmu);
    return finish_call();
}

static int lumped_entry(int npoin, int npoin_w, const int* lpoin_w, const double* Ml, double* R,
                        int ncomp) {
    SOD_TRY(ensure_init());
    const int* dl;
    const double* dM;
    double* dR;
    SOD_TRY(resolve_t(lpoin_w, "lpoin_w", &dl));
    SOD_TRY(resolve_t(Ml, "Ml", &dM));
    SOD_TRY(resolve_w(R, "R", &dR));
    launch_lumped(npoin, npoin_w, ncomp, dl, dM, dR);
    return finish_call();
}

int sod_lumped_solver_scal(int npoin, int npoin_w, const int* lpoin_w, const double* Ml, double* R) {
    return lumped_entry(npoin, npoin_w, lpoin_w, Ml, R, 1);
}

int sod_lumped_solver_vect(int npoin, int npoin_w, const int* lpoin_w, const double* Ml, double* R) {
    return lumped_entry(npoin, npoin_w, lpoin_w, Ml, R, 3);
}

int sod_smart_visc_spectral(int nelem, int npoin, int npoin_w, const int* connec,
                            const int* lpoin_w, const double* Reta, const double* Ngp,
                            const double* coord, const double* dNgp, const double* gpvol,
                            const double* wgp, double gamma_gas, const double* rho, const double* u,
                            const double* csound, const double* Tem, const double* eta,
                            const double* helem, const double* helem_k, const double* Ml,
                            double* mu_e, const int* invAtoIJK, const int* gmshAtoI,
                            const int* gmshAtoJ, const int* gmshAtoK, double* mue_l) {
    (void)Ngp;
    (void)coord;
    (void)dNgp;
    (void)Ml;
    (void)wgp;
    SOD_TRY(ensure_init());
    if (npoin_w <= 0 && rt().params.flag_normalise_entropy == 1)
        return fail(SOD_ERR_ARG, "smart_visc_spectral: npoin_w must be positive");
    // the element kernel uses the library's GLL weights and node order
    SOD_TRY(check_tables(nullptr, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK));
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
    void* work = nullptr;
    SOD_TRY(scratch(sizeof(double) * (RED_BLOCKS + 8), &work, 0));
    launch_smart_visc(nelem, npoin, npoin_w, dc, dl, dR, dgv, gamma_gas, drho, du, dcs, dT, deta, dhl,
                      dh, dmue, dmuel, static_cast<double*>(work));
    return finish_call();
}

int sod_adapt_dt_cfl(int nelem, int npoin, const int* connec, const double* helem, const double* u,
                     const double* csound, double cfl_conv, double* dt, const double* cfl_diff,
                     const double* mu_fluid, const double* mu_sgs, const double* rho) {
    SOD_TRY(ensure_init());
    if (dt == nullptr) return fail(SOD_ERR_ARG, "adapt_dt_cfl: dt is NULL");
    const int has_diff = cfl_diff && mu_fluid && mu_sgs && rho;
    const int* dc;
    const double *dh, *du, *dcs, *dmu = nullptr, *dsgs = nullptr, *drho = nullptr;
    SOD_TRY(resolve_t(connec, "connec", &dc));
    SOD_TRY(resolve_t(helem, "helem", &dh));
    SOD_TRY(resolve_t(u, "u", &du));
    SOD_TRY(resolve_t(csound, "csound", &dcs));
    if (has_diff) {
        SOD_TRY(resolve_t(mu_fluid, "mu_fluid", &dmu));
        SOD_TRY(resolve_t(mu_sgs, "mu_sgs", &dsgs));
        SOD_TRY(resolve_t(rho, "rho", &drho));
    }
    void* work = nullptr;
    SOD_TRY(scratch(64, &work, 1));
    SOD_TRY(launch_adapt_dt_cached(nelem, npoin, dc, dh, du, dcs, cfl_conv, has_diff, has_diff ? *cfl_diff : 0.0,
                                   dmu, dsgs, drho, static_cast<double*>(work)));
    SOD_HIP(hipGetLastError());
    SOD_HIP(hipStreamSynchronize(rt().stream));
    SOD_HIP(hipMemcpy(dt, work, sizeof(double), hipMemcpyDeviceToHost));
    return SOD_OK;
}

int sod_volAvg_EK(int nelem, int npoin, const int* connec, const double* gpvol, const double* Ngp,
                  double rho0, const double* rho, const double* u, double* EK) {
    SOD_TRY(ensure_init());
    if (EK == nullptr) return fail(SOD_ERR_ARG, "volAvg_EK: EK is NULL");
    const int* dc;
    const double *dgv, *dN, *drho, *du;
    SOD_TRY(resolve_t(connec, "connec", &dc));
    SOD_TRY(resolve_t(gpvol, "gpvol", &dgv));
    SOD_TRY(resolve_t(Ngp, "Ngp", &dN));
    SOD_TRY(resolve_t(rho, "rho", &drho));
    SOD_TRY(resolve_t(u, "u", &du));
    void* work = nullptr;
    SOD_TRY(scratch(sizeof(double) * (RED_BLOCKS + 8), &work, 2));
    double* partial = static_cast<double*>(work);
    hipLaunchKernelGGL(k_ek_partial, dim3(RED_BLOCKS), dim3(256), 0, rt().stream, nelem, npoin, dc, dgv,
                       dN, drho, du, partial);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(BLK), 0, rt().stream, RED_BLOCKS, partial,
                       partial + RED_BLOCKS);
    SOD_HIP(hipGetLastError());
    double s = 0.0;
    SOD_HIP(hipStreamSynchronize(rt().stream));
    SOD_HIP(hipMemcpy(&s, partial + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost));
    // mod_analysis.f90:37 uses 3.14159 for pi
    *EK = s / (rho0 * pow(2.0 * 3.14159, 3.0));
    return SOD_OK;
}

int sod_elem_jacobian(int nelem, int npoin, const int* connec, const double* coord,
                      const double* dNgp, const double* wgp, double* gpvol, double* He) {
    SOD_TRY(ensure_init());
    const int* dc;
    const double *dx, *dN, *dw;
    double *dgv, *dHe;
    SOD_TRY(resolve_t(connec, "connec", &dc));
    SOD_TRY(resolve_t(coord, "coord", &dx));
    SOD_TRY(resolve_t(dNgp, "dNgp", &dN));
    SOD_TRY(resolve_t(wgp, "wgp", &dw));
    SOD_TRY(resolve_w(gpvol, "gpvol", &dgv));
    SOD_TRY(resolve_w(He, "He", &dHe));
    if (nelem > 0)
        hipLaunchKernelGGL(k_elem_jacobian, dim3(grid_for((long)nelem * 64, BLK)), dim3(BLK), 0,
                           rt().stream, nelem, npoin, dc, dx, dN, dw, dgv, dHe);
    return finish_call();
}

int sod_lumped_mass_spectral(int nelem, int npoin, const int* connec, const double* gpvol, double* Ml) {
    SOD_TRY(ensure_init());
    const int* dc;
    const double* dgv;
    double* dM;
    SOD_TRY(resolve_t(connec, "connec", &dc));
    SOD_TRY(resolve_t(gpvol, "gpvol", &dgv));
    SOD_TRY(resolve_w(Ml, "Ml", &dM));
    SOD_HIP(hipMemsetAsync(dM, 0, sizeof(double) * (size_t)npoin, rt().stream));
    if (nelem > 0)
        hipLaunchKernelGGL(k_lumped_mass, dim3(grid_for((long)nelem * 64, BLK)), dim3(BLK), 0,
                           rt().stream, nelem, dc, dgv, dM);
    return finish_call();
}

int sod_char_length(int nelem, int npoin, const int* connec, const double* coord, double* helem) {
    SOD_TRY(ensure_init());
    const int* dc;
    const double* dx;
    double* dh;
    SOD_TRY(resolve_t(connec, "connec", &dc));
    SOD_TRY(resolve_t(coord, "coord", &dx));
    SOD_TRY(resolve_w(helem, "helem", &dh));
    if (nelem > 0)
        hipLaunchKernelGGL(k_char_length, dim3(grid_for(nelem, BLK)), dim3(BLK), 0, rt().stream, nelem,
                           npoin, dc, dx, dh);
    return finish_call();
}

int sod_char_length_spectral(int nelem, int npoin, const int* connec, const double* coord,
                             double* helem_l) {
    SOD_TRY(ensure_init());
    const int* dc;
    const double* dx;
    double* dh;
    SOD_TRY(resolve_t(connec, "connec", &dc));
    SOD_TRY(resolve_t(coord, "coord", &dx));
    SOD_TRY(resolve_w(helem_l, "helem_l", &dh));
    if (nelem > 0)
        hipLaunchKernelGGL(k_char_length_spectral, dim3(grid_for(nelem, 4)), dim3(256), 0, rt().stream,
                           nelem, npoin, dc, dx, dh);
    return finish_call();
}

}  // extern "C"
