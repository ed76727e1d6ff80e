/*
 * sod2d_oracle.c -- CPU restatement of SOD2D's explicit compressible SEM
 * right-hand side and RK4 step (HEX64, p = 3, FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, smoke() and the
 * cpu_baseline leg of bench.py.  The product library never links or calls it.
 *
 * The reference is Fortran whose path modules `use` the MPI and HDF5 Fortran
 * modules, which this image does not provide in a form its Fortran compiler
 * can read, so the reference cannot be compiled here without stand-ins and
 * this file is the oracle.  Each routine restates the cited reference loop
 * nest in the same loop and summation order; it is compiled without FMA
 * contraction (-ffp-contract=off) so that it rounds like a scalar build of
 * the reference.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   orc_gll_roots_weights     closed form: +-1, +-1/sqrt5, 1/6, 5/6 rounded
 *                             through FP32
 *   orc_hex64_tables          bijection/corner checks of the a->ijk maps, the
 *                             closed-form 4x4 derivative matrix, Kronecker
 *                             property of Ngp, row sums of dlxigp_ip = 0,
 *                             dNgp against a finite-difference of Ngp
 *   orc_invert_jacobian       the reference's unit test fixture
 *                             (unitt_3d_jacobian_inverse: det(He) = 0.0625)
 *                             and He*Je = I on random matrices
 *   orc_elem_jacobian         affine elements: He = (2/h) I, gpvol = w (h/2)^3;
 *                             curved elements: sum gpvol = box volume
 *   orc_lumped_mass_spectral  sum Ml = volume; affine closed form per node
 *   orc_char_length(_spectral) closed forms on uniform meshes
 *   orc_full_convec_ijk(_H)   free-stream (uniform state -> 0), conservation
 *                             sum_A R_A = 0 on periodic meshes, linear rho with
 *                             constant u -> R_mass = Ml u.grad(rho), an
 *                             independent dense-operator (Kronecker) evaluation
 *                             of the split form, convergence of
 *                             R/Ml to the analytic flux divergence (for _H:
 *                             to div((E+p)u)), and, with FP64 GLL tables on
 *                             affine periodic meshes, the quadratic identities
 *                             of the Kennedy-Gruber form (discrete kinetic
 *                             energy and sum phi R_ener - phi^2/2 R_mass = 0
 *                             with phi = h + k) and of the _H energy split,
 *                             which the plain divergence form and transposed
 *                             or regrouped terms fail (test_oracle_pins.py)
 *   orc_generic_scalar_convec_ijk  same family of checks
 *   orc_full_diffusion_ijk    constant/linear-u states -> 0, sum_A R_A = 0,
 *                             dense-operator weak form, convergence of -R/Ml
 *                             to the analytic div(tau)
 *   orc_sutherland_viscosity  closed form at T0
 *   orc_lumped_solver_*       closed form
 *   orc_smart_visc_spectral   Reta = 0 -> mu_e = 0; saturated limiter -> the
 *                             filter reproduces cglob*beta_e exactly; a
 *                             brute-force 3x3x3 filter on a lattice field;
 *                             0 <= mu_e <= cglob*beta_e; closed forms of each
 *                             limiter branch on uniform states (ce, the Mach
 *                             limiter tanh(M^15 pi), the Jacobian limiter
 *                             1 - (Jmin/Jmax)^2, saturation at beta_e) and of
 *                             the normalisation over working nodes only
 *   orc_adapt_dt_cfl          closed form on a uniform state
 *   orc_volAvg_EK             closed form for TGV on the exact quadrature
 *   orc_rk_4_main             RK4 amplification polynomial on a linear
 *                             advection mode, conservation of mass/momentum/
 *                             energy, uniform state stays uniform
 * No reference output exists in this image; what the checks above leave
 * open is the order of floating-point additions inside each routine (it
 * follows the reference's loop nests line by line) and the mirror table's end
 * entries of the entropy filter, which the brute-force filter test restates
 * from the table mod_filters.f90:42-51 rather than from a mathematical
 * property.
 */
#include "sod2d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NN 64
#define P1 4

/* Built without OpenMP (liboracle.so, the checker) these are empty and every
 * loop runs in the reference's order.  liboracle_omp.so, built with
 * -fopenmp for the multi-core CPU baseline only, runs the element and node
 * loops on all threads with atomic scatters (the reference's own assembly
 * under OpenACC, elem_convec.f90:442-455), so its sums are reordered. */
#ifdef _OPENMP
#define ORC_PRAGMA(x) _Pragma(#x)
#else
#define ORC_PRAGMA(x)
#endif

/* Fortran-style accessors for the reference layouts (all 0-based here). */
#define HE(i, j, g, e) He[(i) + 3 * (j) + 9 * (g) + 576 * (size_t)(e)]
#define GPVOL(g, e) gpvol[(g) + 64 * (size_t)(e)]
#define DLX(g, d, ii) dlxigp_ip[(g) + 64 * (d) + 192 * (ii)]
#define INV(i, j, k) (invAtoIJK[(i) + 4 * (j) + 16 * (k)] - 1)
#define CONNEC(e, a) (connec[(size_t)(e) + (size_t)nelem * (a)] - 1)

void orc_default_params(orc_params *p)
{
    /* mod_numerical_params.f90:20-28, TGVSolver.f90:104 (Rgas) */
    p->ce = 0.1;
    p->cmax = 0.5;
    p->cglob = 1.0;
    p->c_rho = 1.0;
    p->c_ener = 1.0;
    p->Rgas_inf = 1004.0 * (1.4 - 1.0) / 1.4;
    p->flag_normalise_entropy = 1;
    p->flag_total_enthalpy = 0;
    p->flag_real_diff = 1;   /* mod_numerical_params.f90:93 */
    p->flag_diff_suth = 1;   /* :94 */
}

/* ------------------------------------------------------------------ A1 -- */

/* VALEPO, mod_maths.f90:113-152: Legendre P_N and its first two derivatives
 * by the three-term recurrence. */
static void valepo(int n, double x, double *y, double *dy, double *d2y)
{
    double yp, dyp, d2yp, ym, dym, d2ym, c1, c2, c4;
    *y = 1.0;
    *dy = 0.0;
    *d2y = 0.0;
    if (n == 0) return;
    *y = x;
    *dy = 1.0;
    *d2y = 0.0;
    if (n == 1) return;
    yp = 1.0;
    dyp = 0.0;
    d2yp = 0.0;
    for (int i = 2; i <= n; ++i) {
        c1 = (double)i;
        c2 = 2.0 * c1 - 1.0;
        c4 = c1 - 1.0;
        ym = *y;
        *y = (c2 * x * (*y) - c4 * yp) / c1;
        yp = ym;
        dym = *dy;
        *dy = (c2 * x * (*dy) - c4 * dyp + c2 * yp) / c1;
        dyp = dym;
        d2ym = *d2y;
        *d2y = (c2 * x * (*d2y) - c4 * d2yp + 2.0 * c2 * dyp) / c1;
        d2yp = d2ym;
    }
}

/* ZELEGL / WELEGL (mod_maths.f90:45-111) followed by the FP32 rounding of
 * getGaussLobattoLegendre_weights_and_roots (mod_maths.f90:31-42). */
void orc_gll_roots_weights(double roots[4], double weights[4])
{
    const int n = ORC_PORDER;
    const double v_pi = 2.0 * asin(1.0); /* mod_constants.f90:26 with rp=8 */
    double et[4], vn[4], wt[4];
    int n2 = (n - 1) / 2;
    double sn = (double)(2 * n - 4 * n2 - 3);
    double y, dy, d2y;
    et[0] = -1.0;
    et[n] = 1.0;
    vn[0] = sn;
    vn[n] = 1.0;
    et[n2 + 1] = 0.0;
    valepo(n, 0.0, &y, &dy, &d2y);
    vn[n2 + 1] = y;
    double c = v_pi / (double)n;
    for (int i = 1; i <= n2; ++i) {
        double etx = cos(c * (double)i);
        for (int it = 1; it <= 8; ++it) {
            valepo(n, etx, &y, &dy, &d2y);
            etx = etx - dy / d2y;
        }
        et[i] = -etx;
        et[n - i] = etx;
        vn[i] = y * sn;
        vn[n - i] = y;
    }
    /* WELEGL */
    double dn = (double)n;
    double cw = 2.0 / (dn * (dn + 1.0));
    for (int i = 0; i <= n2; ++i) {
        double yy = vn[i];
        double wtx = cw / (yy * yy);
        wt[i] = wtx;
        wt[n - i] = wtx;
    }
    /* n-1 == 2*n2 for n = 3: no centre weight. */
    for (int i = 0; i < 4; ++i) {
        roots[i] = (double)(float)et[i];  /* real(roots_d): default kind 4 */
        weights[i] = (double)(float)wt[i];
    }
}

/* eval_lagrangePoly, mod_maths.f90:178-214 (lorder maps sod2d 1-D index to
 * the increasing-order grid index). */
static void lagrange_poly(const double xi[4], double xp, double l[4])
{
    int lorder[4] = {1, 4, 2, 3};
    for (int i = 0; i <= ORC_PORDER; ++i) {
        l[i] = 1.0;
        for (int j = 0; j <= ORC_PORDER; ++j) {
            if (j != lorder[i] - 1)
                l[i] = l[i] * ((xp - xi[j]) / (xi[lorder[i] - 1] - xi[j]));
        }
    }
}

/* eval_lagrangePolyDeriv, mod_maths.f90:216-256. */
static void lagrange_deriv(const double xi[4], double xp, double dl[4])
{
    int lorder[4] = {1, 4, 2, 3};
    for (int i = 0; i <= ORC_PORDER; ++i) {
        int L = lorder[i] - 1;
        dl[i] = 0.0;
        for (int j = 0; j <= ORC_PORDER; ++j) {
            double aux = 1.0;
            if (j != L) {
                for (int m = 0; m <= ORC_PORDER; ++m) {
                    if (m != j && m != L) aux = aux * (xp - xi[m]) / (xi[L] - xi[m]);
                }
                dl[i] = dl[i] + (1.0 / (xi[L] - xi[j])) * aux;
            }
        }
    }
}

/* genHighOrderHexGmsh for p = 3 (mod_ijk_indices.f90:574-627), with the
 * edge/face tables of elem_hex.f90:17-18 and genHighOrderQuadGmsh
 * (mod_ijk_indices.f90:629-655). Output: lattice (i,j,k) in 0..3 of each
 * gmsh node, gmsh order. */
static void gmsh_hex_lattice(int tab[64][3])
{
    static const int edges[12][2] = {{1, 2}, {1, 4}, {1, 5}, {2, 3}, {2, 6}, {3, 4},
                                     {3, 7}, {4, 8}, {5, 6}, {5, 8}, {6, 7}, {7, 8}};
    static const int faces[6][4] = {{1, 4, 3, 2}, {1, 2, 6, 5}, {1, 5, 8, 4},
                                    {2, 3, 7, 6}, {3, 4, 8, 7}, {5, 6, 7, 8}};
    const int p = 3;
    static const int corners[8][3] = {{0, 0, 0}, {3, 0, 0}, {3, 3, 0}, {0, 3, 0},
                                      {0, 0, 3}, {3, 0, 3}, {3, 3, 3}, {0, 3, 3}};
    /* genHighOrderQuadGmsh(1) + 1 and genHighOrderHexGmsh(1) + 1 */
    static const int tface[4][2] = {{1, 1}, {2, 1}, {2, 2}, {1, 2}};
    static const int tvol[8][3] = {{1, 1, 1}, {2, 1, 1}, {2, 2, 1}, {1, 2, 1},
                                   {1, 1, 2}, {2, 1, 2}, {2, 2, 2}, {1, 2, 2}};
    int inode = 0;
    for (int c = 0; c < 8; ++c) {
        for (int d = 0; d < 3; ++d) tab[inode][d] = corners[c][d];
        ++inode;
    }
    for (int ie = 0; ie < 12; ++ie) {
        int i0 = edges[ie][0] - 1, i1 = edges[ie][1] - 1;
        int u[3];
        for (int d = 0; d < 3; ++d) u[d] = (tab[i1][d] - tab[i0][d]) / p;
        for (int i = 1; i <= p - 1; ++i) {
            for (int d = 0; d < 3; ++d) tab[inode][d] = tab[i0][d] + i * u[d];
            ++inode;
        }
    }
    for (int f = 0; f < 6; ++f) {
        int i0 = faces[f][0] - 1, i1 = faces[f][1] - 1, i3 = faces[f][3] - 1;
        int u[3], v[3];
        for (int d = 0; d < 3; ++d) {
            u[d] = (tab[i1][d] - tab[i0][d]) / p;
            v[d] = (tab[i3][d] - tab[i0][d]) / p;
        }
        for (int i = 0; i < 4; ++i) {
            for (int d = 0; d < 3; ++d)
                tab[inode][d] = tab[i0][d] + u[d] * tface[i][0] + v[d] * tface[i][1];
            ++inode;
        }
    }
    for (int i = 0; i < 8; ++i) {
        for (int d = 0; d < 3; ++d) tab[inode][d] = tvol[i][d];
        ++inode;
    }
}

void orc_hex64_tables(int atoIJK[64], double xgp[64 * 3], double wgp[64],
                      double Ngp[64 * 64], double dNgp[3 * 64 * 64],
                      double dlxigp_ip[64 * 3 * 4], int invAtoIJK[64],
                      int gmshAtoI[64], int gmshAtoJ[64], int gmshAtoK[64])
{
    int tab[64][3];
    int lat2gmsh[4][4][4];
    gmsh_hex_lattice(tab);
    for (int n = 0; n < 64; ++n) lat2gmsh[tab[n][0]][tab[n][1]][tab[n][2]] = n + 1;
    /* set_allocate_array_ijk_sod2d_criteria (mod_ijk_indices.f90:27-52):
     * sod2d 1-D index -> gmsh lattice index */
    const int s2g[4] = {0, 3, 1, 2};
    /* set_allocate_hexahedronHO_ijk_indices (mod_ijk_indices.f90:61-114):
     * loop kp, ip, jp; a2ijk = gmsh2ijk (mod_meshConversorTool.f90:150-152) */
    int c = 0;
    for (int kp = 0; kp < 4; ++kp)
        for (int ip = 0; ip < 4; ++ip)
            for (int jp = 0; jp < 4; ++jp) atoIJK[c++] = lat2gmsh[s2g[ip]][s2g[jp]][s2g[kp]];
    /* atoIJKInverse, mod_geom.f90:123-144 */
    c = 0;
    for (int k = 0; k < 4; ++k)
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                int a = atoIJK[c++];
                invAtoIJK[i + 4 * j + 16 * k] = a;
                gmshAtoI[a - 1] = i + 1;
                gmshAtoJ[a - 1] = j + 1;
                gmshAtoK[a - 1] = k + 1;
            }
    /* GaussLobattoLegendre_hex, quadrature_rules.f90:51-78 */
    double xi[4], w1d[4];
    orc_gll_roots_weights(xi, w1d);
    const int lorder[4] = {1, 4, 2, 3};
    c = 0;
    for (int k = 0; k < 4; ++k)
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                int a = atoIJK[c++] - 1;
                xgp[a + 64 * 0] = xi[lorder[i] - 1];
                xgp[a + 64 * 1] = xi[lorder[j] - 1];
                xgp[a + 64 * 2] = xi[lorder[k] - 1];
                wgp[a] = w1d[lorder[i] - 1] * w1d[lorder[j] - 1] * w1d[lorder[k] - 1];
            }
    /* hex_highorder + TripleTensorProduct on the GLL grid
     * (elem_hex.f90:22-43, mod_maths.f90:258-298), evaluated at each Gauss
     * point as in CFDSolverBase.f90:1218-1224 */
    for (int g = 0; g < 64; ++g) {
        double s = xgp[g], t = xgp[g + 64], z = xgp[g + 128];
        double lx[4], ly[4], lz[4], dlx[4], dly[4], dlz[4];
        lagrange_poly(xi, s, lx);
        lagrange_poly(xi, t, ly);
        lagrange_poly(xi, z, lz);
        lagrange_deriv(xi, s, dlx);
        lagrange_deriv(xi, t, dly);
        lagrange_deriv(xi, z, dlz);
        for (int ii = 0; ii < 4; ++ii) {
            dlxigp_ip[g + 64 * 0 + 192 * ii] = dlx[ii];
            dlxigp_ip[g + 64 * 1 + 192 * ii] = dly[ii];
            dlxigp_ip[g + 64 * 2 + 192 * ii] = dlz[ii];
        }
        c = 0;
        for (int k = 0; k < 4; ++k)
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) {
                    int a = atoIJK[c++] - 1;
                    Ngp[g + 64 * a] = lx[i] * ly[j] * lz[k];
                    dNgp[0 + 3 * a + 192 * g] = dlx[i] * ly[j] * lz[k];
                    dNgp[1 + 3 * a + 192 * g] = lx[i] * dly[j] * lz[k];
                    dNgp[2 + 3 * a + 192 * g] = lx[i] * ly[j] * dlz[k];
                }
    }
}

/* ------------------------------------------------------------------ A4 -- */

/* Cofactor inverse of jacobian_oper.f90:74-131 (3-D branch). */
double orc_invert_jacobian(const double Je_[9], double He_[9])
{
#define JE(i, j) Je_[((i)-1) + 3 * ((j)-1)]
    double det = (JE(1, 1) * JE(2, 2) * JE(3, 3) + JE(1, 2) * JE(2, 3) * JE(3, 1) +
                  JE(1, 3) * JE(2, 1) * JE(3, 2) - JE(3, 1) * JE(2, 2) * JE(1, 3) -
                  JE(3, 2) * JE(2, 3) * JE(1, 1) - JE(3, 3) * JE(2, 1) * JE(1, 2));
    double a[10], b[10];
    a[1] = JE(2, 2) * JE(3, 3) - JE(3, 2) * JE(2, 3);
    a[2] = JE(2, 1) * JE(3, 3) - JE(3, 1) * JE(2, 3);
    a[3] = JE(2, 1) * JE(3, 2) - JE(3, 1) * JE(2, 2);
    a[4] = JE(1, 2) * JE(3, 3) - JE(3, 2) * JE(1, 3);
    a[5] = JE(1, 1) * JE(3, 3) - JE(3, 1) * JE(1, 3);
    a[6] = JE(1, 1) * JE(3, 2) - JE(3, 1) * JE(1, 2);
    a[7] = JE(1, 2) * JE(2, 3) - JE(2, 2) * JE(1, 3);
    a[8] = JE(1, 1) * JE(2, 3) - JE(2, 1) * JE(1, 3);
    a[9] = JE(1, 1) * JE(2, 2) - JE(2, 1) * JE(1, 2);
    a[2] = -a[2];
    a[4] = -a[4];
    a[6] = -a[6];
    a[8] = -a[8];
    for (int i = 1; i <= 9; ++i) b[i] = a[i];
    b[2] = a[4];
    b[3] = a[7];
    b[4] = a[2];
    b[6] = a[8];
    b[7] = a[3];
    b[8] = a[6];
    for (int i = 1; i <= 9; ++i) b[i] = (1.0 / det) * b[i];
    /* He(1,1)=b1 He(1,2)=b2 He(1,3)=b3 He(2,1)=b4 ... column-major store */
    He_[0 + 3 * 0] = b[1];
    He_[0 + 3 * 1] = b[2];
    He_[0 + 3 * 2] = b[3];
    He_[1 + 3 * 0] = b[4];
    He_[1 + 3 * 1] = b[5];
    He_[1 + 3 * 2] = b[6];
    He_[2 + 3 * 0] = b[7];
    He_[2 + 3 * 1] = b[8];
    He_[2 + 3 * 2] = b[9];
    return det;
#undef JE
}

/* elem_jacobian, jacobian_oper.f90:9-138.  Je(i,j) = sum_a dN(i,a) x_a(j). */
void orc_elem_jacobian(int nelem, int npoin, const int *connec, const double *coord,
                       const double *dNgp, const double *wgp, double *gpvol, double *He)
{
    for (int e = 0; e < nelem; ++e) {
        for (int g = 0; g < 64; ++g) {
            double Je[9], Hl[9];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double s = 0.0;
                    for (int a = 0; a < 64; ++a)
                        s += dNgp[i + 3 * a + 192 * g] * coord[CONNEC(e, a) + (size_t)npoin * j];
                    Je[i + 3 * j] = 0.0 + s;
                }
            double det = orc_invert_jacobian(Je, Hl);
            for (int t = 0; t < 9; ++t) He[t + 9 * g + 576 * (size_t)e] = Hl[t];
            GPVOL(g, e) = wgp[g] * det;
        }
    }
}

/* ------------------------------------------------------------------ A6 -- */

/* lumped_mass_spectral, mass_matrix.f90:11-42 */
void orc_lumped_mass_spectral(int nelem, int npoin, const int *connec, const double *gpvol,
                              double *Ml)
{
    memset(Ml, 0, sizeof(double) * (size_t)npoin);
    for (int e = 0; e < nelem; ++e)
        for (int a = 0; a < 64; ++a) Ml[CONNEC(e, a)] += GPVOL(a, e);
}

/* ------------------------------------------------------------------ A7 -- */

/* char_length + hexa_edges, mod_geom.f90:13-37, elem_hex.f90:45-75 */
void orc_char_length(int nelem, int npoin, const int *connec, const double *coord,
                     double *helem)
{
    static const int ed[12][2] = {{2, 1}, {3, 2}, {4, 3}, {1, 4}, {6, 5}, {7, 6},
                                  {8, 7}, {5, 8}, {5, 1}, {6, 2}, {7, 3}, {8, 4}};
    for (int e = 0; e < nelem; ++e) {
        double dist2 = 1000000000000.0;
        for (int k = 0; k < 12; ++k) {
            int n1 = CONNEC(e, ed[k][0] - 1), n0 = CONNEC(e, ed[k][1] - 1);
            double d[3];
            for (int c = 0; c < 3; ++c)
                d[c] = coord[n1 + (size_t)npoin * c] - coord[n0 + (size_t)npoin * c];
            double aux = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            dist2 = fmin(dist2, aux);
        }
        helem[e] = dist2;
    }
}

/* char_length_spectral, mod_geom.f90:39-66 */
void orc_char_length_spectral(int nelem, int npoin, const int *connec, const double *coord,
                              double *helem_l)
{
    for (int e = 0; e < nelem; ++e)
        for (int a = 0; a < 64; ++a) {
            double aux = 1000000000000.0;
            int na = CONNEC(e, a);
            for (int b = 0; b < 64; ++b) {
                if (a == b) continue;
                int nb = CONNEC(e, b);
                double d[3];
                for (int c = 0; c < 3; ++c)
                    d[c] = coord[na + (size_t)npoin * c] - coord[nb + (size_t)npoin * c];
                double d2 = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                aux = fmin(d2, aux);
            }
            helem_l[(size_t)e + (size_t)nelem * a] = aux;
        }
}

/* ----------------------------------------------------------------- A13 -- */

/* full_convec_ijk, elem_convec.f90:211-460 */
void orc_full_convec_ijk(int nelem, int npoin, const int *connec, const double *He,
                         const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK,
                         const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                         const double *u, const double *q, const double *rho,
                         const double *pr, const double *E, double *Rmass, double *Rmom,
                         double *Rener)
{
    const size_t np = (size_t)npoin;
    memset(Rmom, 0, sizeof(double) * 3 * np);
    memset(Rmass, 0, sizeof(double) * np);
    memset(Rener, 0, sizeof(double) * np);
    int ip[NN];
    double Re_mom[NN][3], Re_mass[NN], Re_ener[NN];
    double ul[NN][3], ql[NN][3], rhol[NN], prl[NN], El[NN], fel[NN][3], fl[NN][3][3],
        fuul[NN][3][3], REl[NN], fuel[NN][3], kl[NN], fkl[NN][3], fukl[NN][3], Rkl[NN];
    ORC_PRAGMA(omp parallel for schedule(static) private(ip, Re_mom, Re_mass, Re_ener, ul, ql, rhol, prl, El,
                                                         fel, fl, fuul, REl, fuel, kl, fkl, fukl, Rkl))
    for (int e = 0; e < nelem; ++e) {
        for (int a = 0; a < NN; ++a) ip[a] = CONNEC(e, a);
        for (int a = 0; a < NN; ++a) {
            kl[a] = 0.0;
            for (int d = 0; d < 3; ++d) kl[a] = kl[a] + u[ip[a] + np * d] * u[ip[a] + np * d] * 0.5;
        }
        for (int d = 0; d < 3; ++d)
            for (int a = 0; a < NN; ++a) {
                double uu = u[ip[a] + np * d];
                ul[a][d] = uu;
                ql[a][d] = q[ip[a] + np * d];
                fel[a][d] = rho[ip[a]] * E[ip[a]] * uu;
                fuel[a][d] = E[ip[a]] * uu;
                fkl[a][d] = rho[ip[a]] * kl[a] * uu;
                fukl[a][d] = kl[a] * uu;
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int a = 0; a < NN; ++a) {
                    fl[a][i][j] = q[ip[a] + np * i] * u[ip[a] + np * j];
                    fuul[a][i][j] = u[ip[a] + np * i] * u[ip[a] + np * j];
                }
        for (int a = 0; a < NN; ++a) {
            rhol[a] = rho[ip[a]];
            El[a] = E[ip[a]];
            REl[a] = rho[ip[a]] * E[ip[a]];
            Rkl[a] = rho[ip[a]] * kl[a];
            prl[a] = pr[ip[a]];
        }
        for (int g = 0; g < NN; ++g) {
            double dlxi[4], dleta[4], dlzeta[4];
            for (int ii = 0; ii < 4; ++ii) {
                dlxi[ii] = DLX(g, 0, ii);
                dleta[ii] = DLX(g, 1, ii);
                dlzeta[ii] = DLX(g, 2, ii);
            }
            int I = gmshAtoI[g] - 1, J = gmshAtoJ[g] - 1, K = gmshAtoK[g] - 1;
            double gRho[3] = {0}, gP[3] = {0}, gE[3] = {0}, gRE[3] = {0}, gk[3] = {0},
                   gRk[3] = {0};
            double gU[3][3] = {{0}}, gF[3][3][3] = {{{0}}}, gFuu[3][3][3] = {{{0}}},
                   gQ[3][3] = {{0}}, gFe[3][3] = {{0}}, gFue[3][3] = {{0}}, gFk[3][3] = {{0}},
                   gFuk[3][3] = {{0}};
            for (int ii = 0; ii < 4; ++ii) {
                int nx = INV(ii, J, K), ny = INV(I, ii, K), nz = INV(I, J, ii);
                gRho[0] += dlxi[ii] * rhol[nx];
                gRho[1] += dleta[ii] * rhol[ny];
                gRho[2] += dlzeta[ii] * rhol[nz];
                gP[0] += dlxi[ii] * prl[nx];
                gP[1] += dleta[ii] * prl[ny];
                gP[2] += dlzeta[ii] * prl[nz];
                gE[0] += dlxi[ii] * El[nx];
                gE[1] += dleta[ii] * El[ny];
                gE[2] += dlzeta[ii] * El[nz];
                gRE[0] += dlxi[ii] * REl[nx];
                gRE[1] += dleta[ii] * REl[ny];
                gRE[2] += dlzeta[ii] * REl[nz];
                gk[0] += dlxi[ii] * kl[nx];
                gk[1] += dleta[ii] * kl[ny];
                gk[2] += dlzeta[ii] * kl[nz];
                gRk[0] += dlxi[ii] * Rkl[nx];
                gRk[1] += dleta[ii] * Rkl[ny];
                gRk[2] += dlzeta[ii] * Rkl[nz];
                for (int id = 0; id < 3; ++id) {
                    gU[id][0] += dlxi[ii] * ul[nx][id];
                    gU[id][1] += dleta[ii] * ul[ny][id];
                    gU[id][2] += dlzeta[ii] * ul[nz][id];
                    gQ[id][0] += dlxi[ii] * ql[nx][id];
                    gQ[id][1] += dleta[ii] * ql[ny][id];
                    gQ[id][2] += dlzeta[ii] * ql[nz][id];
                    gFe[id][0] += dlxi[ii] * fel[nx][id];
                    gFe[id][1] += dleta[ii] * fel[ny][id];
                    gFe[id][2] += dlzeta[ii] * fel[nz][id];
                    gFue[id][0] += dlxi[ii] * fuel[nx][id];
                    gFue[id][1] += dleta[ii] * fuel[ny][id];
                    gFue[id][2] += dlzeta[ii] * fuel[nz][id];
                    gFk[id][0] += dlxi[ii] * fkl[nx][id];
                    gFk[id][1] += dleta[ii] * fkl[ny][id];
                    gFk[id][2] += dlzeta[ii] * fkl[nz][id];
                    gFuk[id][0] += dlxi[ii] * fukl[nx][id];
                    gFuk[id][1] += dleta[ii] * fukl[ny][id];
                    gFuk[id][2] += dlzeta[ii] * fukl[nz][id];
                    for (int jd = 0; jd < 3; ++jd) {
                        gF[id][jd][0] += dlxi[ii] * fl[nx][id][jd];
                        gF[id][jd][1] += dleta[ii] * fl[ny][id][jd];
                        gF[id][jd][2] += dlzeta[ii] * fl[nz][id][jd];
                        gFuu[id][jd][0] += dlxi[ii] * fuul[nx][id][jd];
                        gFuu[id][jd][1] += dleta[ii] * fuul[ny][id][jd];
                        gFuu[id][jd][2] += dlzeta[ii] * fuul[nz][id][jd];
                    }
                }
            }
            double gradRho[3] = {0}, gradP[3] = {0}, gradE[3] = {0}, gradRE[3] = {0},
                   gradk[3] = {0}, gradRk[3] = {0}, gradU[3][3] = {{0}}, gradQ[3][3] = {{0}},
                   divF[3] = {0}, divFuu[3] = {0};
            double divQ = 0.0, divFe = 0.0, divFue = 0.0, divFk = 0.0, divFuk = 0.0, divU;
            for (int id = 0; id < 3; ++id)
                for (int jd = 0; jd < 3; ++jd) {
                    double h = HE(id, jd, g, e);
                    gradRho[id] += h * gRho[jd];
                    gradP[id] += h * gP[jd];
                    gradE[id] += h * gE[jd];
                    gradRE[id] += h * gRE[jd];
                    divFe += h * gFe[id][jd];
                    divFue += h * gFue[id][jd];
                    gradk[id] += h * gk[jd];
                    gradRk[id] += h * gRk[jd];
                    divFk += h * gFk[id][jd];
                    divFuk += h * gFuk[id][jd];
                    for (int kd = 0; kd < 3; ++kd) {
                        double h2 = HE(jd, kd, g, e);
                        gradU[id][jd] += h2 * gU[id][kd];
                        gradQ[id][jd] += h2 * gQ[id][kd];
                        divF[id] += h2 * gF[id][jd][kd];
                        divFuu[id] += h2 * gFuu[id][jd][kd];
                    }
                }
            divU = gradU[0][0] + gradU[1][1] + gradU[2][2];
            divQ = gradQ[0][0] + gradQ[1][1] + gradQ[2][2];
            Re_mass[g] = 0.5 * (divQ + rhol[g] * divU);
            Re_ener[g] = 0.25 * (divFe + El[g] * divQ + REl[g] * divU + rhol[g] * divFue) +
                         0.25 * (divFk + kl[g] * divQ + Rkl[g] * divU + rhol[g] * divFuk);
            for (int id = 0; id < 3; ++id) {
                Re_mom[g][id] = 0.25 * (divF[id] + ul[g][id] * divQ + ql[g][id] * divU +
                                        rhol[g] * divFuu[id]) +
                                gradP[id];
                Re_mass[g] = Re_mass[g] + 0.5 * (ul[g][id] * gradRho[id]);
                Re_ener[g] = Re_ener[g] +
                             0.25 * (ql[g][id] * gradE[id] + ul[g][id] * gradRE[id] +
                                     fuel[g][id] * gradRho[id]) +
                             0.25 * (ql[g][id] * gradk[id] + ul[g][id] * gradRk[id] +
                                     fukl[g][id] * gradRho[id]);
                for (int jd = 0; jd < 3; ++jd)
                    Re_mom[g][id] = Re_mom[g][id] +
                                    0.25 * (fuul[g][id][jd] * gradRho[jd] +
                                            ql[g][jd] * gradU[id][jd] + ul[g][jd] * gradQ[id][jd]);
                Re_mom[g][id] = GPVOL(g, e) * Re_mom[g][id];
            }
            Re_mass[g] = GPVOL(g, e) * Re_mass[g];
            Re_ener[g] = GPVOL(g, e) * Re_ener[g];
        }
        for (int d = 0; d < 3; ++d)
            for (int a = 0; a < NN; ++a) {
                ORC_PRAGMA(omp atomic)
                Rmom[ip[a] + np * d] += Re_mom[a][d];
            }
        for (int a = 0; a < NN; ++a) {
            ORC_PRAGMA(omp atomic)
            Rmass[ip[a]] += Re_mass[a];
            ORC_PRAGMA(omp atomic)
            Rener[ip[a]] += Re_ener[a];
        }
    }
}

/* full_convec_ijk_H, elem_convec.f90:20-208 (total-enthalpy flux (E+p)u) */
void orc_full_convec_ijk_H(int nelem, int npoin, const int *connec, const double *He,
                           const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK,
                           const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                           const double *u, const double *q, const double *rho,
                           const double *pr, const double *E, double *Rmass, double *Rmom,
                           double *Rener)
{
    const size_t np = (size_t)npoin;
    memset(Rmom, 0, sizeof(double) * 3 * np);
    memset(Rmass, 0, sizeof(double) * np);
    memset(Rener, 0, sizeof(double) * np);
    int ip[NN];
    double Re_mom[NN][3], Re_mass[NN], Re_ener[NN];
    double ul[NN][3], ql[NN][3], rhol[NN], prl[NN], El[NN], fel[NN][3], fl[NN][3][3],
        fuul[NN][3][3];
    ORC_PRAGMA(omp parallel for schedule(static) private(ip, Re_mom, Re_mass, Re_ener, ul, ql, rhol, prl, El,
                                                         fel, fl, fuul))
    for (int e = 0; e < nelem; ++e) {
        for (int a = 0; a < NN; ++a) ip[a] = CONNEC(e, a);
        for (int d = 0; d < 3; ++d)
            for (int a = 0; a < NN; ++a) {
                ul[a][d] = u[ip[a] + np * d];
                ql[a][d] = q[ip[a] + np * d];
                fel[a][d] = (E[ip[a]] + pr[ip[a]]) * u[ip[a] + np * d];
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int a = 0; a < NN; ++a) {
                    fl[a][i][j] = q[ip[a] + np * i] * u[ip[a] + np * j];
                    fuul[a][i][j] = u[ip[a] + np * i] * u[ip[a] + np * j];
                }
        for (int a = 0; a < NN; ++a) {
            rhol[a] = rho[ip[a]];
            El[a] = E[ip[a]];
            prl[a] = pr[ip[a]];
        }
        for (int g = 0; g < NN; ++g) {
            double dlxi[4], dleta[4], dlzeta[4];
            for (int ii = 0; ii < 4; ++ii) {
                dlxi[ii] = DLX(g, 0, ii);
                dleta[ii] = DLX(g, 1, ii);
                dlzeta[ii] = DLX(g, 2, ii);
            }
            int I = gmshAtoI[g] - 1, J = gmshAtoJ[g] - 1, K = gmshAtoK[g] - 1;
            double gRho[3] = {0}, gP[3] = {0}, gE[3] = {0};
            double gU[3][3] = {{0}}, gF[3][3][3] = {{{0}}}, gFuu[3][3][3] = {{{0}}},
                   gQ[3][3] = {{0}}, gFe[3][3] = {{0}};
            for (int ii = 0; ii < 4; ++ii) {
                int nx = INV(ii, J, K), ny = INV(I, ii, K), nz = INV(I, J, ii);
                gRho[0] += dlxi[ii] * rhol[nx];
                gRho[1] += dleta[ii] * rhol[ny];
                gRho[2] += dlzeta[ii] * rhol[nz];
                gP[0] += dlxi[ii] * prl[nx];
                gP[1] += dleta[ii] * prl[ny];
                gP[2] += dlzeta[ii] * prl[nz];
                gE[0] += dlxi[ii] * (prl[nx] + El[nx]);
                gE[1] += dleta[ii] * (prl[ny] + El[ny]);
                gE[2] += dlzeta[ii] * (prl[nz] + El[nz]);
                for (int id = 0; id < 3; ++id) {
                    gU[id][0] += dlxi[ii] * ul[nx][id];
                    gU[id][1] += dleta[ii] * ul[ny][id];
                    gU[id][2] += dlzeta[ii] * ul[nz][id];
                    gQ[id][0] += dlxi[ii] * ql[nx][id];
                    gQ[id][1] += dleta[ii] * ql[ny][id];
                    gQ[id][2] += dlzeta[ii] * ql[nz][id];
                    gFe[id][0] += dlxi[ii] * fel[nx][id];
                    gFe[id][1] += dleta[ii] * fel[ny][id];
                    gFe[id][2] += dlzeta[ii] * fel[nz][id];
                    for (int jd = 0; jd < 3; ++jd) {
                        gF[id][jd][0] += dlxi[ii] * fl[nx][id][jd];
                        gF[id][jd][1] += dleta[ii] * fl[ny][id][jd];
                        gF[id][jd][2] += dlzeta[ii] * fl[nz][id][jd];
                        gFuu[id][jd][0] += dlxi[ii] * fuul[nx][id][jd];
                        gFuu[id][jd][1] += dleta[ii] * fuul[ny][id][jd];
                        gFuu[id][jd][2] += dlzeta[ii] * fuul[nz][id][jd];
                    }
                }
            }
            double gradRho[3] = {0}, gradP[3] = {0}, gradE[3] = {0}, gradU[3][3] = {{0}},
                   gradQ[3][3] = {{0}}, divF[3] = {0}, divFuu[3] = {0};
            double divQ, divFe = 0.0, divU;
            for (int id = 0; id < 3; ++id)
                for (int jd = 0; jd < 3; ++jd) {
                    double h = HE(id, jd, g, e);
                    gradRho[id] += h * gRho[jd];
                    gradP[id] += h * gP[jd];
                    gradE[id] += h * gE[jd];
                    divFe += h * gFe[id][jd];
                    for (int kd = 0; kd < 3; ++kd) {
                        double h2 = HE(jd, kd, g, e);
                        gradU[id][jd] += h2 * gU[id][kd];
                        gradQ[id][jd] += h2 * gQ[id][kd];
                        divF[id] += h2 * gF[id][jd][kd];
                        divFuu[id] += h2 * gFuu[id][jd][kd];
                    }
                }
            divU = gradU[0][0] + gradU[1][1] + gradU[2][2];
            divQ = gradQ[0][0] + gradQ[1][1] + gradQ[2][2];
            Re_mass[g] = 0.5 * (divQ + rhol[g] * divU);
            Re_ener[g] = 0.5 * (divFe + (El[g] + prl[g]) * divU);
            for (int id = 0; id < 3; ++id) {
                Re_mom[g][id] = 0.25 * (divF[id] + ul[g][id] * divQ + ql[g][id] * divU +
                                        rhol[g] * divFuu[id]) +
                                gradP[id];
                Re_mass[g] = Re_mass[g] + 0.5 * (ul[g][id] * gradRho[id]);
                Re_ener[g] = Re_ener[g] + 0.5 * (ul[g][id] * gradE[id]);
                for (int jd = 0; jd < 3; ++jd)
                    Re_mom[g][id] = Re_mom[g][id] +
                                    0.25 * (fuul[g][id][jd] * gradRho[jd] +
                                            ql[g][jd] * gradU[id][jd] + ul[g][jd] * gradQ[id][jd]);
                Re_mom[g][id] = GPVOL(g, e) * Re_mom[g][id];
            }
            Re_mass[g] = GPVOL(g, e) * Re_mass[g];
            Re_ener[g] = GPVOL(g, e) * Re_ener[g];
        }
        for (int d = 0; d < 3; ++d)
            for (int a = 0; a < NN; ++a) {
                ORC_PRAGMA(omp atomic)
                Rmom[ip[a] + np * d] += Re_mom[a][d];
            }
        for (int a = 0; a < NN; ++a) {
            ORC_PRAGMA(omp atomic)
            Rmass[ip[a]] += Re_mass[a];
            ORC_PRAGMA(omp atomic)
            Rener[ip[a]] += Re_ener[a];
        }
    }
}

/* ----------------------------------------------------------------- A10 -- */

/* generic_scalar_convec_ijk, elem_convec.f90:462-570 */
void orc_generic_scalar_convec_ijk(int nelem, int npoin, const int *connec, const double *He,
                                   const double *gpvol, const double *dlxigp_ip,
                                   const int *invAtoIJK, const int *gmshAtoI,
                                   const int *gmshAtoJ, const int *gmshAtoK, const double *q,
                                   const double *eta, const double *u, double *Rconvec)
{
    const size_t np = (size_t)npoin;
    memset(Rconvec, 0, sizeof(double) * np);
    int ip[NN];
    double ul[NN][3], fel[NN][3], etal[NN], Re[NN];
    ORC_PRAGMA(omp parallel for schedule(static) private(ip, ul, fel, etal, Re))
    for (int e = 0; e < nelem; ++e) {
        for (int a = 0; a < NN; ++a) ip[a] = CONNEC(e, a);
        for (int d = 0; d < 3; ++d)
            for (int a = 0; a < NN; ++a) {
                ul[a][d] = u[ip[a] + np * d];
                fel[a][d] = q[ip[a] + np * d];
            }
        for (int a = 0; a < NN; ++a) etal[a] = eta[ip[a]];
        for (int g = 0; g < NN; ++g) {
            int I = gmshAtoI[g] - 1, J = gmshAtoJ[g] - 1, K = gmshAtoK[g] - 1;
            double gE[3] = {0}, gU[3][3] = {{0}}, gFe[3][3] = {{0}};
            for (int ii = 0; ii < 4; ++ii) {
                int nx = INV(ii, J, K), ny = INV(I, ii, K), nz = INV(I, J, ii);
                double dx = DLX(g, 0, ii), dy = DLX(g, 1, ii), dz = DLX(g, 2, ii);
                gE[0] += dx * etal[nx];
                gE[1] += dy * etal[ny];
                gE[2] += dz * etal[nz];
                for (int id = 0; id < 3; ++id) {
                    gU[id][0] += dx * ul[nx][id];
                    gU[id][1] += dy * ul[ny][id];
                    gU[id][2] += dz * ul[nz][id];
                    gFe[id][0] += dx * fel[nx][id];
                    gFe[id][1] += dy * fel[ny][id];
                    gFe[id][2] += dz * fel[nz][id];
                }
            }
            double gradE[3] = {0}, divU = 0.0, divFe = 0.0;
            for (int id = 0; id < 3; ++id)
                for (int jd = 0; jd < 3; ++jd) {
                    double h = HE(id, jd, g, e);
                    gradE[id] += h * gE[jd];
                    divU += h * gU[id][jd];
                    divFe += h * gFe[id][jd];
                }
            Re[g] = 0.5 * (divFe + etal[g] * divU);
            for (int id = 0; id < 3; ++id) Re[g] = Re[g] + 0.5 * (ul[g][id] * gradE[id]);
            Re[g] = GPVOL(g, e) * Re[g];
        }
        for (int a = 0; a < NN; ++a) {
            ORC_PRAGMA(omp atomic)
            Rconvec[ip[a]] += Re[a];
        }
    }
}

/* ----------------------------------------------------------------- A12 -- */

/* full_diffusion_ijk, elem_diffu.f90:12-198 */
void orc_full_diffusion_ijk(int nelem, int npoin, const int *connec, const double *He,
                            const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK,
                            const int *gmshAtoI, const int *gmshAtoJ, const int *gmshAtoK,
                            double Cp, double Pr, const double *rho_n, const double *rho,
                            const double *u, const double *Tem, const double *mu_fluid,
                            const double *mu_e, const double *mu_sgs, const orc_params *prm,
                            double *Rmass, double *Rmom, double *Rener)
{
    const size_t np = (size_t)npoin, ne = (size_t)nelem;
    const double twoThirds = 2.0 / 3.0;
    memset(Rmass, 0, sizeof(double) * np);
    memset(Rmom, 0, sizeof(double) * 3 * np);
    memset(Rener, 0, sizeof(double) * np);
    int ip[NN];
    double ul[NN][3], rhol[NN], Teml[NN], mufluidl[NN], rhonl[NN];
    double tauXl[NN][3], tauYl[NN][3], tauZl[NN][3], gradTl[NN][3], gradRhol[NN][3], tauUl[NN][3];
    ORC_PRAGMA(omp parallel for schedule(static) private(ip, ul, rhol, Teml, mufluidl, rhonl, tauXl, tauYl, tauZl,
                                                         gradTl, gradRhol, tauUl))
    for (int e = 0; e < nelem; ++e) {
        for (int a = 0; a < NN; ++a) ip[a] = CONNEC(e, a);
        for (int a = 0; a < NN; ++a) {
            rhol[a] = rho[ip[a]];
            rhonl[a] = rho_n[ip[a]];
            Teml[a] = Tem[ip[a]];
            mufluidl[a] = mu_fluid[ip[a]];
        }
        for (int a = 0; a < NN; ++a)
            for (int d = 0; d < 3; ++d) ul[a][d] = u[ip[a] + np * d];
        for (int g = 0; g < NN; ++g) {
            double mu_fgp = mufluidl[g] + rhonl[g] * mu_sgs[e + ne * g];
            double mu_egp = mu_e[e + ne * g];
            int I = gmshAtoI[g] - 1, J = gmshAtoJ[g] - 1, K = gmshAtoK[g] - 1;
            double gRho[3] = {0}, gT[3] = {0}, gU[3][3] = {{0}};
            for (int ii = 0; ii < 4; ++ii) {
                int nx = INV(ii, J, K), ny = INV(I, ii, K), nz = INV(I, J, ii);
                double dx = DLX(g, 0, ii), dy = DLX(g, 1, ii), dz = DLX(g, 2, ii);
                gRho[0] += dx * rhol[nx];
                gRho[1] += dy * rhol[ny];
                gRho[2] += dz * rhol[nz];
                gT[0] += dx * Teml[nx];
                gT[1] += dy * Teml[ny];
                gT[2] += dz * Teml[nz];
                for (int id = 0; id < 3; ++id) {
                    gU[id][0] += dx * ul[nx][id];
                    gU[id][1] += dy * ul[ny][id];
                    gU[id][2] += dz * ul[nz][id];
                }
            }
            double gradRho[3] = {0}, gradT[3] = {0}, gradU[3][3] = {{0}};
            for (int id = 0; id < 3; ++id)
                for (int jd = 0; jd < 3; ++jd) {
                    double h = HE(id, jd, g, e);
                    gradRho[id] += h * gRho[jd];
                    gradT[id] += h * gT[jd];
                    for (int kd = 0; kd < 3; ++kd) gradU[id][jd] += HE(jd, kd, g, e) * gU[id][kd];
                }
            double divU = gradU[0][0] + gradU[1][1] + gradU[2][2];
            double tauU[3] = {0}, tau[3][3];
            for (int id = 0; id < 3; ++id) {
                for (int jd = 0; jd < 3; ++jd) {
                    tauU[id] = tauU[id] + (mu_fgp + mu_egp) * (gradU[id][jd] + gradU[jd][id]) * ul[g][jd];
                    tau[id][jd] = (mu_fgp + mu_egp) * (gradU[id][jd] + gradU[jd][id]);
                }
                tauU[id] = tauU[id] - (mu_fgp)*twoThirds * divU * ul[g][id];
                tau[id][id] = tau[id][id] - (mu_fgp)*twoThirds * divU;
            }
            for (int d = 0; d < 3; ++d) {
                tauXl[g][d] = tau[0][d];
                tauYl[g][d] = tau[1][d];
                tauZl[g][d] = tau[2][d];
                gradTl[g][d] = gradT[d];
                gradRhol[g][d] = gradRho[d];
                tauUl[g][d] = tauU[d];
            }
        }
        for (int g = 0; g < NN; ++g) {
            double nu_e = prm->c_rho * mu_e[e + ne * g] / rhonl[g];
            double kappa_e = mufluidl[g] * Cp / Pr + prm->c_ener * mu_e[e + ne * g] * Cp / Pr +
                             Cp * rhonl[g] * mu_sgs[e + ne * g] / 0.9;
            int I = gmshAtoI[g] - 1, J = gmshAtoJ[g] - 1, K = gmshAtoK[g] - 1;
            double divDe = 0.0, divDr = 0.0, divDm[3] = {0};
            for (int ii = 0; ii < 4; ++ii)
                for (int id = 0; id < 3; ++id) {
                    int qx = INV(ii, J, K), qy = INV(I, ii, K), qz = INV(I, J, ii);
                    double wx = HE(id, 0, qx, e) * GPVOL(qx, e) * DLX(qx, 0, I);
                    double wy = HE(id, 1, qy, e) * GPVOL(qy, e) * DLX(qy, 1, J);
                    double wz = HE(id, 2, qz, e) * GPVOL(qz, e) * DLX(qz, 2, K);
                    divDr = divDr + wx * gradRhol[qx][id];
                    divDr = divDr + wy * gradRhol[qy][id];
                    divDr = divDr + wz * gradRhol[qz][id];
                    divDe = divDe + wx * (tauUl[qx][id] + kappa_e * gradTl[qx][id]);
                    divDe = divDe + wy * (tauUl[qy][id] + kappa_e * gradTl[qy][id]);
                    divDe = divDe + wz * (tauUl[qz][id] + kappa_e * gradTl[qz][id]);
                    divDm[0] = divDm[0] + wx * tauXl[qx][id];
                    divDm[0] = divDm[0] + wy * tauXl[qy][id];
                    divDm[0] = divDm[0] + wz * tauXl[qz][id];
                    divDm[1] = divDm[1] + wx * tauYl[qx][id];
                    divDm[1] = divDm[1] + wy * tauYl[qy][id];
                    divDm[1] = divDm[1] + wz * tauYl[qz][id];
                    divDm[2] = divDm[2] + wx * tauZl[qx][id];
                    divDm[2] = divDm[2] + wy * tauZl[qy][id];
                    divDm[2] = divDm[2] + wz * tauZl[qz][id];
                }
            ORC_PRAGMA(omp atomic)
            Rmass[ip[g]] += nu_e * divDr;
            ORC_PRAGMA(omp atomic)
            Rener[ip[g]] += divDe;
            for (int d = 0; d < 3; ++d) {
                ORC_PRAGMA(omp atomic)
                Rmom[ip[g] + np * d] += divDm[d];
            }
        }
    }
}

/* ----------------------------------------------------------- A11, A14 -- */

/* sutherland_viscosity, mod_fluid_viscosity.f90:21-41 */
void orc_sutherland_viscosity(int npoin, const double *Tem, const double *mu_factor,
                              double *mu_fluid)
{
    const double C1 = 0.000001458, S = 110.4;
    ORC_PRAGMA(omp parallel for schedule(static))
    for (int n = 0; n < npoin; ++n)
        mu_fluid[n] = mu_factor[n] * C1 * pow(Tem[n], 1.5) / (Tem[n] + S);
}

/* lumped_solver_scal / _vect, mod_solver.f90:23-58 */
void orc_lumped_solver_scal(int npoin, int npoin_w, const int *lpoin_w, const double *Ml,
                            double *R)
{
    (void)npoin;
    ORC_PRAGMA(omp parallel for schedule(static))
    for (int i = 0; i < npoin_w; ++i) {
        int n = lpoin_w[i] - 1;
        R[n] = R[n] / Ml[n];
    }
}

void orc_lumped_solver_vect(int npoin, int npoin_w, const int *lpoin_w, const double *Ml,
                            double *R)
{
    ORC_PRAGMA(omp parallel for schedule(static))
    for (int i = 0; i < npoin_w; ++i) {
        int n = lpoin_w[i] - 1;
        for (int d = 0; d < 3; ++d) R[n + (size_t)npoin * d] = R[n + (size_t)npoin * d] / Ml[n];
    }
}

/* ----------------------------------------------------------------- A17 -- */

static double fmax_f(double a, double b) { return a > b ? a : b; }
static double fmin_f(double a, double b) { return a < b ? a : b; }

/* smart_visc_spectral, mod_entropy_viscosity.f90:14-123, with the filter
 * tables of init_filters, mod_filters.f90:15-55.  The dummy named helem(nelem,
 * nnode) receives helem_l and helem_k receives helem (time_integ.f90:492). */
void orc_smart_visc_spectral(int nelem, int npoin, int npoin_w, const int *connec,
                             const int *lpoin_w, const double *Reta, const double *gpvol,
                             const double *wgp, double gamma_gas, const double *rho,
                             const double *u, const double *csound, const double *Tem,
                             const double *eta, const double *helem_l, const double *helem,
                             const int *invAtoIJK, const orc_params *prm, double *mu_e,
                             double *mue_l)
{
    const size_t np = (size_t)npoin, ne = (size_t)nelem;
    const double v_pi = 2.0 * asin(1.0);
    const double Rgas = prm->Rgas_inf;
    /* convertIJK(0:p+2) for p = 3 (mod_filters.f90:42-51) */
    const int convertIJK[6] = {3, 1, 3, 4, 2, 3};
    const double wts[3] = {1.0 / 4.0, 2.0 / 4.0, 1.0 / 4.0};
    double norm;
    memcpy(mue_l, mu_e, sizeof(double) * ne * 64);
    if (prm->flag_normalise_entropy == 1) {
        double maxEta_r = 0.0;
        ORC_PRAGMA(omp parallel for schedule(static) reduction(+ : maxEta_r))
        for (int i = 0; i < npoin_w; ++i) maxEta_r = maxEta_r + eta[lpoin_w[i] - 1];
        double maxEta = maxEta_r / (double)npoin_w;
        double norm_r = 0.0;
        ORC_PRAGMA(omp parallel for schedule(static) reduction(max : norm_r))
        for (int i = 0; i < npoin_w; ++i)
            norm_r = fmax_f(norm_r, fabs(eta[lpoin_w[i] - 1] - maxEta));
        norm = norm_r;
    } else {
        norm = 1.0;
    }
    ORC_PRAGMA(omp parallel for schedule(static))
    for (int e = 0; e < nelem; ++e) {
        double maxJe = 0.0, minJe = 1000000.0, maxV = 0.0, maxC = 0.0;
        for (int g = 0; g < 64; ++g) {
            int n = CONNEC(e, g);
            minJe = fmin_f(minJe, GPVOL(g, e) / wgp[g]);
            maxJe = fmax_f(maxJe, GPVOL(g, e) / wgp[g]);
            double u0 = u[n], u1 = u[n + np], u2 = u[n + 2 * np];
            maxV = fmax_f(maxV, sqrt(u0 * u0 + u1 * u1 + u2 * u2));
            maxC = fmax_f(maxC, csound[n]);
        }
        double M = maxV / maxC;
        double M15 = pow(M, 15.0);
        double ceM = fmax_f(tanh(M15 * v_pi), prm->ce);
        double r = minJe / maxJe;
        double ced = fmax_f(1.0 - r * r, prm->ce);
        ced = fmax_f(ced, ceM);
        double betae = 0.0;
        for (int a = 0; a < 64; ++a) {
            int n = CONNEC(e, a);
            double u0 = u[n], u1 = u[n + np], u2 = u[n + 2 * np];
            double aux2 = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
            double aux3 = sqrt(Rgas * gamma_gas * Tem[n]);
            double aux1 = aux2 + aux3;
            betae = fmax_f(betae, (rho[n] * helem[e]) * (prm->cmax / (double)ORC_PORDER) * aux1);
        }
        for (int a = 0; a < 64; ++a) {
            int n = CONNEC(e, a);
            double R1 = rho[n] * fabs(Reta[n]) / norm;
            double h = helem_l[e + ne * a];
            double Ve = ced * R1 * (h * h);
            mue_l[e + ne * a] = prm->cglob * fmin_f(Ve, betae);
        }
        for (int ii = 1; ii <= 4; ++ii)
            for (int jj = 1; jj <= 4; ++jj)
                for (int kk = 1; kk <= 4; ++kk) {
                    double aux1 = 0.0;
                    for (int ll = -1; ll <= 1; ++ll)
                        for (int mm = -1; mm <= 1; ++mm)
                            for (int nn = -1; nn <= 1; ++nn) {
                                int a = INV(convertIJK[ii + ll] - 1, convertIJK[jj + mm] - 1,
                                            convertIJK[kk + nn] - 1);
                                aux1 = aux1 + wts[ll + 1] * wts[mm + 1] * wts[nn + 1] *
                                                  mue_l[e + ne * a];
                            }
                    int a = INV(convertIJK[ii] - 1, convertIJK[jj] - 1, convertIJK[kk] - 1);
                    mu_e[e + ne * a] = aux1;
                }
    }
}

/* ----------------------------------------------------------------- A18 -- */

/* adapt_dt_cfl, mod_time_ops.f90:11-61 (both CFL terms present) */
double orc_adapt_dt_cfl(int nelem, int npoin, const int *connec, const double *helem,
                        const double *u, const double *csound, double cfl_conv,
                        double cfl_diff, const double *mu_fluid, const double *mu_sgs,
                        const double *rho)
{
    const size_t np = (size_t)npoin, ne = (size_t)nelem;
    double dt_conv = 100000000000000.0, dt_diff = 100000000000000.0, dt_l = 100000000000000.0;
    ORC_PRAGMA(omp parallel for schedule(static) reduction(min : dt_conv, dt_diff))
    for (int e = 0; e < nelem; ++e) {
        double L3 = 0.0;
        for (int a = 0; a < 64; ++a) {
            int n = CONNEC(e, a);
            double umag = fabs(u[n]);
            umag = fmax_f(umag, fabs(u[n + np]));
            umag = fmax_f(umag, fabs(u[n + 2 * np]));
            double aux1 = umag + csound[n];
            L3 = fmax_f(L3, aux1);
        }
        double aux2 = cfl_conv * (helem[e] / (double)(2.0 * ORC_PORDER + 1)) / L3;
        dt_conv = fmin_f(dt_conv, aux2);
        double max_MU = 0.0;
        for (int a = 0; a < 64; ++a) {
            int n = CONNEC(e, a);
            max_MU = fmax_f(max_MU, (rho[n] * mu_sgs[e + ne * a] + mu_fluid[n]) / rho[n]);
        }
        double hh = helem[e] / (double)(2.0 * ORC_PORDER + 1);
        double aux4 = cfl_diff * (hh * hh) / max_MU;
        dt_diff = fmin_f(dt_diff, aux4);
    }
    /* the running minimum of the loop (dt_l = min(dt_conv, dt_diff) after
     * every element, mod_time_ops.f90:54) ends at this value */
    dt_l = fmin_f(dt_l, fmin_f(dt_conv, dt_diff));
    return dt_l;
}

/* ----------------------------------------------------------------- A21 -- */

/* volAvg_EK, mod_analysis.f90:10-44 */
double orc_volAvg_EK(int nelem, int npoin, const int *connec, const double *gpvol,
                     const double *Ngp, double rho0, const double *rho, const double *u)
{
    const size_t np = (size_t)npoin;
    double EK_l = 0.0;
    for (int e = 0; e < nelem; ++e) {
        double R1 = 0.0;
        for (int g = 0; g < 64; ++g)
            for (int a = 0; a < 64; ++a) {
                int n = CONNEC(e, a);
                double u0 = u[n], u1 = u[n + np], u2 = u[n + 2 * np];
                R1 = R1 + GPVOL(g, e) * Ngp[g + 64 * a] * 0.5 * rho[n] * (u0 * u0 + u1 * u1 + u2 * u2);
            }
        EK_l = EK_l + R1;
    }
    EK_l = EK_l / (rho0 * pow(2.0 * 3.14159, 3.0));
    return EK_l;
}

/* ------------------------------------------------------------- A8-A17 -- */

/* rk_4_main, time_integ.f90:137-508, periodic branch. */
void orc_rk_4_main(int nelem, int npoin, int npoin_w, const int *connec, const int *lpoin_w,
                   const double *He, const double *gpvol, const double *dlxigp_ip,
                   const double *wgp, const int *invAtoIJK, const int *gmshAtoI,
                   const int *gmshAtoJ, const int *gmshAtoK, const double *Ml,
                   const double *helem, const double *helem_l, double dt, double Rgas,
                   double gamma_gas, double Cp, double Prt, const double *mu_factor,
                   const orc_params *prm, double *rho, double *u, double *q, double *pr,
                   double *E, double *Tem, double *e_int, double *eta, double *csound,
                   double *machno, double *mu_fluid, double *mu_e, double *mu_sgs,
                   double *mue_l)
{
    const size_t np = (size_t)npoin;
    /* init_rk4_solver, time_integ.f90:81-84 */
    const double a_i[4] = {0.0, 0.5, 0.5, 1.0};
    const double b_i[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    double *buf = (double *)calloc(np * 33, sizeof(double));
    double *Rmass = buf, *Rener = buf + np, *Reta = buf + 2 * np, *Rmom = buf + 3 * np;
    double *aux_rho = buf + 6 * np, *aux_pr = buf + 7 * np, *aux_E = buf + 8 * np,
           *aux_Tem = buf + 9 * np, *aux_e_int = buf + 10 * np, *aux_eta = buf + 11 * np,
           *aux_h = buf + 12 * np, *aux_u = buf + 13 * np, *aux_q = buf + 16 * np;
    double *Rmass_sum = buf + 19 * np, *Rener_sum = buf + 20 * np, *Reta_sum = buf + 21 * np,
           *Rdiff_mass = buf + 22 * np, *Rdiff_ener = buf + 23 * np, *Rmom_sum = buf + 24 * np,
           *Rdiff_mom = buf + 27 * np, *f_eta = buf + 30 * np;

    for (int istep = 0; istep < 4; ++istep) {
        /* time_integ.f90:239-248 */
        ORC_PRAGMA(omp parallel for schedule(static))
        for (size_t n = 0; n < np; ++n) {
            aux_rho[n] = rho[n] - dt * a_i[istep] * Rmass[n];
            aux_E[n] = E[n] - dt * a_i[istep] * Rener[n];
            for (int d = 0; d < 3; ++d)
                aux_q[n + np * d] = q[n + np * d] - dt * a_i[istep] * Rmom[n + np * d];
        }
        /* time_integ.f90:266-285 */
        ORC_PRAGMA(omp parallel for schedule(static))
        for (int i = 0; i < npoin_w; ++i) {
            size_t n = (size_t)(lpoin_w[i] - 1);
            double umag = 0.0;
            for (int d = 0; d < 3; ++d) {
                aux_u[n + np * d] = aux_q[n + np * d] / aux_rho[n];
                umag = umag + (aux_u[n + np * d] * aux_u[n + np * d]);
            }
            aux_e_int[n] = (aux_E[n] / aux_rho[n]) - 0.5 * umag;
            aux_pr[n] = aux_rho[n] * (gamma_gas - 1.0) * aux_e_int[n];
            aux_h[n] = (gamma_gas / (gamma_gas - 1.0)) * aux_pr[n] / aux_rho[n];
            aux_Tem[n] = aux_pr[n] / (aux_rho[n] * Rgas);
            aux_eta[n] = (aux_rho[n] / (gamma_gas - 1.0)) * log(aux_pr[n] / pow(aux_rho[n], gamma_gas));
            for (int d = 0; d < 3; ++d) f_eta[n + np * d] = aux_u[n + np * d] * aux_eta[n];
        }
        /* time_integ.f90:288-304 */
        orc_generic_scalar_convec_ijk(nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                                      gmshAtoI, gmshAtoJ, gmshAtoK, f_eta, aux_eta, aux_u, Reta);
        orc_lumped_solver_scal(npoin, npoin_w, lpoin_w, Ml, Reta);
        /* time_integ.f90:311-316: with the gate off mu_fluid is the caller's,
         * unchanged through the step */
        if (prm->flag_real_diff == 1 && prm->flag_diff_suth == 1)
            orc_sutherland_viscosity(npoin, aux_Tem, mu_factor, mu_fluid);
        /* time_integ.f90:321 */
        orc_full_diffusion_ijk(nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK, gmshAtoI,
                               gmshAtoJ, gmshAtoK, Cp, Prt, aux_rho, aux_rho, aux_u, aux_Tem,
                               mu_fluid, mu_e, mu_sgs, prm, Rdiff_mass, Rdiff_mom, Rdiff_ener);
        /* time_integ.f90:360-365 */
        if (prm->flag_total_enthalpy)
            orc_full_convec_ijk_H(nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK,
                                  gmshAtoI, gmshAtoJ, gmshAtoK, aux_u, aux_q, aux_rho, aux_pr,
                                  aux_E, Rmass, Rmom, Rener);
        else
            orc_full_convec_ijk(nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK, gmshAtoI,
                                gmshAtoJ, gmshAtoK, aux_u, aux_q, aux_rho, aux_pr, aux_h, Rmass,
                                Rmom, Rener);
        /* time_integ.f90:369-373 */
        ORC_PRAGMA(omp parallel for schedule(static))
        for (size_t n = 0; n < np; ++n) {
            Rmass[n] = Rmass[n] + Rdiff_mass[n];
            Rener[n] = Rener[n] + Rdiff_ener[n];
        }
        ORC_PRAGMA(omp parallel for schedule(static))
        for (size_t n = 0; n < 3 * np; ++n) Rmom[n] = Rmom[n] + Rdiff_mom[n];
        /* time_integ.f90:391-393 */
        orc_lumped_solver_scal(npoin, npoin_w, lpoin_w, Ml, Rmass);
        orc_lumped_solver_scal(npoin, npoin_w, lpoin_w, Ml, Rener);
        orc_lumped_solver_vect(npoin, npoin_w, lpoin_w, Ml, Rmom);
        /* time_integ.f90:399-409 */
        ORC_PRAGMA(omp parallel for schedule(static))
        for (size_t n = 0; n < np; ++n) {
            Rmass_sum[n] = Rmass_sum[n] + b_i[istep] * Rmass[n];
            Rener_sum[n] = Rener_sum[n] + b_i[istep] * Rener[n];
            Reta_sum[n] = Reta_sum[n] + b_i[istep] * Reta[n];
            for (int d = 0; d < 3; ++d)
                Rmom_sum[n + np * d] = Rmom_sum[n + np * d] + b_i[istep] * Rmom[n + np * d];
        }
    }
    /* time_integ.f90:417-426 */
    ORC_PRAGMA(omp parallel for schedule(static))
    for (size_t n = 0; n < np; ++n) {
        rho[n] = rho[n] - dt * Rmass_sum[n];
        E[n] = E[n] - dt * Rener_sum[n];
        for (int d = 0; d < 3; ++d) q[n + np * d] = q[n + np * d] - dt * Rmom_sum[n + np * d];
    }
    /* time_integ.f90:445-463 */
    ORC_PRAGMA(omp parallel for schedule(static))
    for (int i = 0; i < npoin_w; ++i) {
        size_t n = (size_t)(lpoin_w[i] - 1);
        double umag = 0.0;
        for (int d = 0; d < 3; ++d) {
            u[n + np * d] = q[n + np * d] / rho[n];
            umag = umag + u[n + np * d] * u[n + np * d];
        }
        e_int[n] = (E[n] / rho[n]) - 0.5 * umag;
        pr[n] = rho[n] * (gamma_gas - 1.0) * e_int[n];
        csound[n] = sqrt(gamma_gas * pr[n] / rho[n]);
        umag = sqrt(umag);
        machno[n] = umag / csound[n];
        Tem[n] = pr[n] / (rho[n] * Rgas);
        eta[n] = (rho[n] / (gamma_gas - 1.0)) * log(pr[n] / pow(rho[n], gamma_gas));
    }
    /* time_integ.f90:467-471 */
    ORC_PRAGMA(omp parallel for schedule(static))
    for (int i = 0; i < npoin_w; ++i) {
        size_t n = (size_t)(lpoin_w[i] - 1);
        Reta[n] = -Reta_sum[n];
    }
    /* time_integ.f90:482-493 */
    if (prm->flag_real_diff == 1 && prm->flag_diff_suth == 1)
        orc_sutherland_viscosity(npoin, Tem, mu_factor, mu_fluid);
    orc_smart_visc_spectral(nelem, npoin, npoin_w, connec, lpoin_w, Reta, gpvol, wgp, gamma_gas,
                            rho, u, csound, Tem, eta, helem_l, helem, invAtoIJK, prm, mu_e,
                            mue_l);
    free(buf);
}
