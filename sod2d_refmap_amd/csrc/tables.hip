// Init-time p=3 tables: GLL nodes/weights, Lagrange derivatives, the gmsh
// high-order node order and the a<->ijk maps.  Host code; the derived
// constants the kernels need are uploaded to __constant__ memory at init.
//
// Reference behaviour reproduced here:
//   * GLL roots and weights are computed in FP64 and then rounded through
//     FP32 (`real(x)` without KIND, mod_maths.f90:39-40), so every downstream
//     table carries FP32-rounded node positions;
//   * 1-D "sod2d" order is (xi=-1, xi=+1, interior increasing)
//     (mod_maths.f90:196-200, mod_ijk_indices.f90:27-52);
//   * element-local node order is gmsh's high-order hexahedron order
//     (a2ijk = gmsh2ijk, mod_meshConversorTool.f90:150-152).
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "sod2d_internal.hpp"

namespace sod {

namespace {

// Legendre P_n(x) and P_n'(x), P_n''(x) by the three-term recurrence.
void legendre(int n, double x, double& p, double& dp, double& d2p) {
    double p0 = 1.0, p1 = x, dp0 = 0.0, dp1 = 1.0, d2p0 = 0.0, d2p1 = 0.0;
    if (n == 0) { p = 1.0; dp = 0.0; d2p = 0.0; return; }
    for (int k = 2; k <= n; ++k) {
        double a = (2.0 * k - 1.0) / k, b = (k - 1.0) / k;
        double p2 = a * x * p1 - b * p0;
        double dp2 = a * (x * dp1 + p1) - b * dp0;
        double d2p2 = a * (x * d2p1 + 2.0 * dp1) - b * d2p0;
        p0 = p1; p1 = p2; dp0 = dp1; dp1 = dp2; d2p0 = d2p1; d2p1 = d2p2;
    }
    p = p1; dp = dp1; d2p = d2p1;
}

// Gmsh high-order hexahedron of order p=3: lattice coordinates of each node
// in gmsh order (corners, 12 edges, 6 faces, interior), 0..3 per direction.
void gmsh_hex3(int lat[64][3]) {
    static const int corner[8][3] = {{0, 0, 0}, {3, 0, 0}, {3, 3, 0}, {0, 3, 0},
                                     {0, 0, 3}, {3, 0, 3}, {3, 3, 3}, {0, 3, 3}};
    // gmsh edge list (corner pairs) and face list (corner quadruples, first,
    // second and last corner span the face's local axes).
    static const int edge[12][2] = {{0, 1}, {0, 3}, {0, 4}, {1, 2}, {1, 5}, {2, 3},
                                    {2, 6}, {3, 7}, {4, 5}, {4, 7}, {5, 6}, {6, 7}};
    static const int face[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {0, 4, 7, 3},
                                   {1, 2, 6, 5}, {2, 3, 7, 6}, {4, 5, 6, 7}};
    // order-1 quad / hex corner patterns used for the face and volume
    // interiors (the recursion bottoms out at order 1 for p = 3)
    static const int quad1[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    int n = 0;
    for (int c = 0; c < 8; ++c, ++n)
        for (int d = 0; d < 3; ++d) lat[n][d] = corner[c][d];
    for (int e = 0; e < 12; ++e) {
        const int* a = corner[edge[e][0]];
        const int* b = corner[edge[e][1]];
        for (int s = 1; s <= 2; ++s, ++n)
            for (int d = 0; d < 3; ++d) lat[n][d] = a[d] + s * (b[d] - a[d]) / 3;
    }
    for (int f = 0; f < 6; ++f) {
        const int* o = corner[face[f][0]];
        const int* pu = corner[face[f][1]];
        const int* pv = corner[face[f][3]];
        for (int s = 0; s < 4; ++s, ++n)
            for (int d = 0; d < 3; ++d)
                lat[n][d] = o[d] + (quad1[s][0] + 1) * (pu[d] - o[d]) / 3 +
                            (quad1[s][1] + 1) * (pv[d] - o[d]) / 3;
    }
    for (int c = 0; c < 8; ++c, ++n)
        for (int d = 0; d < 3; ++d) lat[n][d] = 1 + corner[c][d] / 3;
}

HostTables g_tab;
std::once_flag g_tab_once;

void build_tables() {
    HostTables& t = g_tab;
    // --- GLL nodes/weights of order 3 (interior nodes: roots of P3').
    const int N = kP;
    double x[4], w[4];
    x[0] = -1.0; x[3] = 1.0;
    double r = std::cos(3.14159265358979323846 / N);   // Chebyshev-Gauss-Lobatto initial guess
    for (int it = 0; it < 12; ++it) {
        double p, dp, d2p;
        legendre(N, r, p, dp, d2p);
        r -= dp / d2p;
    }
    x[1] = -r; x[2] = r;
    for (int i = 0; i < 4; ++i) {
        double p, dp, d2p;
        legendre(N, x[i], p, dp, d2p);
        w[i] = 2.0 / (N * (N + 1.0) * p * p);
    }
    for (int i = 0; i < 4; ++i) {
        t.roots[i] = (double)(float)x[i];
        t.weights[i] = (double)(float)w[i];
    }
    // sod2d 1-D order -> increasing grid index
    const int ord[4] = {0, 3, 1, 2};
    // --- derivative matrix D[m][i] = l'_i(xi_m) (i, m in sod2d order), by the
    // product rule on the FP32-rounded grid (mod_maths.f90:243-256).
    for (int m = 0; m < 4; ++m)
        for (int i = 0; i < 4; ++i) {
            const double xp = t.roots[ord[m]];
            const int L = ord[i];
            double s = 0.0;
            for (int j = 0; j < 4; ++j) {
                if (j == L) continue;
                double prod = 1.0;
                for (int k = 0; k < 4; ++k)
                    if (k != j && k != L) prod = prod * (xp - t.roots[k]) / (t.roots[L] - t.roots[k]);
                s = s + (1.0 / (t.roots[L] - t.roots[j])) * prod;
            }
            t.D[m][i] = s;
        }
    // Lagrange values on the grid are exactly the identity (collocation).
    // --- gmsh order and the a<->ijk maps.
    int lat[64][3];
    gmsh_hex3(lat);
    int g_of_lat[4][4][4];
    for (int a = 0; a < 64; ++a) g_of_lat[lat[a][0]][lat[a][1]][lat[a][2]] = a;
    const int s2lat[4] = {0, 3, 1, 2};   // sod2d index -> lattice coordinate
    int c = 0;
    for (int k = 0; k < 4; ++k)
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                int a = g_of_lat[s2lat[i]][s2lat[j]][s2lat[k]];
                t.atoIJK[c++] = a + 1;
                t.invAtoIJK[i + 4 * j + 16 * k] = a + 1;
                t.gmshAtoI[a] = i + 1;
                t.gmshAtoJ[a] = j + 1;
                t.gmshAtoK[a] = k + 1;
                t.lane2node[i + 4 * j + 16 * k] = a;
                t.node2lane[a] = i + 4 * j + 16 * k;
            }
    // --- per-node tables (xgp, wgp, Ngp, dNgp, dlxigp_ip), gmsh order.
    double lv[4][4];   // lv[m][i] = l_i(xi_m) = delta (computed, not assumed)
    for (int m = 0; m < 4; ++m)
        for (int i = 0; i < 4; ++i) {
            const double xp = t.roots[ord[m]];
            const int L = ord[i];
            double prod = 1.0;
            for (int j = 0; j < 4; ++j)
                if (j != L) prod = prod * ((xp - t.roots[j]) / (t.roots[L] - t.roots[j]));
            lv[m][i] = prod;
        }
    std::memset(t.Ngp, 0, sizeof(t.Ngp));
    for (int g = 0; g < 64; ++g) {
        const int I = t.gmshAtoI[g] - 1, J = t.gmshAtoJ[g] - 1, K = t.gmshAtoK[g] - 1;
        t.xgp[g] = t.roots[ord[I]];
        t.xgp[g + 64] = t.roots[ord[J]];
        t.xgp[g + 128] = t.roots[ord[K]];
        t.wgp[g] = t.weights[ord[I]] * t.weights[ord[J]] * t.weights[ord[K]];
        for (int ii = 0; ii < 4; ++ii) {
            t.dlxigp_ip[g + 64 * 0 + 192 * ii] = t.D[I][ii];
            t.dlxigp_ip[g + 64 * 1 + 192 * ii] = t.D[J][ii];
            t.dlxigp_ip[g + 64 * 2 + 192 * ii] = t.D[K][ii];
        }
        for (int a = 0; a < 64; ++a) {
            const int i = t.gmshAtoI[a] - 1, j = t.gmshAtoJ[a] - 1, k = t.gmshAtoK[a] - 1;
            t.Ngp[g + 64 * a] = lv[I][i] * lv[J][j] * lv[K][k];
            t.dNgp[0 + 3 * a + 192 * g] = t.D[I][i] * lv[J][j] * lv[K][k];
            t.dNgp[1 + 3 * a + 192 * g] = lv[I][i] * t.D[J][j] * lv[K][k];
            t.dNgp[2 + 3 * a + 192 * g] = lv[I][i] * lv[J][j] * t.D[K][k];
        }
    }
}

}  // namespace

std::vector<TableUploader>& uploaders();
std::vector<TableDumper>& dumpers();

const HostTables& host_tables() {
    std::call_once(g_tab_once, build_tables);
    return g_tab;
}

int upload_tables() {
    const HostTables& t = host_tables();
    DevTables d;
    std::memcpy(d.D, t.D, sizeof(d.D));
    std::memcpy(d.wgp, t.wgp, sizeof(d.wgp));
    static const int ord[4] = {0, 3, 1, 2};
    for (int i = 0; i < 4; ++i) d.w1d[i] = t.weights[ord[i]];
    std::memcpy(d.lane2node, t.lane2node, sizeof(d.lane2node));
    std::memcpy(d.node2lane, t.node2lane, sizeof(d.node2lane));
    for (TableUploader fn : uploaders())
        if (fn(d) != 0) return fail(SOD_ERR_HIP, "uploading the p=3 tables to the device failed");
    // self-check: every translation unit reads back exactly what was sent
    DevTables* dev = nullptr;
    SOD_HIP(hipMalloc(&dev, sizeof(DevTables)));
    int bad = -1;
    for (size_t u = 0; u < dumpers().size() && bad < 0; ++u) {
        DevTables back;
        SOD_HIP(hipMemset(dev, 0xff, sizeof(DevTables)));
        if (dumpers()[u](dev, nullptr) != 0) bad = (int)u;
        SOD_HIP(hipDeviceSynchronize());
        SOD_HIP(hipMemcpy(&back, dev, sizeof(DevTables), hipMemcpyDeviceToHost));
        if (std::memcmp(&back, &d, sizeof(DevTables)) != 0) bad = (int)u;
    }
    SOD_HIP(hipFree(dev));
    if (bad >= 0)
        return fail(SOD_ERR_HIP, "constant tables of translation unit %d (of %zu) did not reach the device",
                    bad, dumpers().size());
    return SOD_OK;
}

std::vector<TableUploader>& uploaders() {
    static std::vector<TableUploader> v;
    return v;
}

std::vector<TableDumper>& dumpers() {
    static std::vector<TableDumper> v;
    return v;
}

int register_table_uploader(TableUploader fn) {
    uploaders().push_back(fn);
    return (int)uploaders().size();
}

int register_table_dumper(TableDumper fn) {
    dumpers().push_back(fn);
    return (int)dumpers().size();
}

}  // namespace sod

extern "C" int sod_hex64_tables(int* atoIJK, double* xgp, double* wgp, double* Ngp, double* dNgp,
                                double* dlxigp_ip, int* invAtoIJK, int* gmshAtoI, int* gmshAtoJ,
                                int* gmshAtoK) {
    const sod::HostTables& t = sod::host_tables();
    if (atoIJK) std::memcpy(atoIJK, t.atoIJK, sizeof(t.atoIJK));
    if (xgp) std::memcpy(xgp, t.xgp, sizeof(t.xgp));
    if (wgp) std::memcpy(wgp, t.wgp, sizeof(t.wgp));
    if (Ngp) std::memcpy(Ngp, t.Ngp, sizeof(t.Ngp));
    if (dNgp) std::memcpy(dNgp, t.dNgp, sizeof(t.dNgp));
    if (dlxigp_ip) std::memcpy(dlxigp_ip, t.dlxigp_ip, sizeof(t.dlxigp_ip));
    if (invAtoIJK) std::memcpy(invAtoIJK, t.invAtoIJK, sizeof(t.invAtoIJK));
    if (gmshAtoI) std::memcpy(gmshAtoI, t.gmshAtoI, sizeof(t.gmshAtoI));
    if (gmshAtoJ) std::memcpy(gmshAtoJ, t.gmshAtoJ, sizeof(t.gmshAtoJ));
    if (gmshAtoK) std::memcpy(gmshAtoK, t.gmshAtoK, sizeof(t.gmshAtoK));
    return SOD_OK;
}
