// Host-side generator of structured periodic HEX64 meshes in the reference's
// partitioned-mesh data model (the arrays its mesh conversion tool writes):
//   * local node order a2ijk = gmsh2ijk (mod_meshConversorTool.f90:150-152);
//   * high-order nodes at the FP32-rounded GLL positions of each element
//     (interpolateOriginalCoordinates, mod_meshConversorTool.f90:3921-3942);
//   * rank-local ids by first touch, elements in order, local nodes in gmsh
//     order (reorder_nodes_in_mshRank, mod_meshConversorTool.f90:2864-3076);
//   * periodic slaves (x=L, y=L or z=L faces) replaced by their masters in
//     connecParWork; working nodes = non-slave ids in increasing order
//     (create_working_lists_parallel, mod_meshConversorTool.f90:3944-4035).
// Host only; writes straight into caller-allocated arrays.
#include <cmath>
#include <cstdint>
#include <vector>

#include "sod2d_internal.hpp"

using namespace sod;

extern "C" int sod_structured_box_mesh(int nx, int ny, int nz, double Lx, double Ly, double Lz,
                                       double curved_amp, const long long* elem_order,
                                       int* connec_orig, int* connec, int* lpoin_w, double* coord,
                                       int* node_lattice) {
    if (nx < 1 || ny < 1 || nz < 1) return fail(SOD_ERR_ARG, "structured mesh: n must be >= 1");
    const HostTables& t = host_tables();
    const int s2lat[4] = {0, 3, 1, 2};
    int li[64], lj[64], lk[64];
    for (int a = 0; a < 64; ++a) {
        li[a] = s2lat[t.gmshAtoI[a] - 1];
        lj[a] = s2lat[t.gmshAtoJ[a] - 1];
        lk[a] = s2lat[t.gmshAtoK[a] - 1];
    }
    const int64_t nelem = (int64_t)nx * ny * nz;
    const int64_t Mx = 3 * (int64_t)nx + 1, My = 3 * (int64_t)ny + 1, Mz = 3 * (int64_t)nz + 1;
    const int64_t Mtot = Mx * My * Mz;
    if (Mtot > INT32_MAX) return fail(SOD_ERR_ARG, "structured mesh: too many nodes for int32 ids");
    std::vector<int32_t> new_of_lat(Mtot, -1);
    std::vector<int64_t> lat_of_new(Mtot);
    int32_t cnt = 0;
    for (int64_t e = 0; e < nelem; ++e) {
        const int64_t lex = elem_order ? elem_order[e] : e;
        const int64_t ex = lex % nx, ey = (lex / nx) % ny, ez = lex / ((int64_t)nx * ny);
        for (int a = 0; a < 64; ++a) {
            const int64_t id = (3 * ex + li[a]) + Mx * ((3 * ey + lj[a]) + My * (3 * ez + lk[a]));
            int32_t v = new_of_lat[id];
            if (v < 0) {
                v = cnt++;
                new_of_lat[id] = v;
                lat_of_new[v] = id;
            }
            connec_orig[e + nelem * a] = v + 1;
        }
    }
    for (int64_t e = 0; e < nelem; ++e) {
        const int64_t lex = elem_order ? elem_order[e] : e;
        const int64_t ex = lex % nx, ey = (lex / nx) % ny, ez = lex / ((int64_t)nx * ny);
        for (int a = 0; a < 64; ++a) {
            const int64_t I = (3 * ex + li[a]) % (3 * nx), J = (3 * ey + lj[a]) % (3 * ny),
                          K = (3 * ez + lk[a]) % (3 * nz);
            connec[e + nelem * a] = new_of_lat[I + Mx * (J + My * K)] + 1;
        }
    }
    const double* r = t.roots;
    auto axis = [&](int64_t idx, int nel, double L) {
        const double h = L / nel;
        int64_t e = idx / 3;
        if (e > nel - 1) e = nel - 1;
        const int64_t loc = idx - 3 * e;
        return (double)e * h + (1.0 + r[loc]) * (0.5 * h);
    };
    const double tp = 2.0 * 3.14159265358979323846;
    int64_t w = 0;
    for (int64_t v = 0; v < Mtot; ++v) {
        const int64_t id = lat_of_new[v];
        const int64_t I = id % Mx, J = (id / Mx) % My, K = id / (Mx * My);
        double x = axis(I, nx, Lx), y = axis(J, ny, Ly), z = axis(K, nz, Lz);
        if (curved_amp != 0.0) {
            const double sx = std::sin(tp * x / Lx), sy = std::sin(tp * y / Ly), sz = std::sin(tp * z / Lz);
            const double x1 = x + curved_amp * (Lx / nx) * sy * sz;
            const double y1 = y + curved_amp * (Ly / ny) * sz * sx;
            const double z1 = z + curved_amp * (Lz / nz) * sx * sy;
            x = x1; y = y1; z = z1;
        }
        coord[v] = x;
        coord[v + Mtot] = y;
        coord[v + 2 * Mtot] = z;
        if (node_lattice) {
            node_lattice[3 * v] = (int)I;
            node_lattice[3 * v + 1] = (int)J;
            node_lattice[3 * v + 2] = (int)K;
        }
        const bool slave = I == 3 * (int64_t)nx || J == 3 * (int64_t)ny || K == 3 * (int64_t)nz;
        if (!slave) lpoin_w[w++] = (int)(v + 1);
    }
    return SOD_OK;
}
