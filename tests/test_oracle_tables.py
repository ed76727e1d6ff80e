"""Pins the oracle's p=3 tables (GLL, Lagrange derivatives, gmsh order,
a<->ijk maps) against closed forms and structural properties, and checks
that the library's host tables are identical to the oracle's."""
import numpy as np

import oracle
from _util import SOD2D_ORDER, deriv_matrix_sod2d, gll3_f32
from sod2d_refmap_amd import mesh as pmesh


def test_gll_closed_form_fp32_rounded():
    r, w = oracle.gll_roots_weights()
    xr, wr = gll3_f32()
    # the reference rounds FP64 Newton roots through FP32 (mod_maths.f90:39-40)
    assert np.array_equal(r, xr)
    assert np.array_equal(w, wr)
    assert r[2] == np.float64(np.float32(0.4472135954999579))
    assert r[2] != 1.0 / np.sqrt(5.0)


def test_derivative_matrix_matches_barycentric_form():
    t = oracle.hex64_tables()
    D = deriv_matrix_sod2d()
    for g in range(64):
        for d, tab in enumerate((t.gmshAtoI, t.gmshAtoJ, t.gmshAtoK)):
            m = tab[g] - 1
            np.testing.assert_allclose(t.dlxigp_ip[g, d, :], D[m, :], rtol=0, atol=2e-14)
    # rows of a differentiation matrix sum to zero (derivative of a constant)
    assert np.abs(t.dlxigp_ip.sum(axis=2)).max() < 1e-14


def test_gmsh_order_structure():
    t = oracle.hex64_tables()
    inv = t.invAtoIJK - 1
    assert sorted(inv.ravel(order="F").tolist()) == list(range(64))
    # sod2d -> lattice offset, corners first in gmsh order
    s2l = SOD2D_ORDER
    lat = np.zeros((64, 3), int)
    for i in range(4):
        for j in range(4):
            for k in range(4):
                lat[inv[i, j, k]] = (s2l[i], s2l[j], s2l[k])
    corners = [(0, 0, 0), (3, 0, 0), (3, 3, 0), (0, 3, 0), (0, 0, 3), (3, 0, 3), (3, 3, 3), (0, 3, 3)]
    assert [tuple(x) for x in lat[:8]] == corners
    # 12 edges x 2 nodes: each pair lies on a corner-corner edge, in order
    edges = [(1, 2), (1, 4), (1, 5), (2, 3), (2, 6), (3, 4), (3, 7), (4, 8), (5, 6), (5, 8), (6, 7), (7, 8)]
    for e, (a, b) in enumerate(edges):
        A, B = np.array(corners[a - 1]), np.array(corners[b - 1])
        for s in (1, 2):
            assert tuple(lat[8 + 2 * e + s - 1]) == tuple(A + s * (B - A) // 3)
    # 6 faces x 4 interior nodes, then the 8 interior nodes
    onbound = lambda x: np.any((x == 0) | (x == 3))
    assert all(onbound(lat[n]) for n in range(32, 56))
    assert all(not onbound(lat[n]) for n in range(56, 64))
    # gmshAto* are the inverse of invAtoIJK
    for a in range(64):
        assert inv[t.gmshAtoI[a] - 1, t.gmshAtoJ[a] - 1, t.gmshAtoK[a] - 1] == a
    # atoIJK enumerates (k, i, j) with j fastest (mod_geom.f90:123-144)
    c = 0
    for k in range(4):
        for i in range(4):
            for j in range(4):
                assert t.atoIJK[c] - 1 == inv[i, j, k]
                c += 1


def test_shape_functions_collocated():
    t = oracle.hex64_tables()
    assert np.array_equal(t.Ngp, np.eye(64))      # Kronecker property at the nodes
    D = deriv_matrix_sod2d()
    I, J, K = t.gmshAtoI - 1, t.gmshAtoJ - 1, t.gmshAtoK - 1
    for g in range(64):
        for a in range(64):
            exp = [D[I[g], I[a]] * (J[g] == J[a]) * (K[g] == K[a]),
                   D[J[g], J[a]] * (I[g] == I[a]) * (K[g] == K[a]),
                   D[K[g], K[a]] * (I[g] == I[a]) * (J[g] == J[a])]
            np.testing.assert_allclose(t.dNgp[:, a, g], exp, atol=2e-14, rtol=0)
    x, w = gll3_f32()
    assert abs(t.wgp.sum() - 8.0) < 1e-6      # weights of the FP32 rule sum to ~8
    for g in range(64):
        np.testing.assert_allclose(t.xgp[g], x[SOD2D_ORDER[[I[g], J[g], K[g]]]], rtol=0, atol=0)


def test_library_tables_bit_identical_to_oracle():
    o = oracle.hex64_tables()
    p = pmesh.hex64_tables()
    for k in ("atoIJK", "invAtoIJK", "gmshAtoI", "gmshAtoJ", "gmshAtoK", "xgp", "wgp", "Ngp", "dNgp",
              "dlxigp_ip"):
        assert np.array_equal(getattr(o, k), getattr(p, k)), k
