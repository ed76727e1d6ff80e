"""Pins the oracle's init-time geometry (Jacobian inverse, gpvol, lumped
mass, element lengths) and the mesh generator's integer data model."""
import numpy as np
import pytest

import oracle
from _util import gll3_f32, mesh_geometry
from sod2d_refmap_amd import mesh as pmesh


def test_jacobian_inverse_reference_fixture():
    # unitt/unitt_3d_jacobian_inverse: tet with this elcod/dN -> det(He) = 0.0625
    elcod = np.array([[0.0, 2.0, 4.0, 2.0], [0.0, 0.0, 0.0, 2.0], [0.0, 2.0, 0.0, 0.0]])
    dN = np.array([[-1.0, 1.0, 0.0, 0.0], [-1.0, 0.0, 1.0, 0.0], [-1.0, 0.0, 0.0, 1.0]])
    Je = dN @ elcod.T          # Je(i,j) = sum_a dN(i,a) x_a(j)
    det, He = oracle.invert_jacobian(Je)
    assert abs(np.linalg.det(He) - 0.0625) < 1e-10
    assert abs(det - 16.0) < 1e-12


@pytest.mark.parametrize("seed", range(5))
def test_jacobian_inverse_random(seed):
    Je = np.random.default_rng(seed).standard_normal((3, 3)) + 3 * np.eye(3)
    det, He = oracle.invert_jacobian(Je)
    np.testing.assert_allclose(He @ Je, np.eye(3), atol=1e-13)
    np.testing.assert_allclose(det, np.linalg.det(Je), rtol=1e-13)


def test_affine_mesh_closed_forms():
    n = 3
    m = pmesh.periodic_box(n)
    geo = mesh_geometry(m)
    h = 2 * np.pi / n
    _, w1 = gll3_f32()
    t = oracle.hex64_tables()
    He = geo["He"]
    np.testing.assert_allclose(He[0, 0], 2.0 / h, rtol=1e-12)
    np.testing.assert_allclose(He[1, 1], 2.0 / h, rtol=1e-12)
    np.testing.assert_allclose(He[2, 2], 2.0 / h, rtol=1e-12)
    off = He.copy()
    for i in range(3):
        off[i, i] = 0.0
    assert np.abs(off).max() < 1e-12
    gv = geo["gpvol"][0]          # (64, nelem)
    np.testing.assert_allclose(gv, np.broadcast_to(t.wgp[:, None] * (h / 2) ** 3, gv.shape), rtol=1e-12)
    # lumped mass on working nodes: product of 1-D weights times (h/2)^3,
    # node shared by 2^s elements where s = number of element-boundary directions
    np.testing.assert_allclose(geo["Ml"].sum(), (2 * np.pi) ** 3 * t.wgp.sum() / 8.0, rtol=1e-12)
    slave = m.is_slave
    assert np.all(geo["Ml"][slave] == 0.0)
    assert np.all(geo["Ml"][~slave] > 0.0)
    # element lengths: min edge = h; min node distance = h (1 + r1)/2 for r1=-1/sqrt5 (f32)
    np.testing.assert_allclose(geo["helem"], h, rtol=1e-12)
    x, _ = gll3_f32()
    # every node's nearest neighbour is the (-1, -1/sqrt5) gap along an axis
    np.testing.assert_allclose(geo["helem_l"], h * (x[1] + 1.0) / 2.0, rtol=1e-12)


def test_curved_mesh_volume():
    m = pmesh.periodic_box(3, curved_amp=0.05, perm_seed=1)
    geo = mesh_geometry(m)
    # the smooth periodic map preserves the box volume; the GLL rule on
    # FP32-rounded weights integrates J to that volume up to the weight rounding
    t = oracle.hex64_tables()
    vol = geo["gpvol"].sum()
    assert abs(vol / ((2 * np.pi) ** 3 * t.wgp.sum() / 8.0) - 1.0) < 1e-5
    assert np.all(geo["gpvol"] > 0)


@pytest.mark.parametrize("n,curved,perm", [(2, 0.0, None), (3, 0.05, 1), ((2, 3, 4), 0.0, 7)])
def test_mesh_integer_invariants(n, curved, perm):
    m = pmesh.periodic_box(n, curved_amp=curved, perm_seed=perm)
    nx, ny, nz = m.n
    assert m.npoin == (3 * nx + 1) * (3 * ny + 1) * (3 * nz + 1)
    assert m.npoin_w == 27 * nx * ny * nz
    # first-touch numbering: scanning connec_orig element by element, local
    # nodes in gmsh order, new ids appear as 1, 2, 3, ...
    seen = np.zeros(m.npoin + 1, bool)
    nxt = 1
    for e in range(m.nelem):
        for a in range(64):
            v = m.connec_orig[e, a]
            if not seen[v]:
                assert v == nxt
                seen[v] = True
                nxt += 1
    assert nxt == m.npoin + 1
    # working connectivity references only masters, every master exactly
    slave = m.is_slave
    assert not np.any(slave[m.connec - 1])
    assert np.array_equal(np.unique(m.connec), m.lpoin_w)
    assert np.all(np.diff(m.lpoin_w) > 0)
    # connec differs from connec_orig only at slave nodes, and maps each
    # slave to the node at the same position modulo the box
    diff = m.connec != m.connec_orig
    assert np.all(slave[m.connec_orig[diff] - 1])
    lo, lw = m.node_lattice[m.connec_orig[diff] - 1], m.node_lattice[m.connec[diff] - 1]
    M = np.array([3 * nx, 3 * ny, 3 * nz])
    assert np.array_equal(lo % M, lw)
    # each element's 64 nodes form its 4x4x4 lattice block
    t = oracle.hex64_tables()
    inv = t.invAtoIJK - 1
    blk = m.node_lattice[m.connec_orig[:, inv] - 1]          # (nelem,4,4,4,3)
    assert np.all(blk[:, 2, 0, 0, 0] - blk[:, 0, 0, 0, 0] == 1)
    assert np.all(blk[:, 1, 0, 0, 0] - blk[:, 0, 0, 0, 0] == 3)


def test_curved_mapping_periodic():
    m = pmesh.periodic_box(3, curved_amp=0.05)
    L = 2 * np.pi
    slave = m.is_slave
    # slave coordinates equal their master's shifted by a box length
    diff = m.connec != m.connec_orig
    s, w = m.connec_orig[diff] - 1, m.connec[diff] - 1
    shift = (m.coord[s] - m.coord[w]) / L
    np.testing.assert_allclose(shift, np.round(shift), atol=1e-12)
    assert slave.sum() == m.npoin - m.npoin_w
