"""SOD2D HDF5 mesh and restart files (libsod2d_io), on the CPU.

The reference ships no HDF5 file, so the reader is pinned by (i) the
layout its own Fortran loader expects (dataset names, element-major
connectivity, (N,3) coordinate rows, FP32 restart storage), written by our
writer and read back bit for bit; (ii) the generator's mesh, which a file
must reproduce exactly (coordinates FP32-rounded, as the reference's loader
leaves them); (iii) partition consistency across a two-partition file
(global ids, communication lists that pair up); (iv) the restart semantics
of the reference (last element wins when mu_e is stored per node, periodic
slaves copied from masters, time through FP32)."""
import ctypes
import os

import numpy as np
import pytest

from sod2d_refmap_amd import cases, h5io, mesh as pmesh


def _f32(a):
    return np.asarray(a).astype(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def curved():
    return pmesh.periodic_box(3, curved_amp=0.05, perm_seed=9)


def test_file_names():
    # set_hdf5_meshFile_name / set_hdf5_baseResultsFile_name + set_hdf5_restartFile_name
    assert h5io.mesh_file_name("meshes/", "cube", 4) == "meshes/cube-4.hdf"
    assert h5io.restart_file_name("out/", "cube", 4, 2) == "out/restart_cube-4_2.h5"
    with pytest.raises(h5io.IOError_):
        buf = ctypes.create_string_buffer(4)
        h5io._call("sod_h5_mesh_file_name", b"p/", b"name", 1, buf, len(buf))


def test_order_tables():
    t = h5io.order_tables()
    assert np.array_equal(t["a2ijk"], pmesh.hex64_tables().atoIJK)
    assert np.array_equal(t["a2ijk"], t["gmsh2ijk"]) and np.array_equal(t["a2ij"], t["gmsh2ij"])
    for k, n in (("vtk2ijk", 64), ("gmsh2ijk", 64), ("vtk2ij", 16), ("gmsh2ij", 16)):
        assert sorted(t[k]) == list(range(1, n + 1)), k
    # VTK Lagrange hexahedron: vertices 1..8 in VTK's corner order, then the
    # 24 edge nodes, 24 face nodes and the 8 body nodes (i fastest)
    p = 3
    corners = [(0, 0, 0), (p, 0, 0), (p, p, 0), (0, p, 0), (0, 0, p), (p, 0, p), (p, p, p), (0, p, p)]
    assert [h5io._vtk_hex_index(*c) for c in corners] == list(range(1, 9))
    nb = lambda c: sum(v in (0, p) for v in c)
    lat = [(i, j, k) for k in range(4) for j in range(4) for i in range(4)]
    idx = {c: h5io._vtk_hex_index(*c) for c in lat}
    assert sorted(idx[c] for c in lat if nb(c) == 2) == list(range(9, 33))
    assert sorted(idx[c] for c in lat if nb(c) == 1) == list(range(33, 57))
    assert [idx[(i, j, k)] for k in (1, 2) for j in (1, 2) for i in (1, 2)] == list(range(57, 65))
    # the first edge runs from vertex 1 to vertex 2 with i increasing
    assert (idx[(1, 0, 0)], idx[(2, 0, 0)]) == (9, 10)
    # gmsh quadrilateral: corners first, counter-clockwise
    s2g = h5io._s2g()
    pos = {(s2g[ip], s2g[jp]): n for n, (ip, jp) in enumerate((a, b) for a in range(4) for b in range(4))}
    assert [t["gmsh2ij"][pos[c]] for c in ((0, 0), (p, 0), (p, p), (0, p))] == [1, 2, 3, 4]


@pytest.mark.parametrize("kind", ["straight", "curved"])
def test_mesh_roundtrip_one_partition(tmp_path, kind, curved):
    m = pmesh.periodic_box(2) if kind == "straight" else curved
    path = str(tmp_path / "box-1.hdf")
    h5io.write_mesh(m, path)
    p = h5io.read_mesh(path)
    i = p.info
    assert (i["nranks"], i["porder"], i["nnode"]) == (1, 3, 64)
    assert (i["npoin"], i["nelem"], i["npoin_w"]) == (m.npoin, m.nelem, m.npoin_w)
    assert (i["total_nodes"], i["total_elems"], i["is_periodic"], i["has_boundaries"]) == (m.npoin, m.nelem, 1, 0)
    r = p.mesh
    assert np.array_equal(r.connec_orig, m.connec_orig) and r.connec_orig.flags.f_contiguous
    assert np.array_equal(r.connec, m.connec)
    assert np.array_equal(r.lpoin_w, m.lpoin_w)
    # the reference reads coordinates through FP32 (rp_vtk = 4)
    assert np.array_equal(r.coord, _f32(m.coord))
    assert 0 < np.abs(r.coord - m.coord).max() < 1e-6
    assert np.array_equal(p.masSla, h5io.periodic_pairs(m))
    assert np.array_equal(p.globalIdSrl, np.arange(1, m.npoin + 1)) and np.array_equal(p.elemGid, np.arange(1, m.nelem + 1))
    # every slave of connecParOrig is replaced by its master in connecParWork
    co, cw = r.connec_orig, r.connec
    mas = dict(zip(p.masSla[:, 1], p.masSla[:, 0]))
    assert all(cw.flat[t] == mas.get(co.flat[t], co.flat[t]) for t in range(co.size))
    assert not np.isin(p.masSla[:, 1], r.lpoin_w).any() and np.isin(p.masSla[:, 0], r.lpoin_w).all()


def test_mesh_two_partitions(tmp_path, curved):
    m = curved
    path = str(tmp_path / "box-2.hdf")
    h5io.write_mesh(m, path, nranks=2)
    parts = [h5io.read_mesh(path, r) for r in range(2)]
    e_seen = 0
    work = []
    for r, p in enumerate(parts):
        i, g = p.info, p.globalIdSrl
        assert i["nranks"] == 2 and i["rank_elem_start"] == e_seen + 1
        e0, ne = e_seen, i["nelem"]
        e_seen += ne
        # local ids map back to the serial mesh through globalIdSrl
        assert np.array_equal(g[p.mesh.connec_orig - 1], m.connec_orig[e0:e0 + ne])
        assert np.array_equal(g[p.mesh.connec - 1], m.connec[e0:e0 + ne])
        assert np.array_equal(p.mesh.coord, _f32(m.coord[g - 1]))
        assert np.array_equal(p.elemGid, np.arange(e0 + 1, e0 + ne + 1))
        assert np.array_equal(p.globalIdPar, np.arange(i["rank_node_start"], i["rank_node_end"] + 1))
        work.append(g[p.mesh.lpoin_w - 1])
        pairs = g[p.masSla - 1]
        ref = {tuple(x) for x in h5io.periodic_pairs(m)}
        assert all(tuple(x) in ref for x in pairs)
    assert e_seen == m.nelem and parts[1].info["rank_node_start"] == parts[0].info["npoin"] + 1
    assert set(np.concatenate(work)) == set(m.lpoin_w)
    # communication lists pair up: the same global nodes, in the same order
    c0, c1 = parts[0].comms, parts[1].comms
    assert list(c0["ranksToComm"]) == [1] and list(c1["ranksToComm"]) == [0]
    n0 = parts[0].globalIdSrl[c0["nodesToComm"] - 1]
    n1 = parts[1].globalIdSrl[c1["nodesToComm"] - 1]
    assert len(n0) > 0 and np.array_equal(n0, n1)
    assert list(c0["commsMemSize"]) == [len(n0)] and list(c0["commsMemPosInNgb"]) == list(c1["commsMemPosInLoc"])
    with pytest.raises(h5io.IOError_):
        h5io.read_mesh(path, 2)


def test_mesh_errors(tmp_path):
    with pytest.raises(h5io.IOError_, match="cannot open"):
        h5io.read_mesh(str(tmp_path / "missing.hdf"))
    t = h5io.order_tables()
    one = np.ones(1, np.int32)
    w = ctypes.c_void_p()
    with pytest.raises(h5io.IOError_, match="p = 3"):
        h5io._call("sod_h5_mesh_create", str(tmp_path / "p4.hdf").encode(), 1, 4, one.ctypes.data,
                   np.ones(1, np.int64).ctypes.data, one.ctypes.data, None, None, None,
                   *[t[k].ctypes.data for k in ("a2ijk", "gmsh2ijk", "vtk2ijk", "a2ij", "gmsh2ij", "vtk2ij")],
                   ctypes.byref(w))
    # a file whose element-local node order is not the one the kernels use
    m = pmesh.periodic_box(2)
    path = str(tmp_path / "perm-1.hdf")
    saved = h5io.order_tables

    def swapped():
        d = saved()
        d["a2ijk"] = d["a2ijk"][::-1].copy()
        return d

    h5io.order_tables = swapped
    try:
        h5io.write_mesh(m, path)
    finally:
        h5io.order_tables = saved
    with pytest.raises(h5io.IOError_, match="node order"):
        h5io.read_mesh(path)


def _last_writer(connec_orig, npoin, vals):
    """node value when the element loop (outside) and node loop (inside)
    store each element-node value: the last (element, node) pair wins."""
    ne = connec_orig.shape[0]
    out = np.zeros(npoin)
    for e in range(ne):
        out[connec_orig[e] - 1] = vals[e]
    return out


def test_restart_roundtrip(tmp_path, curved):
    m = curved
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas, perturb_seed=3, perturb_amp=1e-2)
    rng = np.random.default_rng(11)
    s["mu_e"] = np.asfortranarray(rng.random((m.nelem, 64)))
    s["mu_sgs"] = np.asfortranarray(rng.random((m.nelem, 64)) * 1e-3)
    path = h5io.restart_file_name(str(tmp_path) + "/", "box", 1, 1)
    h5io.save_restart(path, m, s, time=0.1, istep=42)
    f = h5io.load_restart_fields(path, m)
    assert f["istep"] == 42 and f["time"] == float(np.float32(0.1))
    pairs = h5io.periodic_pairs(m)
    for k in ("rho", "pr", "E"):
        want = _f32(s[k])
        want[pairs[:, 1] - 1] = want[pairs[:, 0] - 1]
        assert np.array_equal(f[k], want), k
    want_u = _f32(s["u"])
    want_u[pairs[:, 1] - 1] = want_u[pairs[:, 0] - 1]
    assert np.array_equal(f["u"], want_u)
    for k in ("mu_e", "mu_sgs"):
        node = _last_writer(m.connec_orig, m.npoin, s[k])
        assert np.array_equal(f[k], _f32(node[m.connec_orig - 1])), k
    # walave_u absent from the file: the loader takes u
    g = h5io.load_restart_fields(path, m, walave=True)
    assert np.array_equal(g["walave_u"], f["u"])
    wa = np.asfortranarray(rng.random((m.npoin, 3)))
    h5io.save_restart(path, m, s, 0.2, 43, walave_u=wa)
    g = h5io.load_restart_fields(path, m, walave=True)
    want = _f32(wa)
    want[pairs[:, 1] - 1] = want[pairs[:, 0] - 1]
    assert np.array_equal(g["walave_u"], want)


def test_restart_partitions_share_one_file(tmp_path, curved):
    """Two node slices written into one file, one after the other, read back
    slice by slice."""
    m = curved
    path = str(tmp_path / "two.h5")
    path_mesh = str(tmp_path / "box-2.hdf")
    h5io.write_mesh(m, path_mesh, nranks=2)
    parts = [h5io.read_mesh(path_mesh, r) for r in range(2)]
    total = parts[0].info["total_nodes"]
    gas = cases.GasParams()
    full = cases.tgv_initial_state(m, gas)
    states = []
    for r, p in enumerate(parts):
        g = p.globalIdSrl - 1
        st = {k: np.asfortranarray(full[k][g]) for k in ("rho", "u", "pr", "E")}
        st["mu_e"] = np.asfortranarray(np.full((p.info["nelem"], 64), 1.0 + r))
        states.append(st)
        h5io.save_restart(path, p.mesh, st, 0.5, 9, total_nodes=total, node_start=p.info["rank_node_start"],
                          create=(r == 0))
    for r, p in enumerate(parts):
        f = h5io.load_restart_fields(path, p.mesh, masSla=p.masSla, node_start=p.info["rank_node_start"])
        want = _f32(states[r]["rho"])
        want[p.masSla[:, 1] - 1] = want[p.masSla[:, 0] - 1]
        assert np.array_equal(f["rho"], want)
        assert np.all(f["mu_e"] == 1.0 + r)


def test_state_from_restart_matches_initial_condition(curved):
    """Loading the TGV initial fields gives the solver state the initial
    condition gives (same formulas; e_int and the Mach number are computed
    differently by the two reference routines)."""
    m = curved
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas)
    f = {k: s[k] for k in ("rho", "u", "pr", "E", "mu_e", "mu_sgs")}
    st = h5io.state_from_restart(f, gas)
    for k in ("q", "Tem", "csound", "eta", "mu_fluid", "mu_factor"):
        assert np.array_equal(st[k], s[k]), k
    assert np.abs(st["e_int"] / s["e_int"] - 1).max() < 1e-13
    uu = np.sum(s["u"] ** 2, axis=1)
    assert np.allclose(st["machno"], np.sqrt(uu) / s["csound"], rtol=0, atol=1e-15)


def test_restart_argument_errors(tmp_path, curved):
    m = curved
    s = cases.tgv_initial_state(m, cases.GasParams())
    bad = pmesh.Mesh(n=None, L=None, nelem=m.nelem, npoin=10, npoin_w=10, connec_orig=m.connec_orig,
                     connec=m.connec, lpoin_w=m.lpoin_w, coord=m.coord, elem_lex=None, node_lattice=None)
    with pytest.raises(h5io.IOError_, match="outside"):
        h5io.save_restart(str(tmp_path / "x.h5"), bad, {k: v[:10] for k, v in s.items() if k in ("rho", "u", "pr", "E")}
                          | {"mu_e": s["mu_e"]}, 0.0, 0)
    with pytest.raises(h5io.IOError_, match="cannot open"):
        h5io.load_restart_fields(str(tmp_path / "none.h5"), m)
