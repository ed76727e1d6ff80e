"""SOD2D mesh files and restart files (partitioned HDF5) through libsod2d_io.

Host-side I/O: a mesh file written by the reference's mesh tool, or by
`write_mesh`, becomes a `Mesh` that the device path runs on; a restart
file of the reference (`restart_<mesh>-<nranks>_<cnt>.h5`) becomes the
level-2 state dict the solvers take, derived from the stored fields as the
reference does after loading (CFDSolverBase.f90:2066-2113, 1081-1095,
1059-1079).

File layout and the reference routines each function follows are cited in
include/sod2d_io.h (mod_hdf5.f90 of the reference).  Coordinates read from a
mesh file and every restart field are FP32-rounded, as the reference's own
loader leaves them.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import cases
from .mesh import Mesh, hex64_tables

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsod2d_io.so")

INFO_KEYS = ("nranks", "porder", "nnode", "total_nodes", "total_elems", "is_periodic", "is_lineal_output",
             "has_boundaries", "rank_node_start", "rank_node_end", "rank_elem_start", "rank_elem_end",
             "npoin", "nelem", "npoin_w", "nper", "num_ranks_with_comms", "num_nodes_to_comm")

_P = ctypes.c_void_p
_I = ctypes.c_int
_LL = ctypes.c_longlong
_SIGS = {
    "sod_io_last_error": ([], ctypes.c_char_p),
    "sod_h5_mesh_file_name": ([ctypes.c_char_p, ctypes.c_char_p, _I, _P, _I], _I),
    "sod_h5_restart_file_name": ([ctypes.c_char_p, ctypes.c_char_p, _I, _I, _P, _I], _I),
    "sod_h5_mesh_open": ([ctypes.c_char_p, _I, _P], _I),
    "sod_h5_mesh_info": ([_P, _P], _I),
    "sod_h5_mesh_read": ([_P] + [_P] * 11, _I),
    "sod_h5_mesh_read_comms": ([_P] + [_P] * 5, _I),
    "sod_h5_mesh_close": ([_P], None),
    "sod_h5_mesh_create": ([ctypes.c_char_p, _I, _I] + [_P] * 12 + [_P], _I),
    "sod_h5_mesh_write_rank": ([_P, _I] + [_P] * 13, _I),
    "sod_h5_mesh_writer_close": ([_P], _I),
    "sod_h5_restart_save": ([ctypes.c_char_p, _I, _LL, _LL, _I, _I, _P, _I, ctypes.c_double] + [_P] * 7, _I),
    "sod_h5_restart_load": ([ctypes.c_char_p, _LL, _I, _I, _P, _I, _P, _P, _P] + [_P] * 7, _I),
}

_lib = None


class IOError_(RuntimeError):
    """A failed libsod2d_io call (the message is the library's)."""


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (sod2d_refmap_amd.build.build_io())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _call(name: str, *args) -> int:
    rc = getattr(lib(), name)(*args)
    if rc is not None and rc < 0:
        raise IOError_(f"{name}: {lib().sod_io_last_error().decode()} (code {rc})")
    return rc


def _p(a) -> int | None:
    return None if a is None else a.ctypes.data


def _name(fn: str, *args) -> str:
    buf = ctypes.create_string_buffer(4096)
    _call(fn, *args, buf, len(buf))
    return buf.value.decode()


def mesh_file_name(path: str, name: str, nranks: int) -> str:
    """<path><name>-<nranks>.hdf (set_hdf5_meshFile_name)."""
    return _name("sod_h5_mesh_file_name", path.encode(), name.encode(), nranks)


def restart_file_name(path: str, mesh_name: str, nranks: int, restart_cnt: int) -> str:
    """<path>restart_<mesh>-<nranks>_<cnt>.h5 (set_hdf5_restartFile_name)."""
    return _name("sod_h5_restart_file_name", path.encode(), mesh_name.encode(), nranks, restart_cnt)


@dataclass
class MeshPartition:
    """One partition of a mesh file: the Mesh the solvers take plus the
    file's other arrays."""
    mesh: Mesh
    info: dict
    masSla: np.ndarray          # (nper, 2) int32 F, 1-based (master, slave)
    globalIdSrl: np.ndarray     # (npoin,) int64
    globalIdPar: np.ndarray     # (npoin,) int64
    elemGid: np.ndarray         # (nelem,) int32
    order: dict                 # a2ijk, gmsh2ijk, vtk2ijk (64,) int32
    comms: dict                 # ranksToComm, commsMemPosInLoc, commsMemSize, commsMemPosInNgb, nodesToComm


def read_mesh(path: str, rank: int = 0) -> MeshPartition:
    """Partition `rank` of a SOD2D mesh file (load_hdf5_meshFile).

    The file's element-local node order must be the one the device path is
    built for (a2ijk = gmsh2ijk of the p=3 hexahedron); another order is an
    error rather than a silent mis-read."""
    h = ctypes.c_void_p()
    _call("sod_h5_mesh_open", os.fsencode(path), rank, ctypes.byref(h))
    try:
        raw = np.zeros(len(INFO_KEYS), np.int64)
        _call("sod_h5_mesh_info", h, raw.ctypes.data)
        info = {k: int(v) for k, v in zip(INFO_KEYS, raw)}
        ne, nn, nw, npr = info["nelem"], info["npoin"], info["npoin_w"], info["nper"]
        co = np.zeros((ne, 64), np.int32, order="F")
        cw = np.zeros((ne, 64), np.int32, order="F")
        lw = np.zeros(nw, np.int32)
        xyz = np.zeros((nn, 3), order="F")
        ms = np.zeros((npr, 2), np.int32, order="F")
        gs = np.zeros(nn, np.int64)
        gp = np.zeros(nn, np.int64)
        eg = np.zeros(ne, np.int32)
        order = {k: np.zeros(64, np.int32) for k in ("a2ijk", "gmsh2ijk", "vtk2ijk")}
        _call("sod_h5_mesh_read", h, _p(co), _p(cw), _p(lw), _p(xyz), _p(ms), _p(gs), _p(gp), _p(eg),
              _p(order["a2ijk"]), _p(order["gmsh2ijk"]), _p(order["vtk2ijk"]))
        nr, nc = info["num_ranks_with_comms"], info["num_nodes_to_comm"]
        comms = {k: np.zeros(nr, np.int32) for k in ("ranksToComm", "commsMemPosInLoc", "commsMemSize",
                                                      "commsMemPosInNgb")}
        comms["nodesToComm"] = np.zeros(nc, np.int32)
        _call("sod_h5_mesh_read_comms", h, *[_p(comms[k]) for k in ("ranksToComm", "commsMemPosInLoc",
                                                                   "commsMemSize", "commsMemPosInNgb",
                                                                   "nodesToComm")])
    finally:
        lib().sod_h5_mesh_close(h)
    want = hex64_tables().atoIJK
    if not np.array_equal(order["a2ijk"], want):
        raise IOError_(f"{path}: element-local node order (/order/a2ijk) differs from the p=3 gmsh order "
                       "this build's kernels use")
    m = Mesh(n=None, L=None, nelem=ne, npoin=nn, npoin_w=nw, connec_orig=co, connec=cw, lpoin_w=lw, coord=xyz,
             elem_lex=None, node_lattice=None)
    return MeshPartition(mesh=m, info=info, masSla=ms, globalIdSrl=gs, globalIdPar=gp, elemGid=eg, order=order,
                         comms=comms)


# ----------------------------------------------------------- order tables

def _s2g(p: int = 3):
    """sod2d 1-D index -> gmsh lattice index: (0, p, 1..p-1)
    (set_allocate_array_ijk_sod2d_criteria, mod_ijk_indices.f90:27-52)."""
    return [0, p] + list(range(1, p))


def _vtk_hex_index(i: int, j: int, k: int, p: int = 3) -> int:
    """1-based point index of lattice node (i, j, k) in VTK's Lagrange
    hexahedron (VTK's published vtkHigherOrderHexahedron::PointIndexFromIJK
    ordering: vertices, edges, faces, body)."""
    ib, jb, kb = i in (0, p), j in (0, p), k in (0, p)
    nb = ib + jb + kb
    if nb == 3:
        return 1 + ((2 if j else 1) if i else (3 if j else 0)) + (4 if k else 0)
    off = 8
    q = p - 1
    if nb == 2:
        if not ib:
            return 1 + off + (i - 1) + ((q + q) if j else 0) + ((2 * (q + q)) if k else 0)
        if not jb:
            return 1 + off + (j - 1) + (q if i else (2 * q + q)) + ((2 * (q + q)) if k else 0)
        off += 4 * q + 4 * q
        return 1 + off + (k - 1) + q * ((3 if j else 1) if i else (2 if j else 0))
    off += 4 * (q + q + q)
    if nb == 1:
        if ib:
            return 1 + off + (j - 1) + q * (k - 1) + (q * q if i else 0)
        off += 2 * q * q
        if jb:
            return 1 + off + (i - 1) + q * (k - 1) + (q * q if j else 0)
        off += 2 * q * q
        return 1 + off + (i - 1) + q * (j - 1) + (q * q if k else 0)
    off += 2 * (q * q + q * q + q * q)
    return 1 + off + (i - 1) + q * ((j - 1) + q * (k - 1))


def _vtk_quad_index(i: int, j: int, p: int = 3) -> int:
    """1-based point index of (i, j) in VTK's Lagrange quadrilateral."""
    ib, jb = i in (0, p), j in (0, p)
    q = p - 1
    if ib and jb:
        return 1 + ((2 if j else 1) if i else (3 if j else 0))
    off = 4
    if not ib and jb:
        return 1 + off + (i - 1) + ((q + q) if j else 0)
    if ib and not jb:
        return 1 + off + (j - 1) + (q if i else (2 * q + q))
    off += 2 * (q + q)
    return 1 + off + (i - 1) + q * (j - 1)


def _gmsh_quad_lattice(p: int = 3):
    """(i, j) lattice of each node of gmsh's order-3 quadrilateral: corners,
    edges 0-1, 1-2, 2-3, 3-0, then the interior order-1 quad."""
    c = [(0, 0), (p, 0), (p, p), (0, p)]
    pts = list(c)
    for a, b in ((0, 1), (1, 2), (2, 3), (3, 0)):
        for s in range(1, p):
            pts.append((c[a][0] + s * (c[b][0] - c[a][0]) // p, c[a][1] + s * (c[b][1] - c[a][1]) // p))
    pts += [(1, 1), (2, 1), (2, 2), (1, 2)]
    return pts


def order_tables(p: int = 3) -> dict:
    """The /order tables of a p=3 mesh file (mod_ijk_indices.f90:61-154):
    for each sod2d position (loops k, i, j for hexahedra and i, j for
    quadrilaterals) the gmsh and VTK point index.  a2ijk is gmsh2ijk and
    a2ij is gmsh2ij (mod_meshConversorTool.f90:150-152)."""
    s2g = _s2g(p)
    hexa = hex64_tables().atoIJK.astype(np.int32)
    vtk = np.array([_vtk_hex_index(s2g[ip], s2g[jp], s2g[kp], p)
                    for kp in range(p + 1) for ip in range(p + 1) for jp in range(p + 1)], np.int32)
    lat = {pt: n + 1 for n, pt in enumerate(_gmsh_quad_lattice(p))}
    gq = np.array([lat[(s2g[ip], s2g[jp])] for ip in range(p + 1) for jp in range(p + 1)], np.int32)
    vq = np.array([_vtk_quad_index(s2g[ip], s2g[jp], p) for ip in range(p + 1) for jp in range(p + 1)], np.int32)
    return dict(a2ijk=hexa, gmsh2ijk=hexa.copy(), vtk2ijk=vtk, a2ij=gq, gmsh2ij=gq.copy(), vtk2ij=vq)


# ------------------------------------------------------------- writing

def periodic_pairs(mesh: Mesh) -> np.ndarray:
    """(nper, 2) 1-based (master, slave) pairs: each slave id of connecParOrig
    with the master id that connecParWork holds in its place, ordered by
    slave id."""
    co, cw = np.asarray(mesh.connec_orig).ravel(order="F"), np.asarray(mesh.connec).ravel(order="F")
    sl = co != cw
    pairs = np.unique(np.stack([co[sl], cw[sl]], axis=1), axis=0)     # (slave, master), sorted by slave
    if len(np.unique(pairs[:, 0])) != len(pairs):
        raise ValueError("a slave node maps to two masters")
    return np.asfortranarray(pairs[:, ::-1].astype(np.int32))


def _partition(mesh: Mesh, nranks: int):
    """Contiguous element blocks; each block numbers the nodes its elements
    touch (and the masters of its slaves) by first touch.  Nodes held by
    several blocks are listed for communication, grouped by neighbour.  One
    block keeps the mesh's own node ids."""
    ne = mesh.nelem
    bounds = [ne * r // nranks for r in range(nranks + 1)]
    co = np.asarray(mesh.connec_orig)
    cw = np.asarray(mesh.connec)
    wset = np.zeros(mesh.npoin + 1, bool)
    wset[mesh.lpoin_w] = True
    parts = []
    for r in range(nranks):
        e0, e1 = bounds[r], bounds[r + 1]
        if nranks == 1:
            glob = np.arange(1, mesh.npoin + 1)
        else:
            ids = np.concatenate([co[e0:e1].ravel(), cw[e0:e1].ravel()])
            _, first = np.unique(ids, return_index=True)
            glob = ids[np.sort(first)]                              # global ids by first touch
        loc = np.zeros(mesh.npoin + 1, np.int64)
        loc[glob] = np.arange(1, len(glob) + 1)
        parts.append(dict(e0=e0, e1=e1, glob=glob, loc=loc))
    owners = {}
    for r, pt in enumerate(parts):
        for g in pt["glob"]:
            owners.setdefault(int(g), []).append(r)
    shared = {g: rs for g, rs in owners.items() if len(rs) > 1}
    for r, pt in enumerate(parts):
        nbrs = sorted({q for g, rs in shared.items() if r in rs for q in rs if q != r})
        lists = {q: sorted(g for g, rs in shared.items() if r in rs and q in rs) for q in nbrs}
        pt["nbrs"], pt["lists"] = nbrs, lists
    for r, pt in enumerate(parts):
        pos, nodes, size, ngb = [], [], [], []
        for q in pt["nbrs"]:
            pos.append(len(nodes) + 1)
            nodes += [int(pt["loc"][g]) for g in pt["lists"][q]]
            size.append(len(pt["lists"][q]))
        pt["comm"] = dict(ranksToComm=np.array(pt["nbrs"], np.int32), commsMemPosInLoc=np.array(pos, np.int32),
                          commsMemSize=np.array(size, np.int32), nodesToComm=np.array(nodes, np.int32))
    for r, pt in enumerate(parts):
        ngb = []
        for q in pt["nbrs"]:
            k = parts[q]["nbrs"].index(r)
            ngb.append(int(parts[q]["comm"]["commsMemPosInLoc"][k]))
        pt["comm"]["commsMemPosInNgb"] = np.array(ngb, np.int32)
    return parts, wset


def write_mesh(mesh: Mesh, path: str, nranks: int = 1, global_ids: np.ndarray | None = None) -> None:
    """Writes `mesh` as a SOD2D mesh file of `nranks` partitions (contiguous
    element blocks).  Node and element global ids are the serial ids
    (`global_ids` overrides the node ids, 1-based, npoin entries)."""
    F = np.asfortranarray
    parts, wset = _partition(mesh, nranks)
    pairs = periodic_pairs(mesh)
    gid = np.arange(1, mesh.npoin + 1, dtype=np.int64) if global_ids is None else np.asarray(global_ids, np.int64)
    co, cw = np.asarray(mesh.connec_orig), np.asarray(mesh.connec)
    xyz = np.asarray(mesh.coord)
    node_start = 0
    data = []
    for pt in parts:
        g, loc, e0, e1 = pt["glob"], pt["loc"], pt["e0"], pt["e1"]
        d = dict(
            connecParOrig=F(loc[co[e0:e1]].astype(np.int32)), connecParWork=F(loc[cw[e0:e1]].astype(np.int32)),
            workingNodesPar=np.sort(loc[g[wset[g]]]).astype(np.int32), coordPar=F(xyz[g - 1]),
            globalIdSrl=gid[g - 1].copy(), globalIdPar=np.arange(node_start + 1, node_start + len(g) + 1,
                                                                  dtype=np.int64),
            elemGid=np.arange(e0 + 1, e1 + 1, dtype=np.int32))
        mine = pairs[np.isin(pairs[:, 1], g)] if len(pairs) else pairs
        d["masSla"] = F(np.stack([loc[mine[:, 0]], loc[mine[:, 1]]], axis=1).astype(np.int32)) if len(mine) \
            else np.zeros((0, 2), np.int32, order="F")
        d.update(pt["comm"])
        data.append(d)
        node_start += len(g)
    ne_r = np.array([pt["e1"] - pt["e0"] for pt in parts], np.int32)
    nn_r = np.array([len(pt["glob"]) for pt in parts], np.int64)
    nw_r = np.array([len(d["workingNodesPar"]) for d in data], np.int32)
    np_r = np.array([d["masSla"].shape[0] for d in data], np.int32)
    ncr = np.array([len(d["ranksToComm"]) for d in data], np.int32)
    ncn = np.array([len(d["nodesToComm"]) for d in data], np.int32)
    t = order_tables()
    w = ctypes.c_void_p()
    _call("sod_h5_mesh_create", os.fsencode(path), nranks, 3, _p(ne_r), _p(nn_r), _p(nw_r),
          _p(np_r) if len(pairs) else None, _p(ncr), _p(ncn), *[_p(t[k]) for k in (
              "a2ijk", "gmsh2ijk", "vtk2ijk", "a2ij", "gmsh2ij", "vtk2ij")], ctypes.byref(w))
    try:
        for r, d in enumerate(data):
            _call("sod_h5_mesh_write_rank", w, r, *[_p(d[k]) for k in (
                "connecParOrig", "connecParWork", "workingNodesPar", "coordPar", "masSla", "globalIdSrl",
                "globalIdPar", "elemGid", "ranksToComm", "commsMemPosInLoc", "commsMemSize", "commsMemPosInNgb",
                "nodesToComm")])
    finally:
        rc = lib().sod_h5_mesh_writer_close(w)
    if rc < 0:
        raise IOError_(f"sod_h5_mesh_writer_close: {lib().sod_io_last_error().decode()}")


# ------------------------------------------------------------- restarts

def save_restart(path: str, mesh: Mesh, state: dict, time: float, istep: int, walave_u=None,
                 total_nodes: int | None = None, node_start: int = 1, create: bool = True) -> None:
    """save_hdf5_restartFile: rho, u, pr, E and the element-node mu_e and
    mu_sgs (stored per node as `mue`, `mut`) of the level-2 state."""
    F = np.asfortranarray
    rho, u, pr, E = (F(np.asarray(state[k], np.float64)) for k in ("rho", "u", "pr", "E"))
    mue = F(np.asarray(state["mu_e"], np.float64))
    mut = F(np.asarray(state.get("mu_sgs", np.zeros_like(mue)), np.float64))
    wa = None if walave_u is None else F(np.asarray(walave_u, np.float64))
    total = mesh.npoin if total_nodes is None else total_nodes
    _call("sod_h5_restart_save", os.fsencode(path), int(create), total, node_start, mesh.npoin, mesh.nelem,
          _p(F(np.asarray(mesh.connec_orig, np.int32))), int(istep), float(time), _p(rho), _p(u), _p(pr), _p(E),
          _p(mue), _p(mut), _p(wa))


def load_restart_fields(path: str, mesh: Mesh, masSla: np.ndarray | None = None, node_start: int = 1,
                        walave: bool = False) -> dict:
    """load_hdf5_restartFile: the stored fields of one partition, widened
    from FP32, with the element-node viscosities scattered back and the
    periodic slaves set from their masters."""
    pairs = periodic_pairs(mesh) if masSla is None else np.asfortranarray(masSla, np.int32)
    n, ne = mesh.npoin, mesh.nelem
    out = dict(rho=np.zeros(n), u=np.zeros((n, 3), order="F"), pr=np.zeros(n), E=np.zeros(n),
               mu_e=np.zeros((ne, 64), order="F"), mu_sgs=np.zeros((ne, 64), order="F"))
    wa = np.zeros((n, 3), order="F") if walave else None
    istep = ctypes.c_int(0)
    time = ctypes.c_double(0.0)
    _call("sod_h5_restart_load", os.fsencode(path), node_start, n, ne,
          _p(np.asfortranarray(mesh.connec_orig, np.int32)), pairs.shape[0], _p(pairs), ctypes.byref(istep),
          ctypes.byref(time), *[_p(out[k]) for k in ("rho", "u", "pr", "E", "mu_e", "mu_sgs")], _p(wa))
    out["istep"], out["time"] = istep.value, time.value
    if walave:
        out["walave_u"] = wa
    return out


def state_from_restart(fields: dict, gas: "cases.GasParams", les: bool = False) -> dict:
    """The level-2 state the solvers take, from loaded restart fields:
    eval_vars_after_load_hdf5_resultsFile (CFDSolverBase.f90:2066-2113),
    mu_factor = flag_mu_factor (CFDSolverBase.f90:1081-1095), mu_sgs = 0
    without LES (CFDSolverBase.f90:997-1001) and the initial Sutherland
    viscosity (CFDSolverBase.f90:1059-1079)."""
    F = np.asfortranarray
    rho, u, pr, E = fields["rho"], fields["u"], fields["pr"], fields["E"]
    g = gas.gamma_gas
    q = rho[:, None] * u
    uu = np.sum(u * u, axis=1)
    e_int = E / rho - 0.5 * uu
    Tem = pr / (rho * gas.Rgas)
    csound = np.sqrt(g * pr / rho)
    eta = (rho / (g - 1.0)) * np.log(pr / (rho ** g))
    machno = np.sqrt(uu) / csound
    mu_factor = np.full(rho.shape[0], gas.mu_factor)
    mu_fluid = cases.sutherland(Tem, mu_factor)
    mu_e = F(fields["mu_e"])
    mu_sgs = F(fields["mu_sgs"]) if les else np.zeros_like(mu_e, order="F")
    return dict(rho=F(rho), u=F(u), q=F(q), pr=F(pr), E=F(E), Tem=F(Tem), e_int=F(e_int), eta=F(eta),
                csound=F(csound), machno=F(machno), mu_fluid=F(mu_fluid), mu_factor=F(mu_factor), mu_e=mu_e,
                mu_sgs=mu_sgs, mue_l=np.zeros_like(mu_e, order="F"))


def load_restart(path: str, mesh: Mesh, gas: "cases.GasParams", masSla: np.ndarray | None = None,
                 node_start: int = 1):
    """(state, time, istep) of a restart file for `mesh`."""
    f = load_restart_fields(path, mesh, masSla, node_start)
    return state_from_restart(f, gas), f["time"], f["istep"]
