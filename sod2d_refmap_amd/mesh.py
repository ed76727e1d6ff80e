"""Synthetic periodic HEX64 meshes in SOD2D's partitioned-mesh data model.

The reference reads its meshes from the partitioned HDF5 files written by its
mesh conversion tool.  This module builds the same arrays for a structured
periodic box directly:

* element-local node order is gmsh's high-order order (a2ijk = gmsh2ijk,
  mod_meshConversorTool.f90:150-152);
* high-order nodes sit at the FP32-rounded GLL positions of each element
  (interpolateOriginalCoordinates, mod_meshConversorTool.f90:3921-3942, maps
  gmsh's equispaced nodes onto the GLL grid);
* rank-local node ids are given by first touch in element order, local nodes
  in gmsh order (reorder_nodes_in_mshRank, mod_meshConversorTool.f90:2864-3076);
* periodic slave nodes (on the x=L, y=L or z=L faces) are replaced by their
  masters in connecParWork, and the working-node list is every non-slave id in
  increasing order (create_working_lists_parallel,
  mod_meshConversorTool.f90:3944-4035).

Arrays follow the reference's Fortran layout: numpy order='F', 1-based ids.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib


@dataclass
class Tables:
    """p=3 tables (sod_hex64_tables)."""
    atoIJK: np.ndarray
    xgp: np.ndarray
    wgp: np.ndarray
    Ngp: np.ndarray
    dNgp: np.ndarray
    dlxigp_ip: np.ndarray
    invAtoIJK: np.ndarray
    gmshAtoI: np.ndarray
    gmshAtoJ: np.ndarray
    gmshAtoK: np.ndarray

    @property
    def gll_roots(self) -> np.ndarray:
        """GLL nodes in increasing order (FP32-rounded, stored as FP64)."""
        return np.sort(np.unique(self.xgp[:, 0]))


_TABLES: Tables | None = None


def hex64_tables() -> Tables:
    global _TABLES
    if _TABLES is None:
        t = Tables(
            atoIJK=np.zeros(64, np.int32), xgp=np.zeros((64, 3), order="F"), wgp=np.zeros(64),
            Ngp=np.zeros((64, 64), order="F"), dNgp=np.zeros((3, 64, 64), order="F"),
            dlxigp_ip=np.zeros((64, 3, 4), order="F"),
            invAtoIJK=np.zeros((4, 4, 4), np.int32, order="F"),
            gmshAtoI=np.zeros(64, np.int32), gmshAtoJ=np.zeros(64, np.int32),
            gmshAtoK=np.zeros(64, np.int32))
        _lib.call("sod_hex64_tables", *[a.ctypes.data for a in (
            t.atoIJK, t.xgp, t.wgp, t.Ngp, t.dNgp, t.dlxigp_ip, t.invAtoIJK, t.gmshAtoI,
            t.gmshAtoJ, t.gmshAtoK)])
        _TABLES = t
    return _TABLES


# sod2d 1-D index (0-based) -> geometric lattice offset 0..3
_S2LAT = np.array([0, 3, 1, 2])


@dataclass
class Mesh:
    n: tuple                 # elements per direction (nx, ny, nz)
    L: tuple                 # box lengths
    nelem: int
    npoin: int               # N_tot, includes periodic slave copies
    npoin_w: int             # working (non-slave) nodes
    connec_orig: np.ndarray  # (nelem,64) int32 F, 1-based, with slave ids
    connec: np.ndarray       # (nelem,64) int32 F, 1-based, masters only (connecParWork)
    lpoin_w: np.ndarray      # (npoin_w,) int32, 1-based, increasing
    coord: np.ndarray        # (npoin,3) float64 F
    elem_lex: np.ndarray     # (nelem,) lexicographic element of each element id
    node_lattice: np.ndarray = field(repr=False)  # (npoin,3) int lattice index of each node

    @property
    def is_slave(self) -> np.ndarray:
        M = np.array([3 * k for k in self.n])
        return np.any(self.node_lattice == M[None, :], axis=1)


def periodic_box(n, L=2.0 * np.pi, curved_amp: float = 0.0, perm_seed: int | None = None) -> Mesh:
    """Periodic box of n^3 (or nx*ny*nz) HEX64 elements.

    curved_amp: if non-zero, smooth node mapping
        x_i += curved_amp * h * sin(2 pi x_j / L) sin(2 pi x_k / L)   (i,j,k cyclic)
    perm_seed: if given, element numbering is permuted by a seeded permutation
        before nodes are numbered (numbering then follows the permuted order).
    """
    nx, ny, nz = (n, n, n) if np.isscalar(n) else tuple(n)
    Lx, Ly, Lz = (L, L, L) if np.isscalar(L) else tuple(L)
    nelem = nx * ny * nz
    elem_lex = np.arange(nelem, dtype=np.int64)
    if perm_seed is not None:
        elem_lex = np.random.default_rng(perm_seed).permutation(nelem).astype(np.int64)
    npoin = (3 * nx + 1) * (3 * ny + 1) * (3 * nz + 1)
    npoin_w = 27 * nelem
    connec_orig = np.empty((nelem, 64), np.int32, order="F")
    connec = np.empty((nelem, 64), np.int32, order="F")
    lpoin_w = np.empty(npoin_w, np.int32)
    coord = np.empty((npoin, 3), order="F")
    node_lattice = np.empty((npoin, 3), np.int32)
    _lib.call("sod_structured_box_mesh", nx, ny, nz, float(Lx), float(Ly), float(Lz), float(curved_amp),
              elem_lex.ctypes.data, connec_orig.ctypes.data, connec.ctypes.data, lpoin_w.ctypes.data,
              coord.ctypes.data, node_lattice.ctypes.data)
    return Mesh(n=(nx, ny, nz), L=(Lx, Ly, Lz), nelem=nelem, npoin=npoin, npoin_w=npoin_w,
                connec_orig=connec_orig, connec=connec, lpoin_w=lpoin_w, coord=coord,
                elem_lex=elem_lex, node_lattice=node_lattice)


def lattice_coordinates(m: Mesh) -> np.ndarray:
    """Unperturbed (straight-sided) coordinates of every node, (npoin,3)."""
    r = hex64_tables().gll_roots
    out = np.empty((m.npoin, 3))
    for d, (nel, length) in enumerate(zip(m.n, m.L)):
        idx = m.node_lattice[:, d].astype(np.int64)
        h = length / nel
        e = np.minimum(idx // 3, nel - 1)
        out[:, d] = e * h + (1.0 + r[idx - 3 * e]) * (0.5 * h)
    return out
