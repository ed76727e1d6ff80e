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
    x0: np.ndarray = field(repr=False, default=None)  # (npoin,3) unperturbed coordinates

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
    t = hex64_tables()
    # lattice offsets (0..3) of each gmsh local node
    li = _S2LAT[t.gmshAtoI - 1]
    lj = _S2LAT[t.gmshAtoJ - 1]
    lk = _S2LAT[t.gmshAtoK - 1]
    nelem = nx * ny * nz
    elem_lex = np.arange(nelem, dtype=np.int64)
    if perm_seed is not None:
        elem_lex = np.random.default_rng(perm_seed).permutation(nelem).astype(np.int64)
    ex = elem_lex % nx
    ey = (elem_lex // nx) % ny
    ez = elem_lex // (nx * ny)
    Mx, My, Mz = 3 * nx + 1, 3 * ny + 1, 3 * nz + 1
    I = 3 * ex[:, None] + li[None, :]
    J = 3 * ey[:, None] + lj[None, :]
    K = 3 * ez[:, None] + lk[None, :]
    lat_id = (I + Mx * (J + My * K)).astype(np.int64)          # (nelem, 64)
    # first touch in (element, gmsh local node) order
    flat = lat_id.ravel()
    uniq, first = np.unique(flat, return_index=True)
    order = np.argsort(first, kind="stable")
    new_of_lat = np.empty(Mx * My * Mz, dtype=np.int64)
    new_of_lat[uniq[order]] = np.arange(uniq.size)
    npoin = uniq.size
    assert npoin == Mx * My * Mz
    lat_of_new = uniq[order]
    Ilat = lat_of_new % Mx
    Jlat = (lat_of_new // Mx) % My
    Klat = lat_of_new // (Mx * My)
    node_lattice = np.stack([Ilat, Jlat, Klat], axis=1)
    connec_orig = np.asfortranarray((new_of_lat[lat_id] + 1).astype(np.int32))
    # periodic masters: lattice index modulo 3n
    Im, Jm, Km = I % (3 * nx), J % (3 * ny), K % (3 * nz)
    connec = np.asfortranarray((new_of_lat[Im + Mx * (Jm + My * Km)] + 1).astype(np.int32))
    slave = (Ilat == 3 * nx) | (Jlat == 3 * ny) | (Klat == 3 * nz)
    lpoin_w = (np.nonzero(~slave)[0] + 1).astype(np.int32)
    # coordinates at FP32-rounded GLL positions
    r = t.gll_roots

    def axis_coord(idx, nel, length):
        h = length / nel
        e = np.minimum(idx // 3, nel - 1)
        loc = idx - 3 * e
        return e * h + (1.0 + r[loc]) * (0.5 * h)

    x = axis_coord(Ilat, nx, Lx)
    y = axis_coord(Jlat, ny, Ly)
    z = axis_coord(Klat, nz, Lz)
    x0 = np.stack([x, y, z], axis=1)
    if curved_amp != 0.0:
        hx, hy, hz = Lx / nx, Ly / ny, Lz / nz
        sx, sy, sz = np.sin(2 * np.pi * x / Lx), np.sin(2 * np.pi * y / Ly), np.sin(2 * np.pi * z / Lz)
        x = x + curved_amp * hx * sy * sz
        y = y + curved_amp * hy * sz * sx
        z = z + curved_amp * hz * sx * sy
    coord = np.asfortranarray(np.stack([x, y, z], axis=1))
    return Mesh(n=(nx, ny, nz), L=(Lx, Ly, Lz), nelem=nelem, npoin=int(npoin),
                npoin_w=int(lpoin_w.size), connec_orig=connec_orig, connec=connec,
                lpoin_w=lpoin_w, coord=coord, elem_lex=elem_lex, node_lattice=node_lattice,
                x0=np.asfortranarray(x0))
