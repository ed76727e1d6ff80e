"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, as the reference the HIP results are
compared with.  The product package (sod2d_refmap_amd) never imports it.

All arrays use the reference's Fortran layouts (numpy order='F') and its
1-based index values; see oracle/sod2d_oracle.h.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_OMP_LIB_PATH = os.path.join(_HERE, "liboracle_omp.so")
_lib = None
_use_omp = False


def use_openmp(on: bool) -> None:
    """Route every call of this module to liboracle_omp.so (the same source
    built with OpenMP; its scatters are atomic, so its sums are reordered).
    Only bench.py's multi-core CPU baseline does this; the checker is the
    serial liboracle.so."""
    global _lib, _use_omp
    if on != _use_omp:
        _use_omp = on
        _lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("ce", ctypes.c_double),
        ("cmax", ctypes.c_double),
        ("cglob", ctypes.c_double),
        ("c_rho", ctypes.c_double),
        ("c_ener", ctypes.c_double),
        ("Rgas_inf", ctypes.c_double),
        ("flag_normalise_entropy", ctypes.c_int),
        ("flag_total_enthalpy", ctypes.c_int),
        ("flag_real_diff", ctypes.c_int),
        ("flag_diff_suth", ctypes.c_int),
    ]


def lib():
    global _lib
    if _lib is None:
        path = _OMP_LIB_PATH if _use_omp else _LIB_PATH
        if not os.path.exists(path):
            raise RuntimeError(
                f"oracle library missing: {path} (run `make -C oracle` or __graft_entry__.build())")
        _lib = ctypes.CDLL(path)
        _lib.orc_adapt_dt_cfl.restype = ctypes.c_double
        _lib.orc_volAvg_EK.restype = ctypes.c_double
        _lib.orc_invert_jacobian.restype = ctypes.c_double
    return _lib


def default_params(**over) -> OrcParams:
    p = OrcParams()
    lib().orc_default_params(ctypes.byref(p))
    for k, v in over.items():
        setattr(p, k, v)
    return p


def _d(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim <= 1), (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32 and (a.flags.f_contiguous or a.ndim <= 1), (a.dtype, a.flags)
    return a.ctypes.data_as(_ip)


def F(a, dtype=np.float64):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


@dataclass
class Tables:
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


def gll_roots_weights():
    r = np.zeros(4)
    w = np.zeros(4)
    lib().orc_gll_roots_weights(_d(r), _d(w))
    return r, w


def hex64_tables() -> Tables:
    t = Tables(
        atoIJK=np.zeros(64, np.int32), xgp=np.zeros((64, 3), order="F"), wgp=np.zeros(64),
        Ngp=np.zeros((64, 64), order="F"), dNgp=np.zeros((3, 64, 64), order="F"),
        dlxigp_ip=np.zeros((64, 3, 4), order="F"), invAtoIJK=np.zeros((4, 4, 4), np.int32, order="F"),
        gmshAtoI=np.zeros(64, np.int32), gmshAtoJ=np.zeros(64, np.int32), gmshAtoK=np.zeros(64, np.int32))
    lib().orc_hex64_tables(_i(t.atoIJK), _d(t.xgp), _d(t.wgp), _d(t.Ngp), _d(t.dNgp), _d(t.dlxigp_ip),
                           _i(t.invAtoIJK), _i(t.gmshAtoI), _i(t.gmshAtoJ), _i(t.gmshAtoK))
    return t


def invert_jacobian(Je):
    Je = F(Je)
    He = np.zeros((3, 3), order="F")
    det = lib().orc_invert_jacobian(_d(Je), _d(He))
    return det, He


def elem_jacobian(connec, coord, t: Tables):
    nelem, npoin = connec.shape[0], coord.shape[0]
    gpvol = np.zeros((1, 64, nelem), order="F")
    He = np.zeros((3, 3, 64, nelem), order="F")
    lib().orc_elem_jacobian(nelem, npoin, _i(connec), _d(coord), _d(t.dNgp), _d(t.wgp), _d(gpvol), _d(He))
    return gpvol, He


def lumped_mass_spectral(connec, gpvol, npoin):
    Ml = np.zeros(npoin)
    lib().orc_lumped_mass_spectral(connec.shape[0], npoin, _i(connec), _d(gpvol), _d(Ml))
    return Ml


def char_length(connec, coord):
    h = np.zeros(connec.shape[0])
    lib().orc_char_length(connec.shape[0], coord.shape[0], _i(connec), _d(coord), _d(h))
    return h


def char_length_spectral(connec, coord):
    h = np.zeros((connec.shape[0], 64), order="F")
    lib().orc_char_length_spectral(connec.shape[0], coord.shape[0], _i(connec), _d(coord), _d(h))
    return h


def _tab_args(t: Tables):
    return (_d(t.dlxigp_ip), _i(t.invAtoIJK), _i(t.gmshAtoI), _i(t.gmshAtoJ), _i(t.gmshAtoK))


def full_convec_ijk(connec, He, gpvol, t: Tables, u, q, rho, pr, E, total_enthalpy=False):
    nelem, npoin = connec.shape[0], rho.shape[0]
    Rmass = np.zeros(npoin)
    Rmom = np.zeros((npoin, 3), order="F")
    Rener = np.zeros(npoin)
    fn = lib().orc_full_convec_ijk_H if total_enthalpy else lib().orc_full_convec_ijk
    fn(nelem, npoin, _i(connec), _d(He), _d(gpvol), *_tab_args(t), _d(u), _d(q), _d(rho), _d(pr),
       _d(E), _d(Rmass), _d(Rmom), _d(Rener))
    return Rmass, Rmom, Rener


def generic_scalar_convec_ijk(connec, He, gpvol, t: Tables, q, eta, u):
    nelem, npoin = connec.shape[0], eta.shape[0]
    R = np.zeros(npoin)
    lib().orc_generic_scalar_convec_ijk(nelem, npoin, _i(connec), _d(He), _d(gpvol), *_tab_args(t),
                                        _d(q), _d(eta), _d(u), _d(R))
    return R


def full_diffusion_ijk(connec, He, gpvol, t: Tables, Cp, Pr, rho_n, rho, u, Tem, mu_fluid, mu_e,
                       mu_sgs, params: OrcParams | None = None):
    nelem, npoin = connec.shape[0], rho.shape[0]
    prm = params or default_params()
    Rmass = np.zeros(npoin)
    Rmom = np.zeros((npoin, 3), order="F")
    Rener = np.zeros(npoin)
    lib().orc_full_diffusion_ijk(nelem, npoin, _i(connec), _d(He), _d(gpvol), *_tab_args(t),
                                 ctypes.c_double(Cp), ctypes.c_double(Pr), _d(rho_n), _d(rho), _d(u),
                                 _d(Tem), _d(mu_fluid), _d(mu_e), _d(mu_sgs), ctypes.byref(prm),
                                 _d(Rmass), _d(Rmom), _d(Rener))
    return Rmass, Rmom, Rener


def sutherland_viscosity(Tem, mu_factor):
    mu = np.zeros_like(Tem)
    lib().orc_sutherland_viscosity(Tem.shape[0], _d(Tem), _d(mu_factor), _d(mu))
    return mu


def lumped_solver_scal(lpoin_w, Ml, R):
    R = R.copy(order="F")
    lib().orc_lumped_solver_scal(R.shape[0], lpoin_w.shape[0], _i(lpoin_w), _d(Ml), _d(R))
    return R


def lumped_solver_vect(lpoin_w, Ml, R):
    R = R.copy(order="F")
    lib().orc_lumped_solver_vect(R.shape[0], lpoin_w.shape[0], _i(lpoin_w), _d(Ml), _d(R))
    return R


def smart_visc_spectral(connec, lpoin_w, Reta, gpvol, wgp, gamma_gas, rho, u, csound, Tem, eta,
                        helem_l, helem, invAtoIJK, params: OrcParams | None = None, mu_e_in=None):
    nelem, npoin = connec.shape[0], rho.shape[0]
    prm = params or default_params()
    mu_e = np.zeros((nelem, 64), order="F") if mu_e_in is None else F(mu_e_in).copy(order="F")
    mue_l = np.zeros((nelem, 64), order="F")
    lib().orc_smart_visc_spectral(nelem, npoin, lpoin_w.shape[0], _i(connec), _i(lpoin_w), _d(Reta),
                                  _d(gpvol), _d(wgp), ctypes.c_double(gamma_gas), _d(rho), _d(u),
                                  _d(csound), _d(Tem), _d(eta), _d(helem_l), _d(helem), _i(invAtoIJK),
                                  ctypes.byref(prm), _d(mu_e), _d(mue_l))
    return mu_e, mue_l


def adapt_dt_cfl(connec, helem, u, csound, cfl_conv, cfl_diff, mu_fluid, mu_sgs, rho):
    return lib().orc_adapt_dt_cfl(connec.shape[0], rho.shape[0], _i(connec), _d(helem), _d(u), _d(csound),
                                  ctypes.c_double(cfl_conv), ctypes.c_double(cfl_diff), _d(mu_fluid),
                                  _d(mu_sgs), _d(rho))


def volAvg_EK(connec, gpvol, Ngp, rho0, rho, u):
    return lib().orc_volAvg_EK(connec.shape[0], rho.shape[0], _i(connec), _d(gpvol), _d(Ngp),
                               ctypes.c_double(rho0), _d(rho), _d(u))


def rk_4_main(connec, lpoin_w, He, gpvol, t: Tables, Ml, helem, helem_l, dt, Rgas, gamma_gas, Cp, Prt,
              mu_factor, state: dict, params: OrcParams | None = None):
    """One RK4 step in place on the dict of level-2 state arrays:
    rho, u, q, pr, E, Tem, e_int, eta, csound, machno, mu_fluid, mu_e, mu_sgs, mue_l."""
    nelem, npoin = connec.shape[0], state["rho"].shape[0]
    prm = params or default_params()
    s = state
    lib().orc_rk_4_main(nelem, npoin, lpoin_w.shape[0], _i(connec), _i(lpoin_w), _d(He), _d(gpvol),
                        _d(t.dlxigp_ip), _d(t.wgp), _i(t.invAtoIJK), _i(t.gmshAtoI), _i(t.gmshAtoJ),
                        _i(t.gmshAtoK), _d(Ml), _d(helem), _d(helem_l), ctypes.c_double(dt),
                        ctypes.c_double(Rgas), ctypes.c_double(gamma_gas), ctypes.c_double(Cp),
                        ctypes.c_double(Prt), _d(mu_factor), ctypes.byref(prm), _d(s["rho"]), _d(s["u"]),
                        _d(s["q"]), _d(s["pr"]), _d(s["E"]), _d(s["Tem"]), _d(s["e_int"]), _d(s["eta"]),
                        _d(s["csound"]), _d(s["machno"]), _d(s["mu_fluid"]), _d(s["mu_e"]),
                        _d(s["mu_sgs"]), _d(s["mue_l"]))
    return state
