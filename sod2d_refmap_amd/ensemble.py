"""Batched ensembles: many independent periodic boxes (uneven sizes, their
own Reynolds and Mach numbers and initial perturbations) advanced together
as one flattened element table by the device time loop
(sod_time_loop_create_batched / sod_time_loop_run_batched): one graph per
step, one launch per kernel for the whole ensemble, each instance with its
own CFL time step and entropy normalisation.

`ensemble_members` draws the members of the seeded TGV ensemble;
`Ensemble` builds and steps them; `sequential_step` is the baseline of
per-instance plain launches (sod_rk_4_main + sod_adapt_dt_cfl per member).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib, cases, mesh as pmesh
from .device import ptr, to_device, to_host
from .solver import PeriodicSolver

SIZES = (8, 12, 16, 24, 32)


def ensemble_members(count=48, sizes=SIZES, size_seed=4, param_seed=5, mode_seed=6, amp=1e-2, curved_amp=0.0):
    """(mesh, gas, state) of each member: n^3 elements with n drawn from
    `sizes`, Re ~ U[400, 3200], Mach ~ U[0.05, 0.3], Taylor-Green plus a
    seeded Fourier perturbation of amplitude `amp`."""
    rs = np.random.default_rng(size_seed)
    rp = np.random.default_rng(param_seed)
    ns = rs.choice(np.asarray(sizes), size=count)
    out = []
    for i, n in enumerate(ns):
        gas = cases.GasParams(Re=float(rp.uniform(400.0, 3200.0)), Mach=float(rp.uniform(0.05, 0.3)))
        m = pmesh.periodic_box(int(n), curved_amp=curved_amp)
        s = cases.tgv_initial_state(m, gas, perturb_seed=mode_seed * 1000 + i, perturb_amp=amp)
        out.append((m, gas, s))
    return out


def _cat_rows(arrs):
    """concatenate (n_i, ...) Fortran arrays along their first (fastest) axis"""
    return np.asfortranarray(np.concatenate([np.asarray(a) for a in arrs], axis=0))


class Ensemble:
    """The members stepped together on the device."""

    _LEVELS = PeriodicSolver._LEVELS

    def __init__(self, members, cfl_conv: float = 0.5, cfl_diff: float = 0.5):
        gas0 = members[0][1]
        for _, g, _ in members:
            if (g.gamma_gas, g.Cp, g.Prt, g.Rgas) != (gas0.gamma_gas, gas0.Cp, gas0.Prt, gas0.Rgas):
                raise ValueError("ensemble members must share gamma, Cp, Prt and Rgas")
        self.gas = gas0
        self.cfl_conv, self.cfl_diff = cfl_conv, cfl_diff
        self.n = len(members)
        eoff, noff, woff = [0], [0], [0]
        geo = {k: [] for k in ("He", "gpvol", "Ml", "helem", "helem_l")}
        conn, lw, dts = [], [], []
        st = {k: [] for k in list(self._LEVELS) + ["csound", "machno", "mu_fluid", "mu_factor", "mu_e", "mu_sgs",
                                                   "mue_l"]}
        for m, g, s in members:
            sol = PeriodicSolver(m, g, s, cfl_conv, cfl_diff)
            dts.append(sol.adapt_dt())
            ne, npn = m.nelem, m.npoin
            geo["He"].append(to_host(sol.He, (3, 3, 64, ne)))
            geo["gpvol"].append(to_host(sol.gpvol, (1, 64, ne)))
            geo["Ml"].append(to_host(sol.Ml, (npn,)))
            geo["helem"].append(to_host(sol.helem, (ne,)))
            geo["helem_l"].append(to_host(sol.helem_l, (ne, 64)))
            sol.close()
            del sol
            conn.append(np.asarray(m.connec) + noff[-1])
            lw.append(np.asarray(m.lpoin_w) + noff[-1])
            for k in st:
                st[k].append(np.asarray(s[k]))
            eoff.append(eoff[-1] + ne)
            noff.append(noff[-1] + npn)
            woff.append(woff[-1] + m.npoin_w)
        self.eoff, self.noff, self.woff = eoff, noff, woff
        self.nelem, self.npoin, self.npoin_w = eoff[-1], noff[-1], woff[-1]
        self.dt = np.asarray(dts, dtype=np.float64)
        self.time = np.zeros(self.n)
        self.connec = to_device(_cat_rows(conn).astype(np.int32))
        self.lpoin_w = to_device(np.concatenate(lw).astype(np.int32))
        self.He = to_device(np.concatenate(geo["He"], axis=3))
        self.gpvol = to_device(np.concatenate(geo["gpvol"], axis=2))
        self.Ml = to_device(np.concatenate(geo["Ml"]))
        self.helem = to_device(np.concatenate(geo["helem"]))
        self.helem_l = to_device(_cat_rows(geo["helem_l"]))
        npn = self.npoin
        self.lv = {}
        for k, nl in self._LEVELS.items():
            a = _cat_rows(st[k])
            comp = 3 if a.ndim == 2 and a.shape[1] == 3 else 1
            full = np.repeat(a.reshape(npn, comp, order="F")[:, :, None], nl, axis=2)
            self.lv[k] = to_device(np.asfortranarray(full))
        for k in ("csound", "machno", "mu_fluid", "mu_factor"):
            setattr(self, k, to_device(np.concatenate(st[k])))
        for k in ("mu_e", "mu_sgs", "mue_l"):
            setattr(self, k, to_device(_cat_rows(st[k])))
        g = self.gas
        h = ctypes.c_void_p()
        off = lambda v: (ctypes.c_int * len(v))(*v)
        self._offs = (off(eoff), off(noff), off(woff))
        _lib.call("sod_time_loop_create_batched", ctypes.byref(h), self.n, *self._offs, self.nelem, npn,
                  self.npoin_w, ptr(self.connec), ptr(self.lpoin_w), ptr(self.He), ptr(self.gpvol), ptr(self.Ml),
                  ptr(self.helem), ptr(self.helem_l), ptr(self.mu_factor), g.Rgas, g.gamma_gas, g.Cp, g.Prt,
                  ptr(self.lv["rho"]), ptr(self.lv["u"]), ptr(self.lv["q"]), ptr(self.lv["pr"]), ptr(self.lv["E"]),
                  ptr(self.lv["Tem"]), ptr(self.csound), ptr(self.machno), ptr(self.lv["e_int"]),
                  ptr(self.lv["eta"]), ptr(self.mu_e), ptr(self.mu_sgs), ptr(self.mu_fluid), ptr(self.mue_l))
        self._loop = h

    def run(self, nsteps: int, keep_levels: bool = True) -> None:
        dt = self.dt.copy()
        tm = self.time.copy()
        _lib.call("sod_time_loop_run_batched", self._loop, nsteps, self.cfl_conv, self.cfl_diff, 0,
                  dt.ctypes.data, tm.ctypes.data, int(keep_levels))
        self.dt, self.time = dt, tm

    def stage_config(self) -> dict:
        """The stage configuration the ensemble's loop runs with (_lib.stage_config)."""
        if self._loop is None:
            raise RuntimeError("the ensemble has not been run yet")
        return _lib.stage_config(self._loop)

    def member_state(self, i: int) -> dict:
        """Level-2 state arrays and element-node viscosities of member i."""
        n0, n1 = self.noff[i], self.noff[i + 1]
        e0, e1 = self.eoff[i], self.eoff[i + 1]
        npn, ne = self.npoin, self.nelem
        out = {}
        for k, nl in self._LEVELS.items():
            per = self.lv[k].numel() // nl
            comp = per // npn
            lvl2 = to_host(self.lv[k][per: 2 * per], (npn, comp) if comp > 1 else (npn,))
            out[k] = np.asfortranarray(lvl2[n0:n1])
        for k in ("csound", "machno", "mu_fluid"):
            out[k] = to_host(getattr(self, k), (npn,))[n0:n1].copy()
        for k in ("mu_e", "mue_l"):
            out[k] = np.asfortranarray(to_host(getattr(self, k), (ne, 64))[e0:e1])
        return out

    def dofs(self) -> int:
        return 5 * self.npoin_w

    def close(self) -> None:
        if getattr(self, "_loop", None) is not None:
            _lib.call("sod_time_loop_destroy", self._loop)
            self._loop = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sequential_solvers(members, cfl_conv: float = 0.5, cfl_diff: float = 0.5):
    sols = []
    for m, g, s in members:
        sol = PeriodicSolver(m, g, s, cfl_conv, cfl_diff)
        sol.adapt_dt()
        sols.append(sol)
    return sols


def sequential_step(sols) -> None:
    """One step of every member through the plain per-instance entry points
    (sod_rk_4_main, then sod_adapt_dt_cfl), one member after the other."""
    for sol in sols:
        sol.rk_4_main(sol.dt)
        sol.time += sol.dt
        sol.adapt_dt()
