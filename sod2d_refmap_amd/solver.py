"""Host-side driver of the periodic explicit path, mirroring how the
reference's CFDSolverBase / CFDSolverPeriodic set up and call it
(CFDSolverBase.f90:2115-2233 init sequence, :1501-1694 time loop,
CFDSolverPeriodic.f90:38-55): tables -> Jacobians -> lumped mass and element
lengths -> initial viscosity/dt -> rk_4_main per step -> adapt_dt_cfl.

Every computation runs in libsod2d_hip.so through the C ABI; device buffers
are torch tensors holding the reference's column-major arrays.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .cases import GasParams
from .device import empty, init, numerical_params, ptr, set_numerical_params, to_device, to_host, zeros
from .mesh import Mesh, hex64_tables


class PeriodicSolver:
    def __init__(self, mesh: Mesh, gas: GasParams, state: dict, cfl_conv: float = 0.5,
                 cfl_diff: float = 0.5, params=None, device: int = 0):
        init(device)
        self.mesh, self.gas = mesh, gas
        self.cfl_conv, self.cfl_diff = cfl_conv, cfl_diff
        if params is None:
            # the library's current numerical parameters (flags a caller set
            # stay in force), with the case's gas constant
            p = numerical_params()
            _lib.call("sod_get_numerical_params", p)
            p.Rgas_inf = gas.Rgas
        else:
            p = params
        set_numerical_params(p)
        self.params = p
        t = hex64_tables()
        self.tab = {k: to_device(getattr(t, k)) for k in (
            "atoIJK", "xgp", "wgp", "Ngp", "dNgp", "dlxigp_ip", "invAtoIJK", "gmshAtoI",
            "gmshAtoJ", "gmshAtoK")}
        ne, npn = mesh.nelem, mesh.npoin
        self.nelem, self.npoin, self.npoin_w = ne, npn, mesh.npoin_w
        self.connec = to_device(mesh.connec)
        self.connec_orig = to_device(mesh.connec_orig)
        self.lpoin_w = to_device(mesh.lpoin_w)
        self.coord = to_device(mesh.coord)
        # evalJacobians / evalMass / evalCharLength (CFDSolverBase.f90:1280-1389, 783-800)
        self.gpvol = empty((64, ne))
        self.He = empty((3, 3, 64, ne))
        _lib.call("sod_elem_jacobian", ne, npn, ptr(self.connec_orig), ptr(self.coord),
                  ptr(self.tab["dNgp"]), ptr(self.tab["wgp"]), ptr(self.gpvol), ptr(self.He))
        self.Ml = empty(npn)
        _lib.call("sod_lumped_mass_spectral", ne, npn, ptr(self.connec), ptr(self.gpvol), ptr(self.Ml))
        self.helem = empty(ne)
        _lib.call("sod_char_length", ne, npn, ptr(self.connec_orig), ptr(self.coord), ptr(self.helem))
        self.helem_l = empty((ne, 64))
        _lib.call("sod_char_length_spectral", ne, npn, ptr(self.connec_orig), ptr(self.coord),
                  ptr(self.helem_l))
        self.load_state(state)
        self.dt = 0.0
        self.time = 0.0

    # ------------------------------------------------------------ state
    _LEVELS = {"rho": 3, "u": 3, "q": 3, "pr": 3, "E": 3, "Tem": 2, "e_int": 2, "eta": 3}

    def load_state(self, s: dict) -> None:
        npn = self.npoin
        self.lv = {}
        for k, nl in self._LEVELS.items():
            a = np.asarray(s[k])
            comp = 3 if a.ndim == 2 and a.shape[1] == 3 else 1
            full = np.repeat(a.reshape(npn, comp, order="F")[:, :, None], nl, axis=2)
            self.lv[k] = to_device(np.asfortranarray(full))
        for k in ("csound", "machno", "mu_fluid", "mu_factor", "mu_e", "mu_sgs", "mue_l"):
            setattr(self, k, to_device(s[k]))
        self.kres = zeros(npn)
        self.etot = zeros(npn)
        self.au = zeros((npn, 3))
        self.ax = [zeros(npn) for _ in range(3)]

    def level(self, name: str, lvl: int = 2) -> torch.Tensor:
        """View of time level `lvl` (1-based, as in the reference) of a state array."""
        t = self.lv[name]
        nl = self._LEVELS[name]
        per = t.numel() // nl
        return t[(lvl - 1) * per: lvl * per]

    def state(self) -> dict:
        npn, ne = self.npoin, self.nelem
        out = {}
        for k in self._LEVELS:
            per = self.lv[k].numel() // self._LEVELS[k]
            comp = per // npn
            shape = (npn, comp) if comp > 1 else (npn,)
            out[k] = to_host(self.level(k), shape)
        for k in ("csound", "machno", "mu_fluid"):
            out[k] = to_host(getattr(self, k), (npn,))
        for k in ("mu_e", "mue_l"):
            out[k] = to_host(getattr(self, k), (ne, 64))
        return out

    # ------------------------------------------------------------ path
    def rk_4_main(self, dt: float) -> None:
        T = self.tab
        g = self.gas
        _lib.call(
            "sod_rk_4_main", 1, 0, self.nelem, 0, self.npoin, self.npoin_w, 0, None, None,
            ptr(T["dlxigp_ip"]), ptr(T["xgp"]), ptr(T["atoIJK"]), ptr(T["invAtoIJK"]),
            ptr(T["gmshAtoI"]), ptr(T["gmshAtoJ"]), ptr(T["gmshAtoK"]), 1, ptr(self.connec),
            ptr(T["Ngp"]), ptr(T["dNgp"]), ptr(self.coord), ptr(T["wgp"]), ptr(self.He), ptr(self.Ml),
            ptr(self.gpvol), dt, ptr(self.helem), ptr(self.helem_l), g.Rgas, g.gamma_gas, g.Cp, g.Prt,
            ptr(self.lv["rho"]), ptr(self.lv["u"]), ptr(self.lv["q"]), ptr(self.lv["pr"]),
            ptr(self.lv["E"]), ptr(self.lv["Tem"]), ptr(self.csound), ptr(self.machno),
            ptr(self.lv["e_int"]), ptr(self.lv["eta"]), ptr(self.mu_e), ptr(self.mu_sgs), ptr(self.kres),
            ptr(self.etot), ptr(self.au), ptr(self.ax[0]), ptr(self.ax[1]), ptr(self.ax[2]),
            ptr(self.lpoin_w), ptr(self.mu_fluid), ptr(self.mu_factor), ptr(self.mue_l))

    def adapt_dt(self) -> float:
        """adapt_dt_cfl with both CFL terms (CFDSolverBase.f90:1132-1141)."""
        import ctypes
        dt = ctypes.c_double(self.dt)
        cd = ctypes.c_double(self.cfl_diff)
        _lib.call("sod_adapt_dt_cfl", self.nelem, self.npoin, ptr(self.connec), ptr(self.helem),
                  ptr(self.level("u")), ptr(self.csound), self.cfl_conv, ctypes.addressof(dt),
                  ctypes.addressof(cd), ptr(self.mu_fluid), ptr(self.mu_sgs), ptr(self.level("rho")))
        self.dt = dt.value
        return self.dt

    def volAvg_EK(self) -> float:
        import ctypes
        ek = ctypes.c_double(0.0)
        _lib.call("sod_volAvg_EK", self.nelem, self.npoin, ptr(self.connec), ptr(self.gpvol),
                  ptr(self.tab["Ngp"]), self.gas.rho0, ptr(self.level("rho")), ptr(self.level("u")),
                  ctypes.addressof(ek))
        return ek.value

    # ------------------------------------------------- device time loop
    def time_loop(self):
        """The device-resident time loop over this solver's arrays
        (sod_time_loop_*), created on first use."""
        import ctypes
        if getattr(self, "_loop", None) is None:
            g = self.gas
            h = ctypes.c_void_p()
            _lib.call("sod_time_loop_create", ctypes.byref(h), self.nelem, self.npoin, self.npoin_w,
                      ptr(self.connec), ptr(self.lpoin_w), ptr(self.He), ptr(self.gpvol), ptr(self.Ml),
                      ptr(self.helem), ptr(self.helem_l), ptr(self.mu_factor), g.Rgas, g.gamma_gas, g.Cp, g.Prt,
                      ptr(self.lv["rho"]), ptr(self.lv["u"]), ptr(self.lv["q"]), ptr(self.lv["pr"]),
                      ptr(self.lv["E"]), ptr(self.lv["Tem"]), ptr(self.csound), ptr(self.machno),
                      ptr(self.lv["e_int"]), ptr(self.lv["eta"]), ptr(self.mu_e), ptr(self.mu_sgs),
                      ptr(self.mu_fluid), ptr(self.mue_l))
            self._loop = h
        return self._loop

    def run_device(self, nsteps: int, istep0: int = 1, log_first: int = 1, log_every: int = 0,
                   fixed_dt: bool = False, keep_levels: bool = True):
        """nsteps of the reference's time loop on the device (dt from the CFL
        condition after each step unless fixed_dt).  Returns the EK values of
        the log steps."""
        import ctypes
        L = self.time_loop()
        if self.dt == 0.0 and not fixed_dt:
            self.adapt_dt()
        nlog = 0
        if log_every > 0:
            nlog = sum(1 for i in range(istep0, istep0 + nsteps) if i >= log_first and (i - log_first) % log_every == 0)
        ek = (ctypes.c_double * max(1, nlog))()
        dt = ctypes.c_double(self.dt)
        tm = ctypes.c_double(self.time)
        _lib.call("sod_time_loop_run", L, istep0, nsteps, self.cfl_conv, self.cfl_diff, int(fixed_dt),
                  ctypes.byref(dt), ctypes.byref(tm), log_first, log_every, self.gas.rho0, ek, max(1, nlog),
                  int(keep_levels))
        self.dt, self.time = dt.value, tm.value
        return [ek[i] for i in range(nlog)]

    def close(self) -> None:
        if getattr(self, "_loop", None) is not None:
            _lib.call("sod_time_loop_destroy", self._loop)
            self._loop = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps: int) -> None:
        """evalTimeIteration's step sequence for the periodic case."""
        if self.dt == 0.0:
            self.adapt_dt()
        for _ in range(nsteps):
            self.rk_4_main(self.dt)
            self.time += self.dt
            self.adapt_dt()
