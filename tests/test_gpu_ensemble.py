"""A batched ensemble of independent boxes (sod_time_loop_create_batched),
member by member and step by step against the oracle (rk_4_main, then
adapt_dt_cfl with the member's own CFL minimum), and against the same members
stepped one after the other through the plain per-instance entry points."""
import numpy as np
import pytest

import oracle
from _util import mesh_geometry, rel_l2
from sod2d_refmap_amd import _lib, cases, mesh as pmesh
from sod2d_refmap_amd.device import numerical_params, set_numerical_params

pytestmark = pytest.mark.gpu

KEYS = ("rho", "q", "E", "u", "pr", "Tem", "eta", "csound", "machno", "mu_fluid", "mu_e", "mue_l")
T = oracle.hex64_tables()


def _members():
    # (n, curved_amp, perm_seed, Re, Mach, perturb_seed, perturb_amp). The
    # 2^3 box at Mach 0.2 takes a 1% perturbation: with 5% the scheme itself
    # (the oracle included) diverges there on the second step.
    specs = [(2, 0.0, None, 800.0, 0.2, 1, 1e-2), (3, 0.05, 9, 1600.0, 0.1, 2, 5e-2),
             (2, 0.0, None, 3000.0, 0.05, 3, 5e-2)]
    out = []
    for n, amp, perm, Re, Ma, seed, pamp in specs:
        m = pmesh.periodic_box(n, curved_amp=amp, perm_seed=perm)
        gas = cases.GasParams(Re=Re, Mach=Ma)
        out.append((m, gas, cases.tgv_initial_state(m, gas, perturb_seed=seed, perturb_amp=pamp)))
    return out


class _OracleMember:
    """One member stepped by the oracle as the reference's time loop does
    (rk_4_main, time += dt, adapt_dt_cfl; CFDSolverBase.f90:1501-1694)."""

    def __init__(self, m, gas, s, params):
        self.m, self.gas, self.p = m, gas, params
        self.g = mesh_geometry(m)
        self.s = {k: v.copy(order="F") for k, v in s.items()}
        self.time = 0.0
        self.dt = self._dt()

    def _dt(self):
        m, s = self.m, self.s
        return oracle.adapt_dt_cfl(m.connec, self.g["helem"], s["u"], s["csound"], 0.5, 0.5, s["mu_fluid"],
                                   np.zeros((m.nelem, 64), order="F"), s["rho"])

    def step(self):
        m, g, gas = self.m, self.g, self.gas
        oracle.rk_4_main(m.connec, m.lpoin_w, g["He"], g["gpvol"], T, g["Ml"], g["helem"], g["helem_l"], self.dt,
                         gas.Rgas, gas.gamma_gas, gas.Cp, gas.Prt, self.s["mu_factor"], self.s, self.p)
        self.time += self.dt
        self.dt = self._dt()


@pytest.mark.parametrize("flags", [{}, {"flag_total_enthalpy": 1}], ids=["default", "total_enthalpy"])
def test_ensemble_matches_oracle_and_members(gpu, flags):
    from sod2d_refmap_amd.ensemble import Ensemble, sequential_solvers, sequential_step
    members = _members()
    gas0 = members[0][1]
    set_numerical_params(numerical_params(Rgas_inf=gas0.Rgas, **flags))
    E = Ensemble(members)
    sols = sequential_solvers(members)
    orc = [_OracleMember(m, g, s, oracle.default_params(Rgas_inf=g.Rgas, **flags)) for m, g, s in members]
    for i, (s, o) in enumerate(zip(sols, orc)):
        assert E.dt[i] == s.dt == o.dt
    for nsteps in (2, 1):
        print(f"[ensemble {flags}] {len(members)} members, {nsteps} steps", flush=True)
        E.run(nsteps)
        for _ in range(nsteps):
            sequential_step(sols)
            for o in orc:
                o.step()
        for i, (s, o) in enumerate(zip(sols, orc)):
            got, seq = E.member_state(i), s.state()
            for k in KEYS:
                assert np.isfinite(got[k]).all(), (i, k)
                err = rel_l2(got[k], o.s[k])
                assert err < 1e-10, ("oracle", i, k, err)
                err = rel_l2(got[k], seq[k])
                assert err < 1e-12, ("per-instance", i, k, err)
            assert abs(E.dt[i] / o.dt - 1) < 1e-12, (i, E.dt[i], o.dt)
            assert abs(E.time[i] / o.time - 1) < 1e-13, (i, E.time[i], o.time)
            assert abs(E.dt[i] / s.dt - 1) < 1e-12 and abs(E.time[i] / s.time - 1) < 1e-12
    E.close()
    _lib.call("sod_synchronize")
