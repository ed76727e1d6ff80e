"""The device time loop (sod_time_loop_*): a graph-captured step replayed
with dt, time, step counter and EK history on the device, step by step
against the oracle (rk_4_main, adapt_dt_cfl, volAvg_EK), and against the
same steps taken through the eager entry points (sod_rk_4_main,
sod_adapt_dt_cfl, sod_volAvg_EK).  The graph is replayed again after eager
work on the same arrays, and every replay is checked."""
import numpy as np
import pytest

import oracle
from _util import mesh_geometry, rel_l2
from sod2d_refmap_amd import cases, mesh as pmesh
from sod2d_refmap_amd.device import to_host

pytestmark = pytest.mark.gpu
T = oracle.hex64_tables()

KEYS = ("rho", "q", "E", "u", "pr", "Tem", "eta", "csound", "machno", "mu_fluid", "mu_e", "mue_l")


def _eager(sol, nsteps, istep0, log_first, log_every):
    eks = []
    before = None
    for istep in range(istep0, istep0 + nsteps):
        before = sol.state()
        sol.rk_4_main(sol.dt)
        sol.time += sol.dt
        if log_every > 0 and istep >= log_first and (istep - log_first) % log_every == 0:
            eks.append(sol.volAvg_EK())
        sol.adapt_dt()
    return eks, before


def test_time_loop_matches_eager_steps(gpu):
    from sod2d_refmap_amd.solver import PeriodicSolver
    m = pmesh.periodic_box(3, curved_amp=0.05, perm_seed=13)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas, perturb_seed=5, perturb_amp=5e-2)
    A = PeriodicSolver(m, gas, s)
    B = PeriodicSolver(m, gas, s)
    A.adapt_dt()
    B.adapt_dt()
    assert A.dt == B.dt
    istep = 1
    for nsteps in (3, 2, 1):
        print(f"[time loop] steps {istep}..{istep + nsteps - 1}", flush=True)
        ek_a, before = _eager(A, nsteps, istep, log_first=2, log_every=2)
        ek_b = B.run_device(nsteps, istep0=istep, log_first=2, log_every=2, keep_levels=True)
        sa, sb = A.state(), B.state()
        for k in KEYS:
            err = rel_l2(sb[k], sa[k])
            assert err < 1e-12, (nsteps, k, err)
        assert abs(B.dt / A.dt - 1) < 1e-12 and abs(B.time / A.time - 1) < 1e-12
        assert len(ek_a) == len(ek_b)
        for x, y in zip(ek_a, ek_b):
            assert abs(y / x - 1) < 1e-12
        # level 1 holds the state at the start of the last step, level 3 a copy of it
        npn = m.npoin
        for k, comp in (("rho", 1), ("q", 3), ("E", 1), ("eta", 1), ("u", 3), ("pr", 1)):
            shape = (npn, comp) if comp > 1 else (npn,)
            l1 = to_host(B.level(k, 1), shape)
            l3 = to_host(B.level(k, 3), shape)
            assert np.array_equal(l1, l3), k
            assert rel_l2(l1, before[k]) < 1e-12, k
        istep += nsteps
    B.close()


def test_time_loop_fixed_dt_without_levels(gpu):
    from sod2d_refmap_amd.solver import PeriodicSolver
    m = pmesh.periodic_box(2)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas)
    A = PeriodicSolver(m, gas, s)
    B = PeriodicSolver(m, gas, s)
    dt = A.adapt_dt() * 0.5
    A.dt = B.dt = dt
    for _ in range(4):
        A.rk_4_main(dt)
    ek = B.run_device(4, fixed_dt=True, keep_levels=False)
    assert ek == [] and B.dt == dt
    sa, sb = A.state(), B.state()
    for k in KEYS:
        assert rel_l2(sb[k], sa[k]) < 1e-12, k
    B.close()


def test_time_loop_vs_oracle(gpu):
    """Each replay of the captured step against the oracle's step of the
    reference's loop (CFDSolverBase.f90:1501-1694): state, dt, time and EK."""
    from sod2d_refmap_amd.solver import PeriodicSolver
    m = pmesh.periodic_box(3, curved_amp=0.05, perm_seed=13)
    g = mesh_geometry(m)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas, perturb_seed=5, perturb_amp=5e-2)
    p = oracle.default_params(Rgas_inf=gas.Rgas)
    zero = np.zeros((m.nelem, 64), order="F")
    ref = {k: v.copy(order="F") for k, v in s.items()}

    def dt_of(st):
        return oracle.adapt_dt_cfl(m.connec, g["helem"], st["u"], st["csound"], 0.5, 0.5, st["mu_fluid"], zero,
                                   st["rho"])

    sol = PeriodicSolver(m, gas, s)
    dt = dt_of(ref)
    assert sol.adapt_dt() == dt
    time = 0.0
    for istep in range(1, 7):
        eks = sol.run_device(1, istep0=istep, log_first=1, log_every=1, keep_levels=True)
        oracle.rk_4_main(m.connec, m.lpoin_w, g["He"], g["gpvol"], T, g["Ml"], g["helem"], g["helem_l"], dt,
                         gas.Rgas, gas.gamma_gas, gas.Cp, gas.Prt, ref["mu_factor"], ref, p)
        time += dt
        ek = oracle.volAvg_EK(m.connec, g["gpvol"], T.Ngp, gas.rho0, ref["rho"], ref["u"])
        dt = dt_of(ref)
        out = sol.state()
        for k in KEYS:
            err = rel_l2(out[k], ref[k])
            assert err < (1e-12 if istep == 1 else 1e-10), (istep, k, err)
        assert abs(eks[0] / ek - 1) < 1e-12, (istep, eks[0], ek)
        assert abs(sol.dt / dt - 1) < 1e-12, (istep, sol.dt, dt)
        assert abs(sol.time / time - 1) < 1e-13, istep
    sol.close()
