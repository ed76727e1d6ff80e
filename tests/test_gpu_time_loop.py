"""The device time loop (sod_time_loop_*): a graph-captured step replayed
with dt, time, step counter and EK history on the device, against the same
steps taken through the eager entry points (sod_rk_4_main, sod_adapt_dt_cfl,
sod_volAvg_EK), whose parity with the oracle is tested elsewhere.  The graph
is replayed again after eager work on the same arrays, and every replay is
checked."""
import numpy as np
import pytest

from _util import rel_l2
from sod2d_refmap_amd import cases, mesh as pmesh
from sod2d_refmap_amd.device import to_host

pytestmark = pytest.mark.gpu

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
