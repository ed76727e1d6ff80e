"""A run that starts from files: a SOD2D mesh file and a restart file read
through libsod2d_io, stepped on the device through sod_rk_4_main, against
the oracle stepping the same loaded mesh and state."""
import numpy as np
import pytest

import oracle
from _util import mesh_geometry, rel_l2
from sod2d_refmap_amd import cases, h5io, mesh as pmesh

pytestmark = pytest.mark.gpu

TOL2 = 1e-11   # two RK4 steps (1e-12 per right-hand side, atomics reorder sums)


def test_steps_from_mesh_and_restart_files(gpu, tmp_path):
    from sod2d_refmap_amd.solver import PeriodicSolver
    gen = pmesh.periodic_box(3, curved_amp=0.05, perm_seed=9)
    gas = cases.GasParams(Re=1600.0, Mach=0.1)
    s0 = cases.tgv_initial_state(gen, gas, perturb_seed=2, perturb_amp=1e-2)
    s0["mu_e"] = np.asfortranarray(np.random.default_rng(5).random((gen.nelem, 64)) * 1e-3)
    mpath = h5io.mesh_file_name(str(tmp_path) + "/", "box3", 1)
    rpath = h5io.restart_file_name(str(tmp_path) + "/", "box3", 1, 1)
    h5io.write_mesh(gen, mpath)
    h5io.save_restart(rpath, gen, s0, time=0.25, istep=10)

    part = h5io.read_mesh(mpath)
    m = part.mesh
    s, t, istep = h5io.load_restart(rpath, m, gas, masSla=part.masSla)
    assert (t, istep) == (0.25, 10)
    print(f"[h5io] mesh {mpath}: {m.nelem} elements, {m.npoin_w} working nodes; restart at step {istep}",
          flush=True)
    g = mesh_geometry(m)
    T = oracle.hex64_tables()
    dt = oracle.adapt_dt_cfl(m.connec, g["helem"], s["u"], s["csound"], 0.5, 0.5, s["mu_fluid"],
                             np.zeros((m.nelem, 64), order="F"), s["rho"])
    ref = {k: v.copy(order="F") for k, v in s.items()}
    p = oracle.default_params(Rgas_inf=gas.Rgas)
    for _ in range(2):
        oracle.rk_4_main(m.connec, m.lpoin_w, g["He"], g["gpvol"], T, g["Ml"], g["helem"], g["helem_l"], dt,
                         gas.Rgas, gas.gamma_gas, gas.Cp, gas.Prt, ref["mu_factor"], ref, p)
    sol = PeriodicSolver(m, gas, s)
    for _ in range(2):
        sol.rk_4_main(dt)
    out = sol.state()
    for k in ("rho", "q", "E", "u", "pr", "Tem", "eta", "csound", "mu_fluid", "mu_e"):
        err = rel_l2(out[k], ref[k])
        print(f"  {k} rel_l2={err:.3e}", flush=True)
        assert err < TOL2, (k, err)
    sol.close()
