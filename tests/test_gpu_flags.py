"""The reference's non-default branches of rk_4_main on every fused entry
point, against the oracle:

  * flag_total_enthalpy: full_convec_ijk_H with E := the conserved E
    (time_integ.f90:361-362) instead of full_convec_ijk with E := h;
  * the Sutherland gate flag_real_diff == 1 .and. flag_diff_suth == 1
    (time_integ.f90:312-316, 482-486): with it off the stages read the
    caller's mu_fluid unchanged, and with flag_real_diff == 0 the CFL step
    drops its diffusive limit (CFDSolverBase.f90:1132-1141).

rk_4_main is run through both implementations (fused stage kernels and the
per-routine sequence), sod_full_rhs against the sum of the oracle's routines,
and the device time loop step by step against the oracle's rk_4_main,
adapt_dt_cfl and volAvg_EK.  The caller's mu_fluid is a seeded field unlike
Sutherland's law, so that reading it or recomputing it cannot agree by
chance.  Tolerances: relative L2 <= 1e-12 per field for one step or RHS,
<= 1e-10 after 10 steps; dt to 1e-12 relative (it is computed from the
state, which carries the step's roundoff).
"""
import ctypes

import numpy as np
import pytest

import oracle
from _util import mesh_geometry, rel_l2, stage_fields
from sod2d_refmap_amd import _lib, cases, mesh as pmesh
from sod2d_refmap_amd.device import numerical_params, set_numerical_params, to_device, to_host

pytestmark = pytest.mark.gpu

T = oracle.hex64_tables()
TOL1 = 1e-12
TOL10 = 1e-10
KEYS = ("rho", "q", "E", "u", "pr", "Tem", "eta", "csound", "machno", "mu_fluid", "mu_e", "mue_l")

FLAGS = {
    "total_enthalpy": dict(flag_total_enthalpy=1),
    "suth_off": dict(flag_diff_suth=0),
    "real_diff_off": dict(flag_real_diff=0),
}
MESHES = {
    "curved3_perm": dict(n=3, curved_amp=0.05, perm_seed=11),
    "aniso324_curved_perm": dict(n=(3, 2, 4), L=(2 * np.pi, 4 * np.pi, 2 * np.pi), curved_amp=0.03, perm_seed=3),
}


@pytest.fixture(scope="module", params=list(MESHES))
def case(request, gpu):
    m = pmesh.periodic_box(**MESHES[request.param])
    g = mesh_geometry(m)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas, perturb_seed=7, perturb_amp=5e-2)
    # a caller's mu_fluid that is not Sutherland's law of the state
    rng = np.random.default_rng(21)
    s["mu_fluid"] = np.asfortranarray(s["mu_fluid"] * rng.uniform(0.5, 2.0, m.npoin))
    print(f"[flags case] {request.param} nelem={m.nelem} npoin={m.npoin}", flush=True)
    return request.param, m, g, gas, s


def _params(gas, flag):
    return numerical_params(Rgas_inf=gas.Rgas, **FLAGS[flag]), oracle.default_params(Rgas_inf=gas.Rgas,
                                                                                       **FLAGS[flag])


def _oracle_step(m, g, gas, s, dt, p):
    oracle.rk_4_main(m.connec, m.lpoin_w, g["He"], g["gpvol"], T, g["Ml"], g["helem"], g["helem_l"], dt, gas.Rgas,
                     gas.gamma_gas, gas.Cp, gas.Prt, s["mu_factor"], s, p)


def _oracle_dt(m, g, s, flag, cfl=0.5):
    zero = np.zeros((m.nelem, 64), order="F")
    cfl_diff = 1e300 if FLAGS[flag].get("flag_real_diff", 1) != 1 else cfl
    return oracle.adapt_dt_cfl(m.connec, g["helem"], s["u"], s["csound"], cfl, cfl_diff, s["mu_fluid"], zero,
                               s["rho"])


@pytest.mark.parametrize("flag", list(FLAGS))
@pytest.mark.parametrize("path", [0, 1], ids=["fused", "refseq"])
@pytest.mark.parametrize("nsteps,tol", [(1, TOL1), (10, TOL10)])
def test_rk_4_main_flags(case, flag, path, nsteps, tol):
    from sod2d_refmap_amd.solver import PeriodicSolver
    name, m, g, gas, s = case
    hp, op = _params(gas, flag)
    sol = PeriodicSolver(m, gas, s, params=hp)
    _lib.call("sod_set_rk4_path", path)
    dt = _oracle_dt(m, g, s, "suth_off")
    ref = {k: v.copy(order="F") for k, v in s.items()}
    for _ in range(nsteps):
        _oracle_step(m, g, gas, ref, dt, op)
        sol.rk_4_main(dt)
    out = sol.state()
    for k in KEYS:
        err = rel_l2(out[k], ref[k])
        print(f"  {flag} rk4[{path}] x{nsteps} {k} rel_l2={err:.3e}", flush=True)
        assert err < tol, (k, err)
    if flag != "total_enthalpy":
        # the gate off leaves the caller's viscosity untouched
        assert np.array_equal(out["mu_fluid"], s["mu_fluid"])


def test_flags_change_the_step(case):
    """Each flag selects a different operator: the oracle's step with it
    differs from the default step by far more than the tolerance (so the
    tests above cannot pass with a flag silently ignored)."""
    name, m, g, gas, s = case
    dt = _oracle_dt(m, g, s, "suth_off")
    base = {k: v.copy(order="F") for k, v in s.items()}
    _oracle_step(m, g, gas, base, dt, oracle.default_params(Rgas_inf=gas.Rgas))
    for flag in ("total_enthalpy", "suth_off"):
        other = {k: v.copy(order="F") for k, v in s.items()}
        _oracle_step(m, g, gas, other, dt, _params(gas, flag)[1])
        d = max(rel_l2(other[k], base[k]) for k in ("rho", "q", "E"))
        assert d > 1e3 * TOL1, (flag, d)


@pytest.mark.parametrize("flag", ["total_enthalpy", "suth_off"])
def test_full_rhs_flags(case, flag):
    name, m, g, gas, s = case
    hp, op = _params(gas, flag)
    set_numerical_params(hp)
    f = stage_fields(s, gas.gamma_gas, gas.Rgas)
    if flag == "suth_off":
        mu, mu_arg = s["mu_fluid"], s["mu_fluid"]
    else:
        mu, mu_arg = oracle.sutherland_viscosity(f["T"], s["mu_factor"]), s["mu_factor"]
    rng = np.random.default_rng(11)
    mu_e = np.asfortranarray(rng.uniform(0, 2e-3, (m.nelem, 64)))
    sgs = np.zeros((m.nelem, 64), order="F")
    th = flag == "total_enthalpy"
    cm, cq, ce = oracle.full_convec_ijk(m.connec, g["He"], g["gpvol"], T, f["u"], s["q"], s["rho"], f["p"],
                                        s["E"] if th else f["h"], th)
    dm, dq, de = oracle.full_diffusion_ijk(m.connec, g["He"], g["gpvol"], T, gas.Cp, gas.Prt, s["rho"], s["rho"],
                                           f["u"], f["T"], mu, mu_e, sgs)
    rt = oracle.generic_scalar_convec_ijk(m.connec, g["He"], g["gpvol"], T, f["feta"], f["eta"], f["u"])
    npn = m.npoin
    t = {k: to_device(v) for k, v in dict(c=m.connec, lw=m.lpoin_w, He=g["He"], gv=g["gpvol"], Ml=g["Ml"],
                                          r=s["rho"], q=s["q"], E=s["E"], mf=mu_arg, mue=mu_e, sgs=sgs).items()}
    import torch
    out = {k: torch.zeros(n, dtype=torch.float64, device="cuda") for k, n in
           (("Rm", npn), ("Rq", 3 * npn), ("Re", npn), ("Rt", npn))}
    _lib.call("sod_full_rhs", m.nelem, npn, m.npoin_w, *[t[k].data_ptr() for k in
                                                          ("c", "lw", "He", "gv", "Ml", "r", "q", "E", "mf", "mue",
                                                           "sgs")],
              gas.Rgas, gas.gamma_gas, gas.Cp, gas.Prt, *[out[k].data_ptr() for k in ("Rm", "Rq", "Re", "Rt")])
    for k, shp, r in (("Rm", (npn,), cm + dm), ("Rq", (npn, 3), cq + dq), ("Re", (npn,), ce + de),
                      ("Rt", (npn,), rt)):
        err = rel_l2(to_host(out[k], shp), r)
        print(f"  full_rhs[{flag}] {k} rel_l2={err:.3e}", flush=True)
        assert err < TOL1, (k, err)
    if th:
        # the adjoint of full_convec_ijk_H is not implemented: refused, not wrong
        z = torch.zeros(3 * npn, dtype=torch.float64, device="cuda")
        zm = torch.zeros(64 * m.nelem, dtype=torch.float64, device="cuda")
        with pytest.raises(_lib.Sod2dError, match="flag_total_enthalpy"):
            _lib.call("sod_full_rhs_vjp", m.nelem, npn, m.npoin_w,
                      *[t[k].data_ptr() for k in ("c", "lw", "He", "gv", "Ml", "r", "q", "E", "mf", "mue", "sgs")],
                      gas.Rgas, gas.gamma_gas, gas.Cp, gas.Prt, *[z.data_ptr()] * 7, zm.data_ptr())


@pytest.mark.parametrize("flag", list(FLAGS))
def test_time_loop_flags_vs_oracle(case, flag):
    """The graph-captured device loop, one replay per step, against the
    oracle's rk_4_main + adapt_dt_cfl (+ volAvg_EK) after every step."""
    from sod2d_refmap_amd.solver import PeriodicSolver
    name, m, g, gas, s = case
    hp, op = _params(gas, flag)
    sol = PeriodicSolver(m, gas, s, params=hp)
    ref = {k: v.copy(order="F") for k, v in s.items()}
    dt = _oracle_dt(m, g, ref, flag)
    sol.dt = dt
    time = 0.0
    for istep in range(1, 5):
        ek = sol.run_device(1, istep0=istep, log_first=1, log_every=1, keep_levels=True)
        _oracle_step(m, g, gas, ref, dt, op)
        time += dt
        ek_ref = oracle.volAvg_EK(m.connec, g["gpvol"], T.Ngp, gas.rho0, ref["rho"], ref["u"])
        dt = _oracle_dt(m, g, ref, flag)
        out = sol.state()
        for k in KEYS:
            err = rel_l2(out[k], ref[k])
            assert err < TOL10, (istep, k, err)
        assert abs(ek[0] / ek_ref - 1) < 1e-12, (istep, ek[0], ek_ref)
        assert abs(sol.dt / dt - 1) < 1e-12, (istep, sol.dt, dt)
        assert abs(sol.time / time - 1) < 1e-14
    sol.close()
