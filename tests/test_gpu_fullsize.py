"""Parity at full size: the 64^3 Taylor-Green case and the 96^3 curvilinear
case (elements permuted, Taylor-Green plus seeded Fourier modes |k| <= 8 of
amplitude 1e-2, Re 3200, Mach 0.2), through checks whose cost does not grow
with the mesh:

- the assembled right-hand side (sod_full_rhs) at the 8 interior nodes of
  sampled elements, which only their own element touches, equals the
  oracle's right-hand side of that element alone (relative L2 <= 1e-12 over
  the samples, per field).  The states carry seeded Fourier modes: the
  unperturbed Taylor-Green state is close to equilibrium, and its residuals
  are 30 to 2000 times smaller than the terms they are formed from (the
  oracle on the 8^3 case), so a comparison there would measure that
  cancellation rather than the kernels;
- momentum is conserved over RK4 steps (split-form convection and weak-form
  viscous terms are conservative), and mass too with the entropy viscosity
  off (cglob = 0), as the oracle shows on small meshes;
- steps from one saved state are bit-reproducible;
- the differentiation case at its full size (24^3, 200 RK4 steps): the RHS
  adjoint against central differences with the entropy viscosity frozen, and
  the checkpointed gradient of the full step against one 200-step segment.

The initial states are sod2d_refmap_amd.cases.tgv_state_on_device:
tgv_initial_state evaluated on the device from the solver's coordinates
(the host version takes minutes at 96^3)."""
import numpy as np
import pytest
import torch

import oracle
from _util import rel_l2, stage_fields
from sod2d_refmap_amd import _lib, cases, mesh as pmesh
from sod2d_refmap_amd.device import numerical_params, ptr, set_numerical_params

pytestmark = pytest.mark.gpu

T = oracle.hex64_tables()
CASES = {
    "tgv64": dict(n=64, curved_amp=0.0, perm_seed=None, Re=1600.0, Mach=0.1, modes_seed=None),
    "curv96_perm_modes": dict(n=96, curved_amp=0.05, perm_seed=1, Re=3200.0, Mach=0.2, modes_seed=2),
}
INTERIOR = np.nonzero((T.gmshAtoI >= 3) & (T.gmshAtoJ >= 3) & (T.gmshAtoK >= 3))[0]   # local volume nodes


@pytest.fixture(scope="module", params=list(CASES))
def big(request, gpu):
    from sod2d_refmap_amd.solver import PeriodicSolver
    kw = CASES[request.param]
    print(f"[fullsize] {request.param}: building the mesh", flush=True)
    m = pmesh.periodic_box(kw["n"], curved_amp=kw["curved_amp"], perm_seed=kw["perm_seed"])
    gas = cases.GasParams(Re=kw["Re"], Mach=kw["Mach"])
    set_numerical_params(numerical_params(Rgas_inf=gas.Rgas))
    print(f"[fullsize] {request.param}: {m.nelem} elements, {m.npoin_w} working nodes; solver set-up", flush=True)
    sol = PeriodicSolver(m, gas, cases.placeholder_state(m, gas))
    cases.tgv_state_on_device(sol, m, gas, kw["modes_seed"])
    print(f"[fullsize] {request.param}: state ready", flush=True)
    yield request.param, m, gas, sol
    sol.close()
    del sol
    torch.cuda.empty_cache()


def test_rhs_of_sampled_elements_matches_oracle(big):
    name, m, gas, sol = big
    ne, npn = m.nelem, m.npoin
    ts, saved = _snapshot(sol)
    if CASES[name]["modes_seed"] is None:
        cases.tgv_state_on_device(sol, m, gas, 3)
    gen = torch.Generator(device="cuda").manual_seed(17)
    mu_e = torch.rand(ne * 64, dtype=torch.float64, device="cuda", generator=gen) * 2e-3
    z = lambda k: torch.zeros(k, dtype=torch.float64, device="cuda")
    Rm, Rq, Re, Rt = z(npn), z(3 * npn), z(npn), z(npn)
    sgs = z(ne * 64)
    _lib.call("sod_full_rhs", ne, npn, m.npoin_w, ptr(sol.connec), ptr(sol.lpoin_w), ptr(sol.He), ptr(sol.gpvol),
              ptr(sol.Ml), ptr(sol.level("rho")), ptr(sol.level("q")), ptr(sol.level("E")), ptr(sol.mu_factor),
              ptr(mu_e), ptr(sgs), gas.Rgas, gas.gamma_gas, gas.Cp, gas.Prt, ptr(Rm), ptr(Rq), ptr(Re), ptr(Rt))
    torch.cuda.synchronize()
    rng = np.random.default_rng(5)
    elems = rng.choice(ne, size=48, replace=False)
    got = {k: [] for k in ("Rm", "Rq", "Re", "Rt")}
    ref = {k: [] for k in got}
    c1 = np.asfortranarray(np.arange(1, 65, dtype=np.int32).reshape(1, 64))
    rho2, q2, E2 = sol.level("rho"), sol.level("q").view(3, npn), sol.level("E")
    for e in elems:
        g = torch.from_numpy(np.asarray(m.connec[e], dtype=np.int64) - 1).cuda()
        st = {"rho": rho2[g].cpu().numpy(), "q": np.asfortranarray(q2[:, g].cpu().numpy().T), "E": E2[g].cpu().numpy()}
        mf = sol.mu_factor[g].cpu().numpy()
        He = np.asfortranarray(sol.He[576 * e:576 * (e + 1)].cpu().numpy().reshape(3, 3, 64, 1, order="F"))
        gv = np.asfortranarray(sol.gpvol[64 * e:64 * (e + 1)].cpu().numpy().reshape(1, 64, 1, order="F"))
        mue = np.asfortranarray(mu_e[e + ne * torch.arange(64, device="cuda")].cpu().numpy().reshape(1, 64))
        f = stage_fields(st, gas.gamma_gas, gas.Rgas)
        mu = oracle.sutherland_viscosity(f["T"], mf)
        cm, cq, ce = oracle.full_convec_ijk(c1, He, gv, T, f["u"], st["q"], st["rho"], f["p"], f["h"])
        dm, dq, de = oracle.full_diffusion_ijk(c1, He, gv, T, gas.Cp, gas.Prt, st["rho"], st["rho"], f["u"], f["T"],
                                               mu, mue, np.zeros((1, 64), order="F"))
        rt = oracle.generic_scalar_convec_ijk(c1, He, gv, T, f["feta"], f["eta"], f["u"])
        gi = g[torch.from_numpy(INTERIOR).cuda()]
        got["Rm"].append(Rm[gi].cpu().numpy())
        got["Rq"].append(Rq.view(3, npn)[:, gi].cpu().numpy().T)
        got["Re"].append(Re[gi].cpu().numpy())
        got["Rt"].append(Rt[gi].cpu().numpy())
        ref["Rm"].append((cm + dm)[INTERIOR])
        ref["Rq"].append((cq + dq)[INTERIOR])
        ref["Re"].append((ce + de)[INTERIOR])
        ref["Rt"].append(rt[INTERIOR])
    errs = {}
    for k in got:
        errs[k] = rel_l2(np.concatenate(got[k]), np.concatenate(ref[k]))
        print(f"  {name} sampled RHS {k} rel_l2={errs[k]:.3e}", flush=True)
    for k, v in ts.items():
        v.copy_(saved[k])
    torch.cuda.synchronize()
    for k, err in errs.items():
        assert err < 1e-12, (k, err)


def _sums(sol, m):
    w = sol.lpoin_w.long() - 1
    Ml = sol.Ml[w]
    npn = m.npoin
    mass = float((Ml * sol.level("rho")[w]).sum())
    mom = (Ml[None, :] * sol.level("q").view(3, npn)[:, w]).sum(1).cpu().numpy()
    return mass, mom


def _snapshot(sol):
    ts = {k: v for k, v in vars(sol).items() if isinstance(v, torch.Tensor)}
    ts.update({"lv." + k: v for k, v in sol.lv.items()})
    return ts, {k: v.clone() for k, v in ts.items()}


def test_conservation_and_reproducibility(big):
    name, m, gas, sol = big
    ts, saved = _snapshot(sol)
    dt = sol.adapt_dt()
    m0, q0 = _sums(sol, m)
    print(f"  {name}: dt={dt:.4e} mass={m0:.6e} momentum={q0}", flush=True)
    finals = []
    for rep in range(2):
        for k, v in ts.items():
            v.copy_(saved[k])
        torch.cuda.synchronize()
        for _ in range(3):
            sol.rk_4_main(dt)
        torch.cuda.synchronize()
        finals.append({k: sol.level(k).clone() for k in ("rho", "q", "E")} | {"mu_e": sol.mu_e.clone()})
    for k in finals[0]:
        assert torch.equal(finals[0][k], finals[1][k]), k
    assert float(sol.mu_e.abs().max()) > 0.0
    _, q1 = _sums(sol, m)
    np.testing.assert_allclose(q1, q0, atol=1e-11 * m0)
    # mass, with the entropy viscosity off
    p0 = numerical_params(Rgas_inf=gas.Rgas)
    try:
        set_numerical_params(numerical_params(Rgas_inf=gas.Rgas, cglob=0.0))
        for k, v in ts.items():
            v.copy_(saved[k])
        torch.cuda.synchronize()
        for _ in range(3):
            sol.rk_4_main(dt)
        torch.cuda.synchronize()
        m1, _ = _sums(sol, m)
        print(f"  {name}: mass drift with cglob=0: {abs(m1 / m0 - 1):.2e}", flush=True)
        assert abs(m1 / m0 - 1) < 1e-12
    finally:
        set_numerical_params(p0)
    for k, v in ts.items():
        v.copy_(saved[k])
    torch.cuda.synchronize()


# ------------------------------------------------ adjoint at full size
def _geometry_of(m):
    from _util import mesh_geometry
    return mesh_geometry(m)


def _adjoint_case():
    """The differentiation case at its full size: Taylor-Green on 24^3
    elements, Re 1600, Mach 0.1, 200 RK4 steps at the initial CFL 0.5 step."""
    from _util import mesh_geometry
    from sod2d_refmap_amd.autograd import EVContext, RHSContext
    m = pmesh.periodic_box(24)
    g = mesh_geometry(m)
    gas = cases.GasParams(Re=1600.0, Mach=0.1)
    s = cases.tgv_initial_state(m, gas)
    cu = lambda a: torch.tensor(np.asarray(a), dtype=torch.float64, device="cuda")
    rho0, p0, u0 = cu(s["rho"]), cu(s["pr"]), cu(s["u"])
    g1 = gas.gamma_gas - 1.0

    def initial(u):   # velocity varied at fixed density and pressure
        return rho0 * 0 + rho0, rho0[:, None] * u, p0 / g1 + 0.5 * rho0 * (u * u).sum(-1)

    dt = oracle.adapt_dt_cfl(m.connec, g["helem"], s["u"], s["csound"], 0.5, 0.5, s["mu_fluid"],
                             np.zeros((m.nelem, 64), order="F"), s["rho"])

    def contexts():
        rc = RHSContext.from_mesh(m, g["He"], g["gpvol"], g["Ml"], s["mu_factor"], gas)
        return rc, EVContext.create(rc, m, g["helem"], g["helem_l"], T)
    return m, gas, u0, initial, dt, contexts


def test_adjoint_200_steps_frozen_viscosity_central_differences(gpu):
    """d(kinetic energy after 200 steps)/d(initial velocity) through the
    hand-written RHS adjoint, with the entropy viscosity held fixed (a smooth
    map), against central differences (step 1e-6) along 8 seeded random
    directions of unit RMS per degree of freedom.  Directions of unit L2 norm
    give directional derivatives of ~1e-7 against an objective of ~0.1, and
    the FP64 rounding of the objective then limits the difference quotient
    to ~1e-5."""
    from sod2d_refmap_amd.autograd import kinetic_energy, rk4_step
    m, gas, u0, initial, dt, contexts = _adjoint_case()
    set_numerical_params(numerical_params(Rgas_inf=gas.Rgas))
    rc, _ = contexts()
    mu_e = torch.tensor(np.random.default_rng(7).uniform(0, 2e-3, (m.nelem, 64)), dtype=torch.float64,
                        device="cuda")

    def run(u):
        r, q, E = initial(u)
        for _ in range(200):
            r, q, E, _ = rk4_step(rc, r, q, E, mu_e, dt)
        return kinetic_energy(rc, r, q)

    u = u0.clone().requires_grad_(True)
    (gu,) = torch.autograd.grad(run(u), u)
    rng = np.random.default_rng(3)
    h = 1e-6
    for it in range(8):
        v = torch.tensor(rng.standard_normal((m.npoin, 3)), dtype=torch.float64, device="cuda")
        v = v * (v.numel() ** 0.5 / torch.linalg.norm(v))
        with torch.no_grad():
            fd = (float(run(u0 + h * v)) - float(run(u0 - h * v))) / (2 * h)
        an = float((gu * v).sum())
        print(f"  frozen-viscosity adjoint, direction {it}: fd={fd:.12e} adjoint={an:.12e}", flush=True)
        assert abs(fd - an) <= 1e-6 * abs(an), (it, fd, an)


def test_adjoint_200_steps_entropy_viscosity_32_directions(gpu):
    """The differentiation case as specified: d(kinetic energy after 200
    steps)/d(initial velocity) through the hand-written adjoints of the whole
    step -- RK4 stages, final EOS and the entropy viscosity with its limiters
    and normalisation -- with a checkpoint every 20 steps, against central
    differences (step 1e-6) along 32 seeded random directions of unit RMS per
    degree of freedom: <= 1e-6 relative on every direction.  The +-h
    trajectories follow the entropy viscosity's branches recorded on the
    unperturbed trajectory (frozen_visc.py), the piecewise-smooth map's
    branch that the adjoint differentiates; without that, limiter switches
    inside the stencil move the difference quotient by 3e-5 to 3e-4."""
    from frozen_visc import FrozenSmartVisc, trajectory
    from sod2d_refmap_amd.autograd import kinetic_energy, run_checkpointed
    m, gas, u0, initial, dt, contexts = _adjoint_case()
    set_numerical_params(numerical_params(Rgas_inf=gas.Rgas))
    rc, ec = contexts()
    g = _geometry_of(m)
    dts = [dt] * 200
    mu0 = torch.zeros((m.nelem, 64), dtype=torch.float64, device="cuda")
    u = u0.clone().requires_grad_(True)
    J = run_checkpointed(ec, *initial(u), mu0, dts, every=20)
    (gu,) = torch.autograd.grad(J, u)
    fz = FrozenSmartVisc(m, g["gpvol"], T.wgp, g["helem"], g["helem_l"], T.invAtoIJK, gas.gamma_gas, gas.Rgas)
    yb, branches = trajectory(ec, fz, (*initial(u0), mu0), dts)
    Jb = float(kinetic_energy(rc, yb[0], yb[1]))
    print(f"[adjoint 24^3] J={float(J):.15e} recorded-branch J={Jb:.15e}", flush=True)
    assert abs(Jb - float(J)) <= 1e-12 * abs(float(J))

    def forward(uu):
        y, _ = trajectory(ec, fz, (*initial(uu), mu0), dts, branches)
        return float(kinetic_energy(rc, y[0], y[1]))

    rng = np.random.default_rng(3)
    h = 1e-6
    worst = 0.0
    for it in range(32):
        v = torch.tensor(rng.standard_normal((m.npoin, 3)), dtype=torch.float64, device="cuda")
        v = v * (v.numel() ** 0.5 / torch.linalg.norm(v))
        fd = (forward(u0 + h * v) - forward(u0 - h * v)) / (2 * h)
        an = float((gu * v).sum())
        err = abs(fd - an) / abs(an)
        worst = max(worst, err)
        print(f"  direction {it}: fd={fd:.12e} adjoint={an:.12e} rel={err:.2e}", flush=True)
        assert err <= 1e-6, (it, fd, an)
    print(f"[adjoint 24^3] worst relative difference over 32 directions: {worst:.2e}", flush=True)


def test_adjoint_200_steps_checkpointed_equals_one_segment(gpu):
    """The full differentiable step (RK4 stages, final EOS and the entropy
    viscosity, its limiters as subgradients): the gradient with a checkpoint
    every 20 steps equals the one taken through a single 200-step segment."""
    from sod2d_refmap_amd.autograd import run_checkpointed
    m, gas, u0, initial, dt, contexts = _adjoint_case()
    set_numerical_params(numerical_params(Rgas_inf=gas.Rgas, flag_normalise_entropy=0))
    try:
        _, ec = contexts()
        mu0 = torch.zeros((m.nelem, 64), dtype=torch.float64, device="cuda")
        grads = []
        for every in (20, 200):
            u = u0.clone().requires_grad_(True)
            J = run_checkpointed(ec, *initial(u), mu0, [dt] * 200, every=every)
            (gu,) = torch.autograd.grad(J, u)
            grads.append(gu)
            print(f"  checkpoint every {every}: J={float(J.detach()):.15e} |grad|={float(gu.norm()):.6e}", flush=True)
        err = float((grads[0] - grads[1]).norm() / grads[1].norm())
        assert err <= 1e-13, err
    finally:
        set_numerical_params(numerical_params(Rgas_inf=gas.Rgas))


# ------------------------------------------------ ensemble at full size
def test_ensemble_of_48_matches_members_stepped_one_by_one(gpu):
    """The batching case at its full size: 48 seeded Taylor-Green boxes of
    8^3 to 32^3 elements (Re, Mach and Fourier modes seeded), 3 steps of the
    batched device loop against the same steps member by member through the
    plain entry points (whose parity with the oracle the small-mesh tests
    show): relative L2 <= 1e-12 per member and field, dt and time equal to
    1e-12."""
    from sod2d_refmap_amd.ensemble import Ensemble, ensemble_members, sequential_solvers, sequential_step
    members = ensemble_members(48)
    print(f"[fullsize] ensemble: {sum(mm.nelem for mm, _, _ in members)} elements in 48 members", flush=True)
    E = Ensemble(members)
    sols = sequential_solvers(members)
    E.run(3, keep_levels=False)
    for _ in range(3):
        sequential_step(sols)
    worst = 0.0
    for i, sol in enumerate(sols):
        a, b = E.member_state(i), sol.state()
        for k in ("rho", "q", "E", "u", "pr", "eta", "mu_fluid", "mu_e"):
            assert np.isfinite(a[k]).all(), (i, k)
            err = rel_l2(a[k], b[k])
            worst = max(worst, err)
            assert err < 1e-12, (i, k, err)
        assert abs(E.dt[i] / sol.dt - 1) < 1e-12 and abs(E.time[i] / sol.time - 1) < 1e-12, i
        sol.close()
    print(f"  ensemble of 48: largest relative L2 difference {worst:.2e}", flush=True)
    E.close()
