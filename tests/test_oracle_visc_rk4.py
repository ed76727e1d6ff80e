"""Pins the oracle's entropy viscosity (smart_visc_spectral), CFL time step,
kinetic-energy average, Sutherland law, lumped solve and the RK4 step."""
import numpy as np
import pytest

import oracle
from _util import mesh_geometry
from sod2d_refmap_amd import cases, mesh as pmesh

T = oracle.hex64_tables()


@pytest.fixture(scope="module")
def tgv3():
    m = pmesh.periodic_box(3, curved_amp=0.05, perm_seed=2)
    g = mesh_geometry(m)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas)
    return m, g, gas, s


def _visc(m, g, gas, s, Reta, **prm):
    p = oracle.default_params(Rgas_inf=gas.Rgas, **prm)
    return oracle.smart_visc_spectral(m.connec, m.lpoin_w, Reta, g["gpvol"], T.wgp, gas.gamma_gas, s["rho"],
                                      s["u"], s["csound"], s["Tem"], s["eta"], g["helem_l"], g["helem"],
                                      T.invAtoIJK, p)


def _beta(m, g, gas, s, cmax=0.5):
    n = m.connec - 1
    vm = np.sqrt((s["u"][n] ** 2).sum(-1))
    c = np.sqrt(gas.Rgas * gas.gamma_gas * s["Tem"][n])
    return (s["rho"][n] * g["helem"][:, None] * (cmax / 3.0) * (vm + c)).max(axis=1)


def test_entropy_viscosity_zero_residual(tgv3):
    m, g, gas, s = tgv3
    mu_e, mue_l = _visc(m, g, gas, s, np.zeros(m.npoin))
    assert np.all(mu_e == 0.0) and np.all(mue_l == 0.0)


def test_entropy_viscosity_saturated_limiter(tgv3):
    m, g, gas, s = tgv3
    mu_e, mue_l = _visc(m, g, gas, s, np.full(m.npoin, 1e30), cglob=0.7)
    beta = _beta(m, g, gas, s)
    np.testing.assert_allclose(mue_l, np.broadcast_to(0.7 * beta[:, None], mue_l.shape), rtol=1e-14)
    # (1/4,1/2,1/4)^3 weights sum to one: a constant element field is preserved
    np.testing.assert_allclose(mu_e, np.broadcast_to(0.7 * beta[:, None], mu_e.shape), rtol=1e-13)


def _smooth_line(v):
    """3-point (1/4,1/2,1/4) smoothing along axis 0 of a geometric-order line
    of 4 values.  End treatment of the reference's index table for p=3:
    position -1 reads position 1 (mirror), position 4 reads position 1."""
    ext = np.concatenate([v[1:2], v, v[1:2]], axis=0)
    return 0.25 * ext[:-2] + 0.5 * ext[1:-1] + 0.25 * ext[2:]


def test_entropy_viscosity_filter_brute_force(tgv3):
    m, g, gas, s = tgv3
    rng = np.random.default_rng(0)
    Reta = np.asfortranarray(rng.uniform(0.0, 1e-3, m.npoin))
    mu_e, mue_l = _visc(m, g, gas, s, Reta)
    beta = _beta(m, g, gas, s)
    assert np.all(mue_l >= 0) and np.all(mue_l <= beta[:, None] * (1 + 1e-14))
    assert np.all(mu_e >= 0) and np.all(mu_e <= beta[:, None] * (1 + 1e-14))
    # unfiltered value: ced * rho |Reta| / norm * helem_l^2 (ced = ce on affine-ish elements)
    eta = s["eta"][m.lpoin_w - 1]
    norm = np.abs(eta - eta.mean()).max()
    n = m.connec - 1
    Ve_min = 0.1 * s["rho"][n] * Reta[n] / norm * g["helem_l"] ** 2
    assert np.all(mue_l <= np.maximum(Ve_min, 0) * (1 / 0.1) + 1e-300)   # ced <= 1
    # separable smoothing in geometric order
    inv = T.invAtoIJK - 1
    geo2sod = np.array([0, 2, 3, 1])            # geometric position -> sod2d index
    lat = inv[np.ix_(geo2sod, geo2sod, geo2sod)]  # gmsh node at geometric (gi,gj,gk)
    f = mue_l[:, lat]                            # (nelem,4,4,4)
    for ax in (1, 2, 3):
        f = np.moveaxis(_smooth_line(np.moveaxis(f, ax, 0)), 0, ax)
    np.testing.assert_allclose(mu_e[:, lat], f, rtol=1e-13, atol=1e-300)


def test_adapt_dt_closed_form():
    m = pmesh.periodic_box(2)
    g = mesh_geometry(m)
    n = m.npoin
    U, c, mu, rho = 0.7, 11.0, 2e-3, 1.3
    u = np.asfortranarray(np.tile([U, -0.2, 0.1], (n, 1)))
    cs = np.full(n, c)
    dt = oracle.adapt_dt_cfl(m.connec, g["helem"], u, cs, 0.5, 0.4, np.full(n, mu),
                             np.zeros((m.nelem, 64), order="F"), np.full(n, rho))
    h = np.pi
    expect = min(0.5 * (h / 7.0) / (U + c), 0.4 * (h / 7.0) ** 2 / (mu / rho))
    assert abs(dt / expect - 1) < 1e-14
    dt2 = oracle.adapt_dt_cfl(m.connec, g["helem"], u, cs, 0.5, 0.4, np.full(n, 20.0),
                              np.zeros((m.nelem, 64), order="F"), np.full(n, rho))
    assert abs(dt2 / (0.4 * (h / 7.0) ** 2 / (20.0 / rho)) - 1) < 1e-14


def test_volavg_ek_closed_form():
    m = pmesh.periodic_box(6)
    g = mesh_geometry(m)
    s = cases.tgv_initial_state(m, cases.GasParams(), ref_fp32=False)
    ek = oracle.volAvg_EK(m.connec, g["gpvol"], T.Ngp, 1.0, np.ones(m.npoin), s["u"])
    # int 1/2 (u^2+v^2) over the box = pi^3; the reference divides by (2*3.14159)^3
    # and integrates with FP32-rounded GLL weights (sum 8*(1+O(1e-8)))
    assert abs(ek / (np.pi ** 3 / (2 * 3.14159) ** 3) - 1) < 1e-5


def test_sutherland_tgv_reference_viscosity():
    gas = cases.GasParams()
    mu = oracle.sutherland_viscosity(np.array([gas.to]), np.array([gas.mu_factor]))
    # mu_factor is chosen so that mu(T0) = rho0 V0 L / Re (TGVSolver.f90:97-103)
    assert abs(mu[0] * gas.Re - 1.0) < 1e-14


def test_lumped_solver_working_nodes_only():
    m = pmesh.periodic_box(2)
    g = mesh_geometry(m)
    R = np.asfortranarray(np.random.default_rng(1).standard_normal((m.npoin, 3)))
    Rv = oracle.lumped_solver_vect(m.lpoin_w, g["Ml"], R)
    w = m.lpoin_w - 1
    np.testing.assert_array_equal(Rv[w], R[w] / g["Ml"][w, None])
    sl = m.is_slave
    np.testing.assert_array_equal(Rv[sl], R[sl])
    Rs = oracle.lumped_solver_scal(m.lpoin_w, g["Ml"], R[:, 0].copy())
    np.testing.assert_array_equal(Rs[w], R[w, 0] / g["Ml"][w])


# ------------------------------------------------------------------ RK4
def _step(m, g, gas, s, dt, **prm):
    p = oracle.default_params(Rgas_inf=gas.Rgas, **prm)
    return oracle.rk_4_main(m.connec, m.lpoin_w, g["He"], g["gpvol"], T, g["Ml"], g["helem"], g["helem_l"], dt,
                            gas.Rgas, gas.gamma_gas, gas.Cp, gas.Prt, s["mu_factor"], s, p)


def test_rk4_uniform_state_is_steady():
    m = pmesh.periodic_box(2, curved_amp=0.05)
    g = mesh_geometry(m)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas)
    n = m.npoin
    for k in ("rho", "pr", "Tem"):
        s[k][:] = s[k].mean()
    s["u"][:] = [0.2, -0.1, 0.05]
    s["q"][:] = s["rho"][:, None] * s["u"]
    s["E"][:] = s["rho"] * (s["pr"] / (s["rho"] * 0.4) + 0.5 * (s["u"] ** 2).sum(1))
    ref = {k: v.copy() for k, v in s.items()}
    _step(m, g, gas, s, 1e-2)
    for k in ("rho", "E"):
        np.testing.assert_allclose(s[k], ref[k], rtol=1e-13)
    np.testing.assert_allclose(s["q"], ref["q"], rtol=1e-12, atol=1e-14)


def test_rk4_conserves_mass_and_momentum():
    m = pmesh.periodic_box(3, perm_seed=4)
    g = mesh_geometry(m)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas, perturb_seed=3)
    w = m.lpoin_w - 1
    Ml = g["Ml"][w]
    m0 = (Ml * s["rho"][w]).sum()
    q0 = (Ml[:, None] * s["q"][w]).sum(0)
    s1 = {k: v.copy(order="F") for k, v in s.items()}
    # momentum: convective split form and viscous weak form are both conservative
    for _ in range(3):
        _step(m, g, gas, s, 5e-3)
    np.testing.assert_allclose((Ml[:, None] * s["q"][w]).sum(0), q0, atol=1e-11 * m0)
    assert np.any(s["mu_e"] > 0)
    # mass: conservative without artificial viscosity; the entropy-viscosity
    # mass diffusion uses nu_e at the test node, which is not conservative
    assert abs((Ml * s["rho"][w]).sum() / m0 - 1) > 1e-10
    for _ in range(3):
        _step(m, g, gas, s1, 5e-3, cglob=0.0)
    assert abs((Ml * s1["rho"][w]).sum() / m0 - 1) < 1e-12


def test_rk4_fourth_order_in_time():
    m = pmesh.periodic_box(2)
    g = mesh_geometry(m)
    gas = cases.GasParams(Re=100.0)
    s0 = cases.tgv_initial_state(m, gas, perturb_seed=1, perturb_amp=5e-2)
    Tend = 0.08

    def run(nsteps):
        s = {k: v.copy(order="F") for k, v in s0.items()}
        for _ in range(nsteps):
            _step(m, g, gas, s, Tend / nsteps, cglob=0.0)    # mu_e == 0: pure RK4 in time
        return s

    ref = run(64)
    errs = []
    for ns in (4, 8, 16):
        s = run(ns)
        errs.append(np.abs(s["q"] - ref["q"]).max())
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 11 < r1 < 21 and 11 < r2 < 21, errs
