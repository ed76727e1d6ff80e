"""The hand-written adjoint of the assembled stage RHS (sod_full_rhs_vjp)
against autograd through the FP64 torch transcription of the reference
element routines (tests/torch_rhs.py, itself checked against the oracle),
and against central finite differences of the forward library call."""
import numpy as np
import pytest
import torch

import oracle
import torch_rhs as TR
from _util import mesh_geometry, rel_l2
from sod2d_refmap_amd import _lib, cases, mesh as pmesh
from sod2d_refmap_amd.device import to_device, to_host

pytestmark = pytest.mark.gpu

T = oracle.hex64_tables()


def _setup(n, sgs, seed):
    m = pmesh.periodic_box(n, curved_amp=0.05, perm_seed=seed)
    g = mesh_geometry(m)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas, perturb_seed=seed + 1, perturb_amp=5e-2)
    rng = np.random.default_rng(seed + 2)
    mu_e = np.asfortranarray(rng.uniform(0, 2e-3, (m.nelem, 64)))
    mu_sgs = np.asfortranarray(rng.uniform(0, 1e-3, (m.nelem, 64)) if sgs else np.zeros((m.nelem, 64)))
    lam = [rng.standard_normal(m.npoin), np.asfortranarray(rng.standard_normal((m.npoin, 3))),
           rng.standard_normal(m.npoin), rng.standard_normal(m.npoin)]
    return m, g, gas, s, mu_e, mu_sgs, lam


class Lib:
    def __init__(self, m, g, gas, s, mu_e, mu_sgs):
        self.m, self.gas = m, gas
        self.keep = {}
        d = lambda k, a: self.keep.setdefault(k, to_device(a)).data_ptr()
        self.static = [d("c", m.connec), d("lw", m.lpoin_w), d("He", g["He"]), d("gv", g["gpvol"]),
                       d("Ml", g["Ml"])]
        self.mf, self.mue, self.sgs = d("mf", s["mu_factor"]), d("mue", mu_e), d("sgs", mu_sgs)

    def _o(self, n):
        return torch.zeros(n, dtype=torch.float64, device="cuda")

    def rhs(self, rho, q, E):
        m, gas = self.m, self.gas
        r, qq, e = to_device(rho), to_device(q), to_device(E)
        out = [self._o(m.npoin), self._o(3 * m.npoin), self._o(m.npoin), self._o(m.npoin)]
        _lib.call("sod_full_rhs", m.nelem, m.npoin, m.npoin_w, *self.static, r.data_ptr(), qq.data_ptr(),
                  e.data_ptr(), self.mf, self.mue, self.sgs, gas.Rgas, gas.gamma_gas, gas.Cp, gas.Prt,
                  *[o.data_ptr() for o in out])
        return [to_host(out[0], (m.npoin,)), to_host(out[1], (m.npoin, 3)), to_host(out[2], (m.npoin,)),
                to_host(out[3], (m.npoin,))]

    def vjp(self, rho, q, E, lam):
        m, gas = self.m, self.gas
        r, qq, e = to_device(rho), to_device(q), to_device(E)
        ls = [to_device(x) for x in lam]
        out = [self._o(m.npoin), self._o(3 * m.npoin), self._o(m.npoin), self._o(64 * m.nelem)]
        _lib.call("sod_full_rhs_vjp", m.nelem, m.npoin, m.npoin_w, *self.static, r.data_ptr(), qq.data_ptr(),
                  e.data_ptr(), self.mf, self.mue, self.sgs, gas.Rgas, gas.gamma_gas, gas.Cp, gas.Prt,
                  *[x.data_ptr() for x in ls], *[o.data_ptr() for o in out])
        return [to_host(out[0], (m.npoin,)), to_host(out[1], (m.npoin, 3)), to_host(out[2], (m.npoin,)),
                to_host(out[3], (m.nelem, 64))]


def _autograd(m, g, gas, s, mu_e, mu_sgs, lam):
    geo = TR.Geometry(m.connec, g["He"], g["gpvol"], T)
    t = lambda a: torch.tensor(np.asarray(a), dtype=torch.float64, requires_grad=True)
    tr, tq, tE = t(s["rho"]), t(s["q"]), t(s["E"])
    tmue = TR.lattice_from_elemnode(geo, mu_e).requires_grad_(True)
    rm, rq, re, rt = TR.full_rhs(geo, tr, tq, tE, torch.tensor(s["mu_factor"]), tmue,
                                 TR.lattice_from_elemnode(geo, mu_sgs), gas.gamma_gas, gas.Rgas, gas.Cp, gas.Prt)
    L = ((rm * torch.tensor(lam[0])).sum() + (rq * torch.tensor(lam[1])).sum() +
         (re * torch.tensor(lam[2])).sum() + (rt * torch.tensor(lam[3])).sum())
    L.backward()
    # lattice [i,j,k] (i-major) -> gmsh node order
    lat = geo.lat.numpy()
    gm = np.zeros((m.nelem, 64))
    gm[:, lat] = tmue.grad.reshape(m.nelem, 64).numpy()
    return [tr.grad.numpy(), tq.grad.numpy(), tE.grad.numpy(), gm]


@pytest.mark.parametrize("n,sgs", [(2, False), (3, True)], ids=["curved2", "curved3_sgs"])
def test_vjp_matches_autograd(gpu, n, sgs):
    m, g, gas, s, mu_e, mu_sgs, lam = _setup(n, sgs, 7 * n)
    print(f"[vjp] n={n} sgs={sgs} npoin={m.npoin}", flush=True)
    L = Lib(m, g, gas, s, mu_e, mu_sgs)
    got = L.vjp(s["rho"], s["q"], s["E"], lam)
    ref = _autograd(m, g, gas, s, mu_e, mu_sgs, lam)
    for name, a, b in zip(("rho", "q", "E", "mu_e"), got, ref):
        err = rel_l2(a, b)
        print(f"  vjp {name} rel_l2={err:.3e}", flush=True)
        assert err < 1e-11, (name, err)


def test_vjp_central_differences(gpu):
    """lam . R(U + h v) - lam . R(U - h v) over 2h against g . v along seeded
    random directions v in (rho, q, E)."""
    m, g, gas, s, mu_e, mu_sgs, lam = _setup(2, False, 5)
    L = Lib(m, g, gas, s, mu_e, mu_sgs)
    gr, gq, gE, _ = L.vjp(s["rho"], s["q"], s["E"], lam)
    rng = np.random.default_rng(42)
    dot = lambda R: sum(float(np.sum(a * b)) for a, b in zip(R, lam))
    for it in range(4):
        v = [rng.standard_normal(m.npoin), rng.standard_normal((m.npoin, 3)), rng.standard_normal(m.npoin)]
        scale = [np.abs(s["rho"]).max(), np.abs(s["q"]).max(), np.abs(s["E"]).max()]
        v = [a * sc for a, sc in zip(v, scale)]
        nv = np.sqrt(sum(float(np.sum(a * a)) for a in v))
        v = [a / nv for a in v]
        h = 1e-6 * np.sqrt(float(np.sum(s["E"] ** 2)))
        up = L.rhs(s["rho"] + h * v[0], np.asfortranarray(s["q"] + h * v[1]), s["E"] + h * v[2])
        dn = L.rhs(s["rho"] - h * v[0], np.asfortranarray(s["q"] - h * v[1]), s["E"] - h * v[2])
        fd = (dot(up) - dot(dn)) / (2 * h)
        an = float(np.sum(gr * v[0]) + np.sum(gq * v[1]) + np.sum(gE * v[2]))
        print(f"  fd direction {it}: fd={fd:.10e} adjoint={an:.10e}", flush=True)
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_smart_visc_vjp_matches_autograd(gpu):
    """sod_smart_visc_spectral_vjp against autograd through the torch
    transcription of smart_visc_spectral (first-index subgradients)."""
    m, g, gas, s, mu_e, mu_sgs, lam = _setup(3, False, 11)
    from sod2d_refmap_amd.device import numerical_params, set_numerical_params
    set_numerical_params(numerical_params(Rgas_inf=gas.Rgas))
    rng = np.random.default_rng(17)
    Reta = rng.uniform(-1e-2, 1e-2, m.npoin)
    lam_mu = np.asfortranarray(rng.standard_normal((m.nelem, 64)))
    t = lambda a: torch.tensor(np.asarray(a), dtype=torch.float64, requires_grad=True)
    tr, tu, tc, tT, te, tR = t(s["rho"]), t(s["u"]), t(s["csound"]), t(s["Tem"]), t(s["eta"]), t(Reta)
    mu, _ = TR.smart_visc(m.connec, m.lpoin_w, tR, g["gpvol"], T.wgp, gas.gamma_gas, tr, tu, tc, tT, te,
                          g["helem_l"], g["helem"], T.invAtoIJK, gas.Rgas)
    (mu * torch.tensor(lam_mu)).sum().backward()
    ref = [tr.grad.numpy(), tu.grad.numpy(), tc.grad.numpy(), tT.grad.numpy(), te.grad.numpy(), tR.grad.numpy()]
    keep = []

    def d(a):
        x = to_device(a)
        keep.append(x)
        return x.data_ptr()

    outs = [torch.zeros(k * m.npoin, dtype=torch.float64, device="cuda") for k in (1, 3, 1, 1, 1, 1)]
    _lib.call("sod_smart_visc_spectral_vjp", m.nelem, m.npoin, m.npoin_w, d(m.connec), d(m.lpoin_w), d(Reta),
              d(g["gpvol"]), gas.gamma_gas, d(s["rho"]), d(s["u"]), d(s["csound"]), d(s["Tem"]), d(s["eta"]),
              d(g["helem_l"]), d(g["helem"]), d(lam_mu), *[o.data_ptr() for o in outs])
    shapes = [(m.npoin,), (m.npoin, 3), (m.npoin,), (m.npoin,), (m.npoin,), (m.npoin,)]
    for name, o, shp, r in zip(("rho", "u", "csound", "Tem", "eta", "Reta"), outs, shapes, ref):
        got = to_host(o, shp)
        nr = np.linalg.norm(r)
        err = np.linalg.norm(got - r) / (nr if nr > 0 else 1.0)
        print(f"  smart_visc vjp {name} |ref|={nr:.3e} rel_err={err:.3e}", flush=True)
        assert err < 1e-11, (name, err)


def fd4(f, h):
    """Fourth-order central difference of the scalar map a -> f(a) at 0."""
    return (8.0 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12.0 * h)


def _ev_context(m, g, gas, s):
    from sod2d_refmap_amd.autograd import EVContext, RHSContext
    rc = RHSContext.from_mesh(m, g["He"], g["gpvol"], g["Ml"], s["mu_factor"], gas)
    return EVContext.create(rc, m, g["helem"], g["helem_l"], T)


def _cuda(a):
    return torch.tensor(np.asarray(a), dtype=torch.float64, device="cuda")


def test_rk4_chain_matches_rk_4_main(gpu):
    """The differentiable step (FullRHS stages + final EOS + SmartVisc) is
    the same map as sod_rk_4_main."""
    from sod2d_refmap_amd.autograd import rk4_step_ev
    from sod2d_refmap_amd.device import numerical_params, set_numerical_params
    from sod2d_refmap_amd.solver import PeriodicSolver
    m = pmesh.periodic_box(3, curved_amp=0.05, perm_seed=2)
    g = mesh_geometry(m)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas, perturb_seed=4, perturb_amp=5e-2)
    s["mu_e"] = np.asfortranarray(np.random.default_rng(3).uniform(0, 1e-3, (m.nelem, 64)))
    sol = PeriodicSolver(m, gas, s)
    dt = sol.adapt_dt()
    set_numerical_params(numerical_params(Rgas_inf=gas.Rgas))
    ec = _ev_context(m, g, gas, s)
    y = (_cuda(s["rho"]), _cuda(s["q"]), _cuda(s["E"]), _cuda(s["mu_e"]))
    for _ in range(2):
        sol.rk_4_main(dt)
        y = rk4_step_ev(ec, *y, dt)
    st = sol.state()
    w = m.lpoin_w - 1
    for name, a, b in (("rho", y[0], st["rho"]), ("q", y[1], st["q"]), ("E", y[2], st["E"]),
                       ("mu_e", y[3], st["mu_e"])):
        a = a.detach().cpu().numpy()
        err = rel_l2(a[w] if name != "mu_e" else a, b[w] if name != "mu_e" else b)
        print(f"  chain vs rk_4_main {name} rel_l2={err:.3e}", flush=True)
        assert err < 1e-12, (name, err)


def test_smart_visc_vjp_alone_against_frozen_branches(gpu):
    """The entropy-viscosity adjoint (sod_smart_visc_spectral_vjp) alone,
    against fourth-order central differences of smart_visc_spectral along the
    branches recorded at the base state (frozen_visc.py), in all six inputs
    at once, with the residual mixing signs and saturating a seventh of the
    nodes: relative 1e-8 (a kink would show as an O(1) difference, a wrong
    derivative of a smooth branch as a difference that does not shrink with
    the step).  The map is strongly curved along these directions (the
    normalisation max |eta - mean| is ~0.04 here): a second-order quotient at
    step 1e-7 carries ~1e-8 of truncation error, the fourth-order one ~1e-11
    (both measured with the torch transcription and autograd on the CPU).
    The recorded-branch map equals the library's forward at the base state."""
    from frozen_visc import FrozenSmartVisc
    from sod2d_refmap_amd.autograd import SmartVisc
    from sod2d_refmap_amd.device import numerical_params, set_numerical_params
    m = pmesh.periodic_box(3, curved_amp=0.05, perm_seed=2)
    g = mesh_geometry(m)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas, perturb_seed=4, perturb_amp=5e-2)
    rng = np.random.default_rng(17)
    Reta = rng.uniform(-1e-2, 1e-2, m.npoin)
    Reta[rng.choice(m.npoin, m.npoin // 7, replace=False)] = 1e3
    base = [_cuda(s["rho"]), _cuda(s["u"]), _cuda(s["csound"]), _cuda(s["Tem"]), _cuda(s["eta"]), _cuda(Reta)]
    scale = [1.0, 1.0, 10.0, 300.0, 10.0, 1e-2]
    for normalise in (1, 0):
        set_numerical_params(numerical_params(Rgas_inf=gas.Rgas, flag_normalise_entropy=normalise))
        ec = _ev_context(m, g, gas, s)
        fz = FrozenSmartVisc(m, g["gpvol"], T.wgp, g["helem"], g["helem_l"], T.invAtoIJK, gas.gamma_gas, gas.Rgas,
                             normalise=bool(normalise))
        x = [t.clone().requires_grad_(True) for t in base]
        mu = SmartVisc.apply(*x, ec)
        mu_t, br = fz(*base)
        assert rel_l2(mu_t.cpu().numpy(), mu.detach().cpu().numpy()) < 1e-13
        assert bool(br["via_ve"].any()) and not bool(br["via_ve"].all())
        lam = torch.tensor(rng.standard_normal((m.nelem, 64)), dtype=torch.float64, device="cuda")
        grads = torch.autograd.grad((mu * lam).sum(), x)
        h = 1e-7
        for it in range(4):
            v = [sc * torch.tensor(rng.standard_normal(tuple(t.shape)), dtype=torch.float64, device="cuda")
                 for sc, t in zip(scale, base)]
            with torch.no_grad():
                f = lambda a: float((fz(*[b + a * d for b, d in zip(base, v)], branches=br)[0] * lam).sum())
                fd = fd4(f, h)
            an = float(sum((gg * d).sum() for gg, d in zip(grads, v)))
            print(f"  smart_visc vjp (normalise {normalise}) direction {it}: fd={fd:.12e} adjoint={an:.12e}",
                  flush=True)
            assert abs(fd - an) <= 1e-8 * abs(an), (normalise, it, fd, an)


def test_rk4_adjoint_checkpointed_fd(gpu):
    """Gradient of the kinetic energy after K RK4 steps with respect to the
    initial velocity (at fixed density and pressure), through the hand-written
    adjoints with checkpointing, against central differences of the forward
    chain (same dt sequence) along seeded random unit directions, with the
    entropy normalisation on and off.  The +-h chains follow the entropy
    viscosity's branches recorded on the unperturbed chain (frozen_visc.py),
    the map the adjoint differentiates: 1e-6 relative, with fourth-order
    central differences at step 1e-4 along the recorded branches.  At step
    1e-6 the objective's FP64 rounding (J ~ 3) puts up to 9e-7 of noise into
    the quotient where the directional derivative is small (7e-4); at 1e-4
    the quotient agrees with autograd through the torch transcription of the
    same chain to 1e-8 (measured on the CPU)."""
    from frozen_visc import FrozenSmartVisc, trajectory
    from sod2d_refmap_amd.autograd import kinetic_energy, run_checkpointed
    for normalise in (0, 1):
        _rk4_adjoint_fd(normalise, FrozenSmartVisc, trajectory, kinetic_energy, run_checkpointed)


def _rk4_adjoint_fd(normalise, FrozenSmartVisc, trajectory, kinetic_energy, run_checkpointed):
    from sod2d_refmap_amd.device import numerical_params, set_numerical_params
    m = pmesh.periodic_box(2, curved_amp=0.05, perm_seed=6)
    g = mesh_geometry(m)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas, perturb_seed=8, perturb_amp=5e-2)
    set_numerical_params(numerical_params(Rgas_inf=gas.Rgas, flag_normalise_entropy=normalise))
    ec = _ev_context(m, g, gas, s)
    rho0, p0, u0 = _cuda(s["rho"]), _cuda(s["pr"]), _cuda(s["u"])
    g1 = gas.gamma_gas - 1.0

    def initial(u):
        return rho0 * 0 + rho0, rho0[:, None] * u, p0 / g1 + 0.5 * rho0 * (u * u).sum(-1)

    dt = oracle.adapt_dt_cfl(m.connec, g["helem"], s["u"], s["csound"], 0.5, 0.5, s["mu_fluid"],
                             np.zeros((m.nelem, 64), order="F"), s["rho"])
    K, every = 6, 2
    dts = [dt] * K
    mu0 = _cuda(np.zeros((m.nelem, 64)))
    u = u0.clone().requires_grad_(True)
    J = run_checkpointed(ec, *initial(u), mu0, dts, every=every)
    (gu,) = torch.autograd.grad(J, u)
    fz = FrozenSmartVisc(m, g["gpvol"], T.wgp, g["helem"], g["helem_l"], T.invAtoIJK, gas.gamma_gas, gas.Rgas,
                         normalise=bool(normalise))
    yb, branches = trajectory(ec, fz, (*initial(u0), mu0), dts)

    def forward(uu):
        y, _ = trajectory(ec, fz, (*initial(uu), mu0), dts, branches)
        return float(kinetic_energy(ec.rc, y[0], y[1]))

    Jv = float(J.detach())
    assert abs(forward(u0) - Jv) <= 1e-13 * abs(Jv)
    rng = np.random.default_rng(32)
    h = 1e-4
    for it in range(4):
        v = torch.tensor(rng.standard_normal((m.npoin, 3)), dtype=torch.float64, device="cuda")
        v = v / torch.linalg.norm(v)
        fd = fd4(lambda a: forward(u0 + a * v), h)
        an = float((gu * v).sum())
        print(f"  rk4 adjoint (normalise {normalise}) direction {it}: fd={fd:.12e} adjoint={an:.12e}", flush=True)
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_vjp_costs_at_most_three_forward_rhs(gpu):
    """sod_full_rhs_vjp against sod_full_rhs on a 32^3 Taylor-Green box
    (32,768 elements): mean time per call from HIP events on the library
    stream over 5 calls each after a warm-up.  The adjoint recomputes the
    forward element work and runs three transposed line passes; it must stay
    within 3x a forward evaluation (a register-spilling build measured 13x)."""
    import ctypes
    from sod2d_refmap_amd.device import numerical_params, set_numerical_params
    from sod2d_refmap_amd.solver import PeriodicSolver
    m = pmesh.periodic_box(32)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas, perturb_seed=3, perturb_amp=1e-2)
    set_numerical_params(numerical_params(Rgas_inf=gas.Rgas))
    sol = PeriodicSolver(m, gas, s)
    n, ne = m.npoin, m.nelem
    z = lambda k: torch.zeros(k, dtype=torch.float64, device="cuda")
    common = [sol.connec.data_ptr(), sol.lpoin_w.data_ptr(), sol.He.data_ptr(), sol.gpvol.data_ptr(),
              sol.Ml.data_ptr(), sol.level("rho").data_ptr(), sol.level("q").data_ptr(), sol.level("E").data_ptr(),
              sol.mu_factor.data_ptr(), sol.mu_e.data_ptr(), sol.mu_sgs.data_ptr(), gas.Rgas, gas.gamma_gas, gas.Cp,
              gas.Prt]
    outs = [z(n), z(3 * n), z(n), z(n)]
    lam = [torch.randn(k * n, dtype=torch.float64, device="cuda") for k in (1, 3, 1, 1)]
    gout = [z(n), z(3 * n), z(n), z(64 * ne)]
    fwd = lambda: _lib.call("sod_full_rhs", ne, n, m.npoin_w, *common, *[o.data_ptr() for o in outs])
    adj = lambda: _lib.call("sod_full_rhs_vjp", ne, n, m.npoin_w, *common, *[x.data_ptr() for x in lam],
                            *[o.data_ptr() for o in gout])
    fwd()
    adj()
    torch.cuda.synchronize()
    _lib.call("sod_profile_enable", 1)
    _lib.call("sod_profile_reset")
    try:
        for _ in range(5):
            fwd()
            adj()
        torch.cuda.synchronize()
        t = {}
        for name in ("full_rhs", "full_rhs_vjp"):
            tot, cnt = ctypes.c_double(0.0), ctypes.c_int(0)
            _lib.call("sod_profile_get", name.encode(), ctypes.byref(tot), ctypes.byref(cnt))
            t[name] = tot.value / max(cnt.value, 1)
    finally:
        _lib.call("sod_profile_enable", 0)
    ratio = t["full_rhs_vjp"] / t["full_rhs"]
    print(f"[vjp cost] 32^3: forward {t['full_rhs']:.4f} ms, vjp {t['full_rhs_vjp']:.4f} ms, ratio {ratio:.2f}",
          flush=True)
    sol.close()
    assert ratio <= 3.0, ratio
