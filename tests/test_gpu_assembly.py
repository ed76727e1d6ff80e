"""The fused step's pull assembly and the drop-in prepared-mesh cache.

The stage kernel stores element-node residuals and the node update sums each
node's contributions in a fixed order, so a step is reproducible bit for bit.
The drop-in entry points keep up to 64 prepared meshes; stepping more meshes
than that in turn evicts the least recently used, and every result must
still match the oracle."""
import numpy as np
import pytest

import oracle
from _util import mesh_geometry, rel_l2
from sod2d_refmap_amd import _lib, cases, mesh as pmesh
from sod2d_refmap_amd.device import numerical_params, set_numerical_params

pytestmark = pytest.mark.gpu

T = oracle.hex64_tables()


def _oracle_step(m, g, gas, s, dt):
    s = {k: v.copy(order="F") for k, v in s.items()}
    p = oracle.default_params(Rgas_inf=gas.Rgas)
    oracle.rk_4_main(m.connec, m.lpoin_w, g["He"], g["gpvol"], T, g["Ml"], g["helem"], g["helem_l"], dt,
                     gas.Rgas, gas.gamma_gas, gas.Cp, gas.Prt, s["mu_factor"], s, p)
    return s


def test_step_is_reproducible(gpu):
    """Three steps from one saved device state, twice on the same solver
    (the geometry, whose lumped mass is assembled with atomics as in the
    reference, is then the same): bit-identical states."""
    import torch
    from sod2d_refmap_amd.solver import PeriodicSolver
    m = pmesh.periodic_box(4, curved_amp=0.05, perm_seed=3)
    gas = cases.GasParams()
    set_numerical_params(numerical_params(Rgas_inf=gas.Rgas))
    s = cases.tgv_initial_state(m, gas, perturb_seed=5, perturb_amp=2e-2)
    sol = PeriodicSolver(m, gas, s)
    dt = sol.adapt_dt()
    tensors = {k: v for k, v in vars(sol).items() if isinstance(v, torch.Tensor)}
    tensors.update({"lv." + k: v for k, v in sol.lv.items()})
    saved = {k: v.clone() for k, v in tensors.items()}
    outs = []
    for rep in range(2):
        print(f"[assembly] reproducibility run {rep}", flush=True)
        for k, v in tensors.items():
            v.copy_(saved[k])
        torch.cuda.synchronize()
        for _ in range(3):
            sol.rk_4_main(dt)
        outs.append(sol.state())
    sol.close()
    for k in ("rho", "q", "E", "eta", "mu_e", "mu_fluid"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_prepared_mesh_cache_eviction(gpu):
    """70 meshes (more than the 64 cached) stepped in turn through the
    drop-in rk_4_main, twice round: each step matches the oracle."""
    from sod2d_refmap_amd.solver import PeriodicSolver
    gas = cases.GasParams()
    set_numerical_params(numerical_params(Rgas_inf=gas.Rgas))
    sols, refs = [], []
    for i in range(70):
        m = pmesh.periodic_box(2, curved_amp=0.03 * (i % 3), perm_seed=i)
        s = cases.tgv_initial_state(m, gas, perturb_seed=i, perturb_amp=1e-2)
        sols.append((m, PeriodicSolver(m, gas, s), s))
    print(f"[assembly] {len(sols)} meshes, two rounds", flush=True)
    for rnd in range(2):
        for i, (m, sol, s) in enumerate(sols):
            g = mesh_geometry(m)
            dt = oracle.adapt_dt_cfl(m.connec, g["helem"], s["u"], s["csound"], 0.5, 0.5, s["mu_fluid"],
                                     np.zeros((m.nelem, 64), order="F"), s["rho"])
            ref = _oracle_step(m, g, gas, s, dt)
            sol.rk_4_main(dt)
            out = sol.state()
            for k in ("rho", "q", "E", "mu_e"):
                err = rel_l2(out[k], ref[k])
                assert err < 1e-12, (rnd, i, k, err)
            # the next round's oracle step starts from the solver's own state
            nxt = {k: np.asfortranarray(v) for k, v in out.items()}
            for k in ("mu_factor", "mu_sgs"):
                nxt[k] = s[k]
            sols[i] = (m, sol, nxt)
    for _, sol, _ in sols:
        sol.close()


def test_mesh_rewritten_in_place(gpu):
    """A mesh array rewritten at the same device address.  The He and gpvol
    of an element the prepared-mesh fingerprint does not sample (the 4096
    samples of 17^3 take elements floor(4913 k / 4096), which skips element
    6) are changed and sod_invalidate_mesh_cache announces it: the next step
    uses the new geometry.  helem of one element is changed without any call: the CFL
    step's cache hashes all of connec and helem and is rebuilt, so dt equals
    the oracle's bit for bit."""
    import torch
    from sod2d_refmap_amd.solver import PeriodicSolver
    m = pmesh.periodic_box(17)
    gas = cases.GasParams()
    set_numerical_params(numerical_params(Rgas_inf=gas.Rgas))
    s = cases.tgv_initial_state(m, gas, perturb_seed=4, perturb_amp=2e-2)
    g = mesh_geometry(m)
    assert 6 not in {(m.nelem * k) // 4096 for k in range(4096)}
    sol = PeriodicSolver(m, gas, s)
    dt = sol.adapt_dt()
    sol.rk_4_main(dt)   # builds the prepared mesh
    sol.load_state(s)
    e = 6
    g2 = {k: v.copy(order="F") for k, v in g.items()}
    g2["He"][:, :, :, e] *= 1.01
    g2["gpvol"][:, :, e] *= 0.99
    sol.He[576 * e: 576 * (e + 1)].mul_(1.01)
    sol.gpvol[64 * e: 64 * (e + 1)].mul_(0.99)
    torch.cuda.synchronize()
    _lib.call("sod_invalidate_mesh_cache")
    print("[assembly] in-place He/gpvol rewrite of element 6, 17^3", flush=True)
    ref = _oracle_step(m, g2, gas, s, dt)
    sol.rk_4_main(dt)
    out = sol.state()
    for k in ("rho", "q", "E", "mu_e"):
        err = rel_l2(out[k], ref[k])
        assert err < 1e-12, (k, err)
    # CFL step: helem of one element shrinks in place, no call
    sol.load_state(s)
    e = 4321
    h2 = g["helem"].copy()
    h2[e] *= 0.25
    sol.helem.copy_(torch.from_numpy(np.ascontiguousarray(g["helem"])).to("cuda"))
    sol.adapt_dt()   # cached h_min of the original helem
    sol.helem[e:e + 1].mul_(0.25)
    torch.cuda.synchronize()
    zero = np.zeros((m.nelem, 64), order="F")
    want = oracle.adapt_dt_cfl(m.connec, h2, s["u"], s["csound"], 0.5, 0.5, s["mu_fluid"], zero, s["rho"])
    old = oracle.adapt_dt_cfl(m.connec, g["helem"], s["u"], s["csound"], 0.5, 0.5, s["mu_fluid"], zero, s["rho"])
    assert want < old
    assert sol.adapt_dt() == want
    sol.close()
