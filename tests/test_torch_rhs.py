"""The FP64 torch transcription of the element routines (tests/torch_rhs.py),
the reference for the adjoint, against the oracle (CPU)."""
import numpy as np
import torch

import oracle
import torch_rhs as TR
from _util import mesh_geometry, rel_l2, stage_fields
from sod2d_refmap_amd import cases, mesh as pmesh

T = oracle.hex64_tables()


def _case():
    m = pmesh.periodic_box(2, curved_amp=0.05, perm_seed=3)
    g = mesh_geometry(m)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas, perturb_seed=9, perturb_amp=5e-2)
    return m, g, gas, s


def test_torch_rhs_matches_oracle():
    m, g, gas, s = _case()
    geo = TR.Geometry(m.connec, g["He"], g["gpvol"], T)
    t = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float64)
    f = stage_fields(s, gas.gamma_gas, gas.Rgas)
    ref = oracle.full_convec_ijk(m.connec, g["He"], g["gpvol"], T, f["u"], s["q"], s["rho"], f["p"], f["h"])
    got = TR.full_convec(geo, t(f["u"]), t(s["q"]), t(s["rho"]), t(f["p"]), t(f["h"]))
    for a, b in zip(got, ref):
        assert rel_l2(a.numpy(), b) < 1e-12
    ref = oracle.generic_scalar_convec_ijk(m.connec, g["He"], g["gpvol"], T, f["feta"], f["eta"], f["u"])
    assert rel_l2(TR.scalar_convec(geo, t(f["feta"]), t(f["eta"]), t(f["u"])).numpy(), ref) < 1e-12
    rng = np.random.default_rng(4)
    mu_e = np.asfortranarray(rng.uniform(0, 2e-3, (m.nelem, 64)))
    mu_sgs = np.asfortranarray(rng.uniform(0, 1e-3, (m.nelem, 64)))
    mu = oracle.sutherland_viscosity(f["T"], s["mu_factor"])
    ref = oracle.full_diffusion_ijk(m.connec, g["He"], g["gpvol"], T, gas.Cp, gas.Prt, s["rho"], s["rho"], f["u"],
                                    f["T"], mu, mu_e, mu_sgs)
    got = TR.full_diffusion(geo, gas.Cp, gas.Prt, t(s["rho"]), t(s["rho"]), t(f["u"]), t(f["T"]), t(mu),
                            TR.lattice_from_elemnode(geo, mu_e), TR.lattice_from_elemnode(geo, mu_sgs))
    for a, b in zip(got, ref):
        assert rel_l2(a.numpy(), b) < 1e-12
    # the assembled stage RHS
    rm, rq, re, rt = TR.full_rhs(geo, t(s["rho"]), t(s["q"]), t(s["E"]), t(s["mu_factor"]),
                                 TR.lattice_from_elemnode(geo, mu_e), TR.lattice_from_elemnode(geo, mu_sgs),
                                 gas.gamma_gas, gas.Rgas, gas.Cp, gas.Prt)
    cm, cq, ce = oracle.full_convec_ijk(m.connec, g["He"], g["gpvol"], T, f["u"], s["q"], s["rho"], f["p"], f["h"])
    dm, dq, de = oracle.full_diffusion_ijk(m.connec, g["He"], g["gpvol"], T, gas.Cp, gas.Prt, s["rho"], s["rho"],
                                           f["u"], f["T"], mu, mu_e, mu_sgs)
    assert rel_l2(rm.numpy(), cm + dm) < 1e-12
    assert rel_l2(rq.numpy(), cq + dq) < 1e-12
    assert rel_l2(re.numpy(), ce + de) < 1e-12


def test_torch_smart_visc_matches_oracle():
    m, g, gas, s = _case()
    rng = np.random.default_rng(8)
    Reta = rng.uniform(-1e-2, 1e-2, m.npoin)
    p = oracle.default_params(Rgas_inf=gas.Rgas)
    ref_mu, ref_l = oracle.smart_visc_spectral(m.connec, m.lpoin_w, Reta, g["gpvol"], T.wgp, gas.gamma_gas, s["rho"],
                                               s["u"], s["csound"], s["Tem"], s["eta"], g["helem_l"], g["helem"],
                                               T.invAtoIJK, p)
    t = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float64)
    mu, ml = TR.smart_visc(m.connec, m.lpoin_w, t(Reta), g["gpvol"], T.wgp, gas.gamma_gas, t(s["rho"]), t(s["u"]),
                           t(s["csound"]), t(s["Tem"]), t(s["eta"]), g["helem_l"], g["helem"], T.invAtoIJK, gas.Rgas)
    assert rel_l2(ml.numpy(), ref_l) < 1e-13
    assert rel_l2(mu.numpy(), ref_mu) < 1e-13
