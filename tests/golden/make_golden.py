"""Writes tests/golden/tgv_curved2.npz: inputs and the oracle's outputs for
every routine on the path, on a 2^3 curved periodic box with permuted
element numbering (Taylor-Green state + seeded perturbation, seeded mu_e).

The reference itself cannot be built in this image, so the expected outputs
are the oracle's (oracle/sod2d_oracle.c, pinned by tests/test_oracle_*.py).
The fixture freezes them: tests/test_golden.py checks that the oracle still
reproduces it bit for bit and that the GPU library matches it to the parity
tolerances.  Regenerate only on purpose:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from sod2d_refmap_amd import cases, mesh  # noqa: E402

NAME = "tgv_curved2"
MESH = dict(n=2, curved_amp=0.05, perm_seed=17)
STATE = dict(perturb_seed=23, perturb_amp=5e-2)
MU_E_SEED = 29
NSTEPS = 10


def inputs():
    m = mesh.periodic_box(**MESH)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas, **STATE)
    s["mu_e"] = np.asfortranarray(np.random.default_rng(MU_E_SEED).uniform(0, 2e-3, (m.nelem, 64)))
    return m, gas, s


def outputs(m, gas, s):
    t = oracle.hex64_tables()
    out = {}
    gpvol, He = oracle.elem_jacobian(m.connec_orig, m.coord, t)
    out.update(gpvol=gpvol, He=He, Ml=oracle.lumped_mass_spectral(m.connec, gpvol, m.npoin),
               helem=oracle.char_length(m.connec_orig, m.coord),
               helem_l=oracle.char_length_spectral(m.connec_orig, m.coord))
    for tot, key in ((False, "convec"), (True, "convec_H")):
        r = oracle.full_convec_ijk(m.connec, He, gpvol, t, s["u"], s["q"], s["rho"], s["pr"], s["E"], tot)
        out.update({f"{key}_Rmass": r[0], f"{key}_Rmom": r[1], f"{key}_Rener": r[2]})
    feta = np.asfortranarray(s["u"] * s["eta"][:, None])
    out["scalar_Rconvec"] = oracle.generic_scalar_convec_ijk(m.connec, He, gpvol, t, feta, s["eta"], s["u"])
    zero = np.zeros((m.nelem, 64), order="F")
    r = oracle.full_diffusion_ijk(m.connec, He, gpvol, t, gas.Cp, gas.Prt, s["rho"], s["rho"], s["u"], s["Tem"],
                                  s["mu_fluid"], s["mu_e"], zero)
    out.update(diffu_Rmass=r[0], diffu_Rmom=r[1], diffu_Rener=r[2])
    Reta = np.asfortranarray(np.random.default_rng(MU_E_SEED + 1).uniform(-1e-2, 1e-2, m.npoin))
    p = oracle.default_params(Rgas_inf=gas.Rgas)
    mu_e, mue_l = oracle.smart_visc_spectral(m.connec, m.lpoin_w, Reta, gpvol, t.wgp, gas.gamma_gas, s["rho"],
                                             s["u"], s["csound"], s["Tem"], s["eta"], out["helem_l"],
                                             out["helem"], t.invAtoIJK, p)
    out.update(visc_Reta=Reta, visc_mu_e=mu_e, visc_mue_l=mue_l)
    dt = oracle.adapt_dt_cfl(m.connec, out["helem"], s["u"], s["csound"], 0.5, 0.5, s["mu_fluid"], zero, s["rho"])
    out["dt"] = np.array([dt])
    out["EK0"] = np.array([oracle.volAvg_EK(m.connec, gpvol, t.Ngp, 1.0, s["rho"], s["u"])])
    st = {k: v.copy(order="F") for k, v in s.items()}
    for step in range(NSTEPS):
        oracle.rk_4_main(m.connec, m.lpoin_w, He, gpvol, t, out["Ml"], out["helem"], out["helem_l"], dt,
                         gas.Rgas, gas.gamma_gas, gas.Cp, gas.Prt, st["mu_factor"], st, p)
        if step == 0 or step == NSTEPS - 1:
            tag = f"rk{step + 1}"
            for k in ("rho", "q", "E", "u", "pr", "Tem", "eta", "csound", "machno", "mu_fluid", "mu_e", "mue_l"):
                out[f"{tag}_{k}"] = st[k].copy(order="F")
    return out


def main():
    m, gas, s = inputs()
    out = outputs(m, gas, s)
    arrays = {f"in_{k}": v for k, v in s.items()}
    arrays.update(in_connec=m.connec, in_connec_orig=m.connec_orig, in_lpoin_w=m.lpoin_w, in_coord=m.coord)
    arrays.update({f"out_{k}": v for k, v in out.items()})
    path = os.path.join(HERE, NAME + ".npz")
    np.savez_compressed(path, **arrays)
    digest = hashlib.sha256(open(path, "rb").read()).hexdigest()
    meta = dict(mesh=MESH, state=STATE, mu_e_seed=MU_E_SEED, nsteps=NSTEPS, sha256=digest,
                gas=dict(gamma=gas.gamma_gas, Rgas=gas.Rgas, Cp=gas.Cp, Prt=gas.Prt))
    with open(os.path.join(HERE, NAME + ".json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(f"wrote {path} ({os.path.getsize(path)} bytes) sha256={digest}")


if __name__ == "__main__":
    main()
