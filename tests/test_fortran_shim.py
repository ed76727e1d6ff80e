"""The Fortran drop-in module (sod2d_refmap_amd/fortran/mod_sod2d_hip.f90):
a Fortran program calls the library through the reference's routine names
and dummy lists (tests/fortran/shim_driver.f90), and its results are
compared with the oracle.

The host-only part (p=3 tables, mesh generator) runs on the CPU; the path
part (init routines, element routines, CFL dt, one rk_4_main step, EK on
host arrays registered as OpenACC-style present data) needs a GPU.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from _util import mesh_geometry, rel_l2
from sod2d_refmap_amd import build as pbuild, cases, mesh as pmesh

DRIVER = pbuild.DRIVER
T = oracle.hex64_tables()


def _need_driver():
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built (run __graft_entry__.build())")


class Reader:
    def __init__(self, path):
        self.b = open(path, "rb").read()
        self.o = 0

    def take(self, dtype, shape):
        n = int(np.prod(shape))
        a = np.frombuffer(self.b, dtype=dtype, count=n, offset=self.o)
        self.o += a.nbytes
        return a.reshape(shape, order="F")

    def done(self):
        return self.o == len(self.b)


def test_fortran_tables_and_mesh(tmp_path):
    _need_driver()
    out = tmp_path / "tab.bin"
    subprocess.run([DRIVER, "tables", str(out)], check=True, timeout=60)
    r = Reader(out)
    i4, f8 = np.int32, np.float64
    assert np.array_equal(r.take(i4, (64,)), T.atoIJK)
    assert np.array_equal(r.take(f8, (64, 3)), T.xgp)
    assert np.array_equal(r.take(f8, (64,)), T.wgp)
    assert np.array_equal(r.take(f8, (64, 64)), T.Ngp)
    assert np.array_equal(r.take(f8, (3, 64, 64)), T.dNgp)
    assert np.abs(r.take(f8, (64, 3, 4)) - T.dlxigp_ip).max() <= 1e-12
    assert np.array_equal(r.take(i4, (4, 4, 4)), T.invAtoIJK)
    for ref in (T.gmshAtoI, T.gmshAtoJ, T.gmshAtoK):
        assert np.array_equal(r.take(i4, (64,)), ref)
    m = pmesh.periodic_box(2)
    assert np.array_equal(r.take(i4, (8, 64)), m.connec_orig)
    assert np.array_equal(r.take(i4, (8, 64)), m.connec)
    assert np.array_equal(r.take(i4, (216,)), m.lpoin_w)
    assert np.array_equal(r.take(f8, (343, 3)), m.coord)
    assert r.done()


@pytest.mark.gpu
def test_fortran_path(tmp_path, gpu):
    _need_driver()
    m = pmesh.periodic_box(3, curved_amp=0.05, perm_seed=5)
    g = mesh_geometry(m)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas, perturb_seed=3, perturb_amp=5e-2)
    rng = np.random.default_rng(21)
    s["mu_e"] = np.asfortranarray(rng.uniform(0, 2e-3, (m.nelem, 64)))
    dt = oracle.adapt_dt_cfl(m.connec, g["helem"], s["u"], s["csound"], 0.5, 0.5, s["mu_fluid"],
                             np.zeros((m.nelem, 64), order="F"), s["rho"])
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([m.nelem, m.npoin, m.npoin_w], np.int32).tofile(f)
        for a in (m.connec_orig, m.connec, m.lpoin_w):
            np.asfortranarray(a, np.int32).ravel(order="F").tofile(f)
        for a in (m.coord, s["rho"], s["u"], s["q"], s["pr"], s["E"], s["Tem"], s["e_int"], s["eta"],
                  s["csound"], s["machno"], s["mu_fluid"], s["mu_factor"], s["mu_e"]):
            np.asfortranarray(a, np.float64).ravel(order="F").tofile(f)
        np.array([gas.Rgas, gas.gamma_gas, gas.Cp, gas.Prt, dt], np.float64).tofile(f)
    print(f"[fortran driver] 3^3 curved mesh, npoin={m.npoin}", flush=True)
    p = subprocess.run([DRIVER, "path", str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    ne, npn = m.nelem, m.npoin
    r = Reader(fout)
    f8 = np.float64
    assert rel_l2(r.take(f8, (1, 64, ne)), g["gpvol"]) < 1e-14
    assert rel_l2(r.take(f8, (3, 3, 64, ne)), g["He"]) < 1e-13
    assert rel_l2(r.take(f8, (npn,)), g["Ml"]) < 1e-14
    np.testing.assert_allclose(r.take(f8, (ne,)), g["helem"], rtol=4e-16, atol=0)
    np.testing.assert_allclose(r.take(f8, (ne, 64)), g["helem_l"], rtol=4e-16, atol=0)
    conv = oracle.full_convec_ijk(m.connec, g["He"], g["gpvol"], T, s["u"], s["q"], s["rho"], s["pr"], s["E"])
    for shp, ref in zip(((npn,), (npn, 3), (npn,)), conv):
        assert rel_l2(r.take(f8, shp), ref) < 1e-12
    feta = np.asfortranarray(s["u"] * s["eta"][:, None])
    ref = oracle.generic_scalar_convec_ijk(m.connec, g["He"], g["gpvol"], T, feta, s["eta"], s["u"])
    assert rel_l2(r.take(f8, (npn,)), ref) < 1e-12
    diff = oracle.full_diffusion_ijk(m.connec, g["He"], g["gpvol"], T, gas.Cp, gas.Prt, s["rho"], s["rho"], s["u"],
                                     s["Tem"], s["mu_fluid"], s["mu_e"], np.zeros((ne, 64), order="F"))
    for shp, ref in zip(((npn,), (npn, 3), (npn,)), diff):
        assert rel_l2(r.take(f8, shp), ref) < 1e-12
    assert r.take(f8, (1,))[0] == dt
    ek = r.take(f8, (1,))[0]
    st = {k: v.copy(order="F") for k, v in s.items()}
    oracle.rk_4_main(m.connec, m.lpoin_w, g["He"], g["gpvol"], T, g["Ml"], g["helem"], g["helem_l"], dt, gas.Rgas,
                     gas.gamma_gas, gas.Cp, gas.Prt, st["mu_factor"], st, oracle.default_params(Rgas_inf=gas.Rgas))
    for k, shp in (("rho", (npn,)), ("q", (npn, 3)), ("E", (npn,)), ("mu_e", (ne, 64))):
        err = rel_l2(r.take(f8, shp), st[k])
        print(f"  fortran rk_4_main {k} rel_l2={err:.3e}", flush=True)
        assert err < 1e-12, (k, err)
    ek_ref = oracle.volAvg_EK(m.connec, g["gpvol"], T.Ngp, 1.0, st["rho"], st["u"])
    assert abs(ek / ek_ref - 1) < 1e-12
    assert r.done()
