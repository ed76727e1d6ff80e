"""The CFL time step regrouped by node equals the reference's element loop
bit for bit (CPU).

adapt_dt_cfl (mod_time_ops.f90:30-57) takes, per element, the maximum over
its nodes of |u|_inf + c and of (rho*mu_sgs + mu)/rho, forms
cfl*(h_e/7)/max and cfl*(h_e/7)^2/max, and the minimum over elements.  With
mu_sgs = 0 the device kernels evaluate the same expressions per node with
h = min over the elements holding the node of h_e.  Here that regrouping is
evaluated in numpy, in the kernels' operation order, and compared with the
oracle's element loop."""
import numpy as np
import pytest

import oracle
from _util import mesh_geometry
from sod2d_refmap_amd import cases, mesh as pmesh


def node_form(m, helem, u, c, mu, rho, cfl_c, cfl_d):
    hmin = np.full(m.npoin, np.inf)
    np.minimum.at(hmin, m.connec.ravel(order="F") - 1, np.repeat(helem, 64).reshape(m.nelem, 64, order="C")
                  .ravel(order="F"))
    held = np.isfinite(hmin)
    umag = np.maximum(np.maximum(np.abs(u[:, 0]), np.abs(u[:, 1])), np.abs(u[:, 2]))
    L3 = np.maximum(0.0, umag + c)
    hh = hmin / 7.0
    dt = cfl_c * hh / L3
    mx = np.maximum(0.0, (rho * 0.0 + mu) / rho)
    dt = np.minimum(dt, cfl_d * (hh * hh) / mx)
    return dt[held].min()


@pytest.mark.parametrize("kw", [dict(n=3), dict(n=4, curved_amp=0.05, perm_seed=9), dict(n=(3, 4, 2), L=(1.0, 2.5, 0.7))])
def test_node_form_equals_element_loop(kw):
    m = pmesh.periodic_box(**kw)
    g = mesh_geometry(m)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas, perturb_seed=4, perturb_amp=5e-2)
    rng = np.random.default_rng(21)
    zero = np.zeros((m.nelem, 64), order="F")
    limits = set()
    for trial in range(6):
        u = np.asfortranarray(s["u"] * rng.uniform(0.2, 3.0) + rng.normal(0, 0.3, s["u"].shape))
        # mu/rho from the flow's own value up to ~10: the diffusive limit binds in the later trials
        mu = np.asfortranarray(s["mu_fluid"] * rng.uniform(0.5, 2.0) * 10.0 ** (trial + 0.5 * (trial > 2)))
        helem = np.asfortranarray(g["helem"] * rng.uniform(0.7, 1.3, m.nelem))
        for cfl_c, cfl_d in ((0.5, 0.5), (0.95, 0.1)):
            ref = oracle.adapt_dt_cfl(m.connec, helem, u, s["csound"], cfl_c, cfl_d, mu, zero, s["rho"])
            conv = oracle.adapt_dt_cfl(m.connec, helem, u, s["csound"], cfl_c, 1e300, mu, zero, s["rho"])
            limits.add("diffusive" if ref < conv else "convective")
            got = node_form(m, helem, u, s["csound"], mu, s["rho"], cfl_c, cfl_d)
            assert got == ref, (trial, cfl_c, got, ref)
    assert limits == {"diffusive", "convective"}
