"""Host-side case builders: the separable matrix-product evaluation of the
Fourier velocity modes equals the direct einsum evaluation (same draws) to
rounding, on curved, permuted and anisotropic boxes."""
import numpy as np
import pytest

from sod2d_refmap_amd import cases, mesh as pmesh


@pytest.mark.parametrize("kw,seed", [(dict(n=3), 1), (dict(n=4, curved_amp=0.05, perm_seed=1), 2),
                                     (dict(n=(3, 2, 4), L=(2 * np.pi, 4 * np.pi, 2 * np.pi), perm_seed=3), 5)])
def test_fast_fourier_modes_equal_direct(kw, seed):
    m = pmesh.periodic_box(**kw)
    a = cases.fourier_velocity_perturbation(m, seed)
    b = cases.fourier_velocity_perturbation_fast(m, seed)
    assert a.shape == b.shape == (m.npoin, 3)
    assert np.abs(a).max() > 0.1
    assert np.abs(a - b).max() <= 1e-14 * np.abs(a).max()


def test_placeholder_state_shapes():
    m = pmesh.periodic_box(2)
    gas = cases.GasParams(Re=3200.0, Mach=0.2)
    s = cases.placeholder_state(m, gas)
    ref = cases.tgv_initial_state(m, gas)
    for k, v in ref.items():
        assert np.shape(s[k]) == np.shape(v), k
    assert np.all(s["mu_factor"] == gas.mu_factor)
