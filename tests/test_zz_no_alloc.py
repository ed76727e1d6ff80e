"""The allocation-free entry point sod_smart_visc_spectral_noalloc against
sod_smart_visc_spectral, through the C ABI: two calls in a row, the second
on a smaller mesh, each equal bit for bit to the allocating entry point
(same kernels, same summation order), with the entropy normalisation on
(the branch that uses the scratch array) and off; and an oracle check of
the first call.  Captured into a HIP graph on a side stream, it records
only kernel launches: replays equal the eager allocating call bit for bit,
and unchecked table pointers are refused under capture."""
import ctypes
import warnings

import numpy as np
import pytest
import torch

import oracle
from _util import mesh_geometry, rel_l2, stage_fields
from sod2d_refmap_amd import _lib, cases, mesh as pmesh
from sod2d_refmap_amd.device import numerical_params, set_numerical_params, to_device, to_host

pytestmark = pytest.mark.gpu

T = oracle.hex64_tables()


def _inputs(n, seed):
    m = pmesh.periodic_box(n, curved_amp=0.04, perm_seed=seed)
    g = mesh_geometry(m)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas, perturb_seed=seed, perturb_amp=5e-2)
    rng = np.random.default_rng(seed)
    Reta = rng.standard_normal(m.npoin) * 1e-2
    return m, g, gas, s, Reta


def _call(entry, m, g, gas, s, Reta, keep):
    dev = lambda a: keep.append(to_device(a)) or keep[-1].data_ptr()
    out = lambda k: keep.append(torch.full((k,), np.nan, dtype=torch.float64, device="cuda")) or keep[-1]
    ne, npn = m.nelem, m.npoin
    mue, muel = out(64 * ne), out(64 * ne)
    _lib.call(entry, ne, npn, m.npoin_w, dev(m.connec), dev(m.lpoin_w), dev(Reta), dev(T.Ngp), dev(m.coord),
              dev(T.dNgp), dev(g["gpvol"]), dev(T.wgp), gas.gamma_gas, dev(s["rho"]), dev(s["u"]),
              dev(s["csound"]), dev(s["Tem"]), dev(s["eta"]), dev(g["helem_l"]), dev(g["helem"]), dev(g["Ml"]),
              mue.data_ptr(), dev(T.invAtoIJK), dev(T.gmshAtoI), dev(T.gmshAtoJ), dev(T.gmshAtoK), muel.data_ptr())
    torch.cuda.synchronize()
    return to_host(mue, (ne, 64)), to_host(muel, (ne, 64))


@pytest.mark.parametrize("normalise", [1, 0])
def test_noalloc_matches_allocating_entry_point(gpu, normalise):
    gas = cases.GasParams()
    p = numerical_params(Rgas_inf=gas.Rgas, flag_normalise_entropy=normalise)
    set_numerical_params(p)
    first = None
    # two calls in a row, the second on a smaller mesh (the scratch keeps the
    # first call's contents)
    for n, seed in ((4, 5), (2, 9)):
        m, g, gas, s, Reta = _inputs(n, seed)
        print(f"[noalloc] normalise={normalise} mesh {n}^3 nelem={m.nelem}", flush=True)
        keep = []
        a_mue, a_muel = _call("sod_smart_visc_spectral", m, g, gas, s, Reta, keep)
        b_mue, b_muel = _call("sod_smart_visc_spectral_noalloc", m, g, gas, s, Reta, keep)
        assert np.all(np.isfinite(b_mue)) and np.all(np.isfinite(b_muel))
        assert np.array_equal(a_mue, b_mue)
        assert np.array_equal(a_muel, b_muel)
        if first is None:
            first = (m, g, gas, s, Reta, b_mue, b_muel)
    m, g, gas, s, Reta, b_mue, b_muel = first
    mu_e_ref, mue_l_ref = oracle.smart_visc_spectral(m.connec, m.lpoin_w, Reta, g["gpvol"], T.wgp, gas.gamma_gas,
                                                     s["rho"], s["u"], s["csound"], s["Tem"], s["eta"],
                                                     g["helem_l"], g["helem"], T.invAtoIJK, p)
    assert rel_l2(b_mue, mu_e_ref) < 1e-12
    assert rel_l2(b_muel, mue_l_ref) < 1e-12


def test_noalloc_refuses_bad_sizes(gpu):
    # npoin_w > npoin: refused before any pointer is read
    args = [1, 1, 2] + [None] * 8 + [1.4] + [None] * 14
    assert _lib.lib().sod_smart_visc_spectral_noalloc(*args) == -1   # SOD_ERR_ARG


def test_noalloc_in_captured_graph(gpu):
    """sod_smart_visc_spectral_noalloc captured into a graph on a side stream
    (sod_set_stream, no synchronisation on return): three replays, with eager
    work on the default stream between them, each equal bit for bit to the
    allocating entry point called eagerly on the same inputs and tables; under
    capture, copies of the node-order tables at new addresses (never checked)
    are refused with SOD_ERR_ARG."""
    gas = cases.GasParams()
    set_numerical_params(numerical_params(Rgas_inf=gas.Rgas))
    m, g, gas, s, Reta = _inputs(3, 7)
    keep = []
    dev = lambda a: keep.append(to_device(a)) or keep[-1].data_ptr()
    ne, npn = m.nelem, m.npoin
    ins = [dev(m.connec), dev(m.lpoin_w), dev(Reta), dev(T.Ngp), dev(m.coord), dev(T.dNgp), dev(g["gpvol"]),
           dev(T.wgp), gas.gamma_gas, dev(s["rho"]), dev(s["u"]), dev(s["csound"]), dev(s["Tem"]), dev(s["eta"]),
           dev(g["helem_l"]), dev(g["helem"]), dev(g["Ml"])]
    tabs = [dev(T.invAtoIJK), dev(T.gmshAtoI), dev(T.gmshAtoJ), dev(T.gmshAtoK)]
    nan = lambda: torch.full((64 * ne,), np.nan, dtype=torch.float64, device="cuda")
    a_mue, a_muel, b_mue, b_muel = nan(), nan(), nan(), nan()
    args = lambda mue, muel, t: [ne, npn, m.npoin_w, *ins, mue.data_ptr(), *t, muel.data_ptr()]
    _lib.call("sod_smart_visc_spectral", *args(a_mue, a_muel, tabs))   # eager; checks the tables
    torch.cuda.synchronize()
    assert torch.isfinite(a_mue).all() and torch.isfinite(a_muel).all()
    side = torch.cuda.Stream()
    try:
        _lib.call("sod_set_stream", ctypes.c_void_p(side.cuda_stream))
        _lib.call("sod_set_sync", 0)
        graph = torch.cuda.CUDAGraph()
        print("[noalloc] capturing on a side stream", flush=True)
        with torch.cuda.graph(graph, stream=side):
            rc = _lib.lib().sod_smart_visc_spectral_noalloc(*args(b_mue, b_muel, tabs))
        assert rc == 0, _lib.lib().sod_last_error()
        for rep in range(3):
            b_mue.fill_(np.nan)
            b_muel.fill_(np.nan)
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            print(f"[noalloc] replay {rep}", flush=True)
            assert torch.equal(b_mue, a_mue) and torch.equal(b_muel, a_muel), rep
        fresh = [dev(T.invAtoIJK), dev(T.gmshAtoI), dev(T.gmshAtoJ), dev(T.gmshAtoK)]
        g2 = torch.cuda.CUDAGraph()
        with warnings.catch_warnings():   # the refused call leaves the graph empty
            warnings.simplefilter("ignore", UserWarning)
            with torch.cuda.graph(g2, stream=side):
                rc = _lib.lib().sod_smart_visc_spectral_noalloc(*args(b_mue, b_muel, fresh))
        assert rc == -1 and b"stream capture" in _lib.lib().sod_last_error()
    finally:
        _lib.call("sod_set_stream", None)
        _lib.call("sod_set_sync", 1)
        torch.cuda.synchronize()
