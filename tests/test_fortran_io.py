"""The Fortran binding of the HDF5 mesh/restart library
(sod2d_refmap_amd/fortran/mod_sod2d_io.f90): a Fortran program
(tests/fortran/io_driver.f90) loads each partition of a mesh file and a
restart through the module, and what it read must equal what the Python
reader reads; the restart it writes must load back in Python.  Host only."""
import os
import subprocess

import numpy as np
import pytest

from sod2d_refmap_amd import build as pbuild, cases, h5io, mesh as pmesh


def _f32(a):
    return np.asarray(a).astype(np.float32).astype(np.float64)


def _run(mesh_path, rank, rst_in, rst_out, dump):
    if not os.path.exists(pbuild.IO_DRIVER):
        pytest.fail(f"{pbuild.IO_DRIVER} is not built (run __graft_entry__.build())")
    r = subprocess.run([pbuild.IO_DRIVER, mesh_path, str(rank), rst_in, rst_out, dump], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ne, nn, nw, nper, istep = (int(v) for v in r.stdout.split()[:5])
    return ne, nn, nw, nper, istep


def _read_dump(path, ne, nn, nw, nper):
    b = open(path, "rb").read()
    o = 0

    def take(dtype, shape):
        nonlocal o
        n = int(np.prod(shape))
        a = np.frombuffer(b, dtype=dtype, count=n, offset=o)
        o += a.nbytes
        return a.reshape(shape, order="F")
    d = dict(co=take(np.int32, (ne, 64)), cw=take(np.int32, (ne, 64)), lw=take(np.int32, (nw,)),
             xyz=take(np.float64, (nn, 3)))
    d["ms"] = take(np.int32, (nper, 2)) if nper else np.zeros((0, 2), np.int32)
    for k, shp in (("rho", (nn,)), ("u", (nn, 3)), ("pr", (nn,)), ("E", (nn,)), ("mue", (ne, 64)),
                   ("mut", (ne, 64)), ("time", (1,))):
        d[k] = take(np.float64, shp)
    assert o == len(b)
    return d


@pytest.mark.parametrize("nranks", [1, 2])
def test_fortran_reads_what_python_reads(tmp_path, nranks):
    m = pmesh.periodic_box(3, curved_amp=0.05, perm_seed=9)
    gas = cases.GasParams()
    s = cases.tgv_initial_state(m, gas, perturb_seed=1, perturb_amp=1e-2)
    rng = np.random.default_rng(3)
    mpath = h5io.mesh_file_name(str(tmp_path) + "/", "box", nranks)
    h5io.write_mesh(m, mpath, nranks=nranks)
    parts = [h5io.read_mesh(mpath, r) for r in range(nranks)]
    rpath = str(tmp_path / "restart_in.h5")
    total = parts[0].info["total_nodes"]
    for r, p in enumerate(parts):
        g = p.globalIdSrl - 1
        st = {k: np.asfortranarray(s[k][g]) for k in ("rho", "u", "pr", "E")}
        st["mu_e"] = np.asfortranarray(rng.random((p.info["nelem"], 64)))
        st["mu_sgs"] = np.asfortranarray(rng.random((p.info["nelem"], 64)) * 1e-3)
        h5io.save_restart(rpath, p.mesh, st, time=0.75, istep=33, total_nodes=total,
                          node_start=p.info["rank_node_start"], create=(r == 0))
    for r, p in enumerate(parts):
        out = str(tmp_path / f"restart_out_{r}.h5")
        dump = str(tmp_path / f"dump_{r}.bin")
        ne, nn, nw, nper, istep = _run(mpath, r, rpath, out, dump)
        assert (ne, nn, nw, nper, istep) == (p.info["nelem"], p.info["npoin"], p.info["npoin_w"], p.info["nper"], 33)
        d = _read_dump(dump, ne, nn, nw, nper)
        assert np.array_equal(d["co"], p.mesh.connec_orig) and np.array_equal(d["cw"], p.mesh.connec)
        assert np.array_equal(d["lw"], p.mesh.lpoin_w) and np.array_equal(d["xyz"], p.mesh.coord)
        assert np.array_equal(d["ms"], p.masSla)
        f = h5io.load_restart_fields(rpath, p.mesh, masSla=p.masSla, node_start=p.info["rank_node_start"])
        for k, kk in (("rho", "rho"), ("u", "u"), ("pr", "pr"), ("E", "E"), ("mue", "mu_e"), ("mut", "mu_sgs")):
            assert np.array_equal(d[k], f[kk]), k
        assert d["time"][0] == f["time"] == float(np.float32(0.75))
        # the restart the Fortran program wrote: doubled density, next step
        g = h5io.load_restart_fields(out, p.mesh, masSla=p.masSla)
        assert g["istep"] == 34 and g["time"] == float(np.float32(1.75))
        assert np.array_equal(g["rho"], 2.0 * f["rho"])
        assert np.array_equal(g["u"], f["u"]) and np.array_equal(g["E"], f["E"])


def test_fortran_reports_library_errors(tmp_path):
    r = subprocess.run([pbuild.IO_DRIVER, str(tmp_path / "missing.hdf"), "0", "x", "y", str(tmp_path / "d")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "cannot open mesh file" in r.stderr
