"""The C-ABI library loads, and exports every entry point include/*.h
declares (no compute call: runs without a GPU)."""
import ctypes
import glob
import os
import re

import numpy as np

from sod2d_refmap_amd import _lib, h5io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


HEADER_LIB = {"sod2d_hip.h": _lib.LIB_PATH, "sod2d_io.h": h5io.LIB_PATH}


def declared_symbols(header: str | None = None):
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        if header is not None and os.path.basename(h) != header:
            continue
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?(?:int|void|char|double)\s*\**\s*(sod_\w+)\s*\(", txt, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert {"sod_full_convec_ijk", "sod_full_diffusion_ijk", "sod_rk_4_main", "sod_smart_visc_spectral",
            "sod_generic_scalar_convec_ijk"} <= names


def test_every_header_has_its_library():
    assert {os.path.basename(h) for h in glob.glob(os.path.join(ROOT, "include", "*.h"))} == set(HEADER_LIB)


def test_library_exports_every_declared_symbol():
    for header, path in HEADER_LIB.items():
        lib = ctypes.CDLL(path)
        missing = [n for n in sorted(declared_symbols(header)) if not hasattr(lib, n)]
        assert not missing, (header, missing)


def test_binding_covers_header():
    assert declared_symbols("sod2d_hip.h") <= set(_lib.exported_symbols())
    assert declared_symbols("sod2d_io.h") == set(h5io._SIGS)


def test_host_tables_without_gpu():
    from sod2d_refmap_amd.mesh import hex64_tables
    t = hex64_tables()
    assert sorted((t.invAtoIJK.ravel() - 1).tolist()) == list(range(64))
    assert np.all(np.abs(t.dlxigp_ip.sum(axis=2)) < 1e-14)


def test_errors_without_init():
    lib = _lib.lib()
    rc = lib.sod_init(4)
    assert rc == -4 and b"porder" in lib.sod_last_error()


def test_binding_argument_counts_match_header():
    """Every ctypes signature has as many arguments as the C prototype."""
    txt = ""
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt += re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
    bad = []
    for name, args in list(_lib._SIGS.items()) + [(k, v[0]) for k, v in h5io._SIGS.items()]:
        m = re.search(r"\b" + name + r"\s*\(([^;]*?)\)\s*;", txt, re.S)
        assert m, name
        params = m.group(1).strip()
        n = 0 if params in ("void", "") else len(params.split(","))
        if n != len(args):
            bad.append((name, n, len(args)))
    assert not bad, bad


def test_stage_configuration_arguments_are_checked():
    """sod_set_autotune / sod_set_stage_config only record a setting (no
    device work): unknown modes and configurations are refused with a
    message, and the default is restored."""
    import pytest
    for bad in (-1, 3, 16):
        with pytest.raises(_lib.Sod2dError, match="mode"):
            _lib.set_autotune(bad)
    for nw, asm in ((8, "pull"), (3, "pull"), (1, "atomic"), (2, "atomic"), (4, "colour"), (1, 3), (0, "colour")):
        with pytest.raises(_lib.Sod2dError, match="no configuration"):
            _lib.set_stage_config(nw, asm)
    for nw, asm in ((1, "pull"), (2, "pull"), (4, "pull"), (4, "atomic"), (1, "colour"), (0, "pull"), (1, False),
                    (4, True)):
        _lib.set_stage_config(nw, asm)
    for mode in (1, 2, 0):
        _lib.set_autotune(mode)


def test_missing_library_fails_loudly():
    """Without the HIP build every entry point raises: the package has no CPU
    fallback.  Run in a child process so that this process keeps its loaded
    library."""
    import subprocess
    import sys
    code = ("import sod2d_refmap_amd._lib as L\n"
            "L.LIB_PATH = '/nonexistent/libsod2d_hip.so'\n"
            "L._lib = None\n"
            "from sod2d_refmap_amd.solver import PeriodicSolver\n"
            "try:\n"
            "    L.call('sod_init', 0)\n"
            "except RuntimeError as e:\n"
            "    print('raised:', e)\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "raised:" in r.stdout and "is not built" in r.stdout
