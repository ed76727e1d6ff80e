"""ctypes binding of libsod2d_hip.so (the C ABI declared in include/sod2d_hip.h).

The product path always runs through this library: if it is missing the
import of any compute function raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsod2d_hip.so")

c_int, c_double, c_size_t, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p
P = c_void_p  # every array argument is passed as an address

SOD_OK = 0


class Sod2dError(RuntimeError):
    """A non-zero return code of the C ABI, with sod_last_error()'s message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libsod2d_hip error {code}: {msg}")
        self.code = code


class NumericalParams(ctypes.Structure):
    """mod_numerical_params.f90 module variables read by the path."""
    _fields_ = [
        ("ce", c_double), ("cmax", c_double), ("cglob", c_double), ("c_rho", c_double),
        ("c_ener", c_double), ("Rgas_inf", c_double), ("flag_normalise_entropy", c_int),
        ("flag_total_enthalpy", c_int), ("flag_rk_order", c_int), ("flag_les", c_int),
        ("flag_real_diff", c_int), ("flag_diff_suth", c_int),
    ]


_SIGS = {
    "sod_init": [c_int],
    "sod_finalize": [],
    "sod_set_stream": [c_void_p],
    "sod_set_sync": [c_int],
    "sod_synchronize": [],
    "sod_set_numerical_params": [ctypes.POINTER(NumericalParams)],
    "sod_get_numerical_params": [ctypes.POINTER(NumericalParams)],
    "sod_enter_data": [P, c_size_t, c_int],
    "sod_exit_data": [P, c_int],
    "sod_update_device": [P, c_size_t],
    "sod_update_host": [P, c_size_t],
    "sod_invalidate_mesh_cache": [],
    "sod_profile_enable": [c_int],
    "sod_profile_reset": [],
    "sod_profile_get": [ctypes.c_char_p, ctypes.POINTER(c_double), ctypes.POINTER(c_int)],
    "sod_hex64_tables": [P] * 10,
    "sod_structured_box_mesh": [c_int] * 3 + [c_double] * 4 + [P] * 6,
    "sod_elem_jacobian": [c_int, c_int] + [P] * 6,
    "sod_lumped_mass_spectral": [c_int, c_int, P, P, P],
    "sod_char_length": [c_int, c_int, P, P, P],
    "sod_char_length_spectral": [c_int, c_int, P, P, P],
    "sod_full_convec_ijk": [c_int, c_int] + [P] * 19,
    "sod_full_convec_ijk_H": [c_int, c_int] + [P] * 19,
    "sod_generic_scalar_convec_ijk": [c_int, c_int] + [P] * 15,
    "sod_full_diffusion_ijk": [c_int, c_int] + [P] * 9 + [c_double, c_double] + [P] * 11,
    "sod_sutherland_viscosity": [c_int, P, P, P],
    "sod_lumped_solver_scal": [c_int, c_int, P, P, P],
    "sod_lumped_solver_vect": [c_int, c_int, P, P, P],
    "sod_smart_visc_spectral": [c_int, c_int, c_int] + [P] * 8 + [c_double] + [P] * 14,
    "sod_smart_visc_spectral_noalloc": [c_int, c_int, c_int] + [P] * 8 + [c_double] + [P] * 14,
    "sod_adapt_dt_cfl": [c_int, c_int, P, P, P, P, c_double, P, P, P, P, P],
    "sod_volAvg_EK": [c_int, c_int, P, P, P, c_double, P, P, P],
    "sod_set_rk4_path": [c_int],
    "sod_full_rhs": [c_int] * 3 + [P] * 11 + [c_double] * 4 + [P] * 4,
    "sod_full_rhs_vjp": [c_int] * 3 + [P] * 11 + [c_double] * 4 + [P] * 8,
    "sod_smart_visc_spectral_vjp": [c_int] * 3 + [P] * 4 + [c_double] + [P] * 14,
    "sod_rk_4_main": ([c_int] * 7 + [P] * 9 + [c_int] + [P] * 8 + [c_double] + [P] * 2 +
                      [c_double] * 4 + [P] * 22),
    "sod_time_loop_create": [ctypes.POINTER(c_void_p)] + [c_int] * 3 + [P] * 8 + [c_double] * 4 + [P] * 14,
    "sod_time_loop_run": [c_void_p, c_int, c_int, c_double, c_double, c_int, ctypes.POINTER(c_double),
                          ctypes.POINTER(c_double), c_int, c_int, c_double, ctypes.POINTER(c_double), c_int, c_int],
    "sod_time_loop_destroy": [c_void_p],
    "sod_time_loop_create_batched": ([ctypes.POINTER(c_void_p), c_int, P, P, P] + [c_int] * 3 + [P] * 8 +
                                     [c_double] * 4 + [P] * 14),
    "sod_time_loop_run_batched": [c_void_p, c_int, c_double, c_double, c_int, P, P, c_int],
    "sod_set_autotune": [c_int],
    "sod_set_stage_config": [c_int, c_int],
    "sod_set_stage_config_lds": [c_int, c_int, c_int],
    "sod_stage_lds_layout": [c_void_p, ctypes.POINTER(c_int)],
    "sod_stage_element_order": [c_void_p, ctypes.POINTER(c_int)],
    "sod_element_colouring": [c_void_p, ctypes.POINTER(c_int), P, c_int],
    "sod_stage_config": [c_void_p, ctypes.POINTER(c_int), ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                         ctypes.POINTER(c_double)],
}

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `python sod2d_refmap_amd/build.py` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        lib_ = ctypes.CDLL(LIB_PATH)
        missing = [name for name in _SIGS if getattr(lib_, name, None) is None]
        if missing:
            raise RuntimeError(f"{LIB_PATH} lacks {', '.join(missing)}: it was built from older sources; "
                               "rebuild it with __graft_entry__.build()")
        for name, args in _SIGS.items():
            fn = getattr(lib_, name)
            fn.argtypes = args
            fn.restype = c_int
        lib_.sod_last_error.restype = ctypes.c_char_p
        lib_.sod_last_error.argtypes = []
        lib_.sod_default_numerical_params.restype = None
        lib_.sod_default_numerical_params.argtypes = [ctypes.POINTER(NumericalParams)]
        lib_.sod_deviceptr.restype = c_void_p
        lib_.sod_deviceptr.argtypes = [c_void_p]
        _lib = lib_
    return _lib


def check(rc: int) -> None:
    if rc != SOD_OK:
        raise Sod2dError(rc, lib().sod_last_error().decode(errors="replace"))


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


ASSEMBLY = {"pull": 0, "atomic": 1, "colour": 2}
_ASSEMBLY_NAMES = {v: k for k, v in ASSEMBLY.items()}
LDS_LAYOUT = {"swizzled": 0, "linear": 1}
_LDS_NAMES = {v: k for k, v in LDS_LAYOUT.items()}


def set_autotune(mode: int) -> None:
    """How sod_rk_4_main and the time loops configure the stage kernel
    (sod_set_autotune): 0 the default, 1 timed choice among the
    configurations with a fixed summation order (tiles with element-node
    stores, colour-ordered scatter), 2 the same with atomic scatter among the
    candidates."""
    call("sod_set_autotune", int(mode))


def set_stage_config(elems_per_workgroup: int, assembly, lds="swizzled") -> None:
    """Pins one stage configuration (sod_set_stage_config_lds).  assembly:
    "pull" (element-node stores), "atomic" or "colour", or the ABI's 0/1/2
    (a bool keeps its meaning of atomic scatter or not); lds: "swizzled" or
    "linear" (or 0/1), the LDS layout of the element's nodal values."""
    a = ASSEMBLY[assembly] if isinstance(assembly, str) else int(assembly)
    lay = LDS_LAYOUT[lds] if isinstance(lds, str) else int(lds)
    call("sod_set_stage_config_lds", int(elems_per_workgroup), a, lay)


def stage_config(loop=None) -> dict:
    """The stage configuration of a time loop's mesh, or (loop None) of the
    mesh last stepped through sod_rk_4_main (sod_stage_config)."""
    nw, at, md, ms, lay = c_int(), c_int(), c_int(), c_double(), c_int()
    call("sod_stage_config", loop, ctypes.byref(nw), ctypes.byref(at), ctypes.byref(md), ctypes.byref(ms))
    call("sod_stage_lds_layout", loop, ctypes.byref(lay))
    reo = c_int()
    call("sod_stage_element_order", loop, ctypes.byref(reo))
    return {"elems_per_workgroup": nw.value, "assembly": _ASSEMBLY_NAMES.get(at.value, str(at.value)),
            "atomic_scatter": at.value == 1, "lds": _LDS_NAMES.get(lay.value, str(lay.value)), "mode": md.value,
            "stage_ms": ms.value, "locality_order": bool(reo.value)}


def exported_symbols() -> list[str]:
    return sorted(list(_SIGS) + ["sod_last_error", "sod_default_numerical_params", "sod_deviceptr"])
