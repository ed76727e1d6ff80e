"""Device buffers for the C ABI.

torch is used only for device memory and streams: a reference array of any
Fortran shape lives on the GPU as a flat tensor holding its column-major
storage, and its data_ptr() is what the C ABI receives.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib

_DT = {np.dtype(np.float64): torch.float64, np.dtype(np.int32): torch.int32}


def to_device(a: np.ndarray, device="cuda") -> torch.Tensor:
    """Flat device tensor holding the column-major storage of `a`."""
    a = np.asarray(a)
    flat = np.ascontiguousarray(a.ravel(order="F"))
    return torch.from_numpy(flat).to(device)


def to_host(t: torch.Tensor, shape) -> np.ndarray:
    """Inverse of to_device: Fortran-ordered numpy array of `shape`."""
    return np.asfortranarray(t.detach().cpu().numpy().reshape(shape, order="F"))


def empty(shape, dtype=torch.float64, device="cuda") -> torch.Tensor:
    n = int(np.prod(shape)) if not np.isscalar(shape) else int(shape)
    return torch.empty(max(n, 1), dtype=dtype, device=device)


def zeros(shape, dtype=torch.float64, device="cuda") -> torch.Tensor:
    n = int(np.prod(shape)) if not np.isscalar(shape) else int(shape)
    return torch.zeros(max(n, 1), dtype=dtype, device=device)


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


_INITIALISED = False


def init(device: int = 0, sync: bool = True) -> None:
    """sod_init(3) on the current torch device, bound to torch's current stream."""
    global _INITIALISED
    torch.cuda.set_device(device)
    _lib.call("sod_init", 3)
    _lib.call("sod_set_stream", torch.cuda.current_stream().cuda_stream)
    _lib.call("sod_set_sync", 1 if sync else 0)
    _INITIALISED = True


def numerical_params(**over) -> _lib.NumericalParams:
    p = _lib.NumericalParams()
    _lib.lib().sod_default_numerical_params(p)
    for k, v in over.items():
        setattr(p, k, v)
    return p


def set_numerical_params(p: _lib.NumericalParams) -> None:
    _lib.call("sod_set_numerical_params", p)
