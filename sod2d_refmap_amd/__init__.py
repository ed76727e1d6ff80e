"""MI355X-native implementation of SOD2D's explicit compressible
spectral-element RHS + RK4 step (HEX64, p=3, FP64) behind the reference's
routine signatures.  Compute runs in libsod2d_hip.so (HIP, gfx950)."""
from ._lib import LIB_PATH, Sod2dError, lib  # noqa: F401

__all__ = ["LIB_PATH", "Sod2dError", "lib"]
