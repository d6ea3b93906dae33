"""trident_amd — MI355X-native (gfx950) local SpGEMM path of Trident.

The product is the HIP library libtrident_hip.so behind the C-ABI in
include/trident_hip.h; this package holds its sources (csrc/), the ctypes
front end (device.py) and the seeded operand generators (gen.py).
"""
from . import gen  # noqa: F401
from ._lib import TR_COLS, TR_ROWS, TridentError  # noqa: F401

__all__ = ["gen", "TR_ROWS", "TR_COLS", "TridentError", "device"]


def __getattr__(name):
    if name == "device":
        from . import device
        return device
    raise AttributeError(name)
