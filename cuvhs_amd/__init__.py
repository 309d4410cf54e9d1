"""MI355X-native VHS RF decoder (drop-in for the cuVHS decode path).

The decoder is native code (cuvhs_amd/lib/libcuvhs.so, HIP for gfx950) behind
the C ABI in include/cuvhs.h; `native` holds its ctypes bindings and the
synthetic capture generator.
"""
from . import native  # noqa: F401
from .native import CuvhsError, Decoder, format_for, generate  # noqa: F401
