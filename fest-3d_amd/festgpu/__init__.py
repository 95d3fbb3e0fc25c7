"""festgpu -- host-side Python view of the MI355X FEST-3D explicit-residual
engine: synthetic cases (cases.py) and the ctypes binding of the C ABI in
include/fest_hip.h (native.py)."""
from . import cases  # noqa: F401
