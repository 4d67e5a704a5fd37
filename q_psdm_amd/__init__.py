"""MI355X-native Q-compensated 3-D Kirchhoff PSDM imaging path.

libqpsdm.so (HIP kernels for gfx950 behind the C ABI in include/) does the
work; this package is its Python binding (lib), the host-side mirror of the
reference's imaging operator (engine) and the synthetic survey generators
(synth).
"""
from .lib import LIB_PATH, QpsdmError
from .lib import lib as load_library

__all__ = ["load_library", "LIB_PATH", "QpsdmError"]
