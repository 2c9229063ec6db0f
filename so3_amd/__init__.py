"""so3_amd -- FP64 Wigner transforms on SO(3) for AMD Instinct MI355X (gfx950).

The transforms run in hand-written HIP kernels behind the reference's C API
(libso3_amd.so, headers in include/so3/).  This package is the Python host
side: parameter objects, the seeded synthetic-data generator and ctypes
wrappers mirroring the reference's Python binding (src/so3/bindings.pyx).
"""
from .params import *  # noqa: F401,F403
