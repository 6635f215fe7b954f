"""aphros_amd: MI355X-native pressure-Poisson solver for Aphros' `linsolver_symm`.

The product is libaphpois.so (HIP kernels for gfx950 behind the C ABI of
include/aphpois.h).  This package holds its ctypes binding (`_lib`), a Python
mirror of the reference's linear-solver plugin interface (`linear`), the
synthetic-system generator (`synth`), the batched/served front ends
(`batch`, `serve`) and the tile autotuner (`autotune`).
"""
from . import _lib  # noqa: F401
