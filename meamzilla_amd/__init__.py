"""MI355X-native evaluation of the MEAM/EAM spline fitting objective.

The objective (weighted squared errors of energies, forces and stresses over a
database of periodic cells) is computed by hand-written gfx950 HIP kernels in
the shared library ``libmeamz.so`` behind the C-ABI declared in
``include/meamz.h``.  This package holds the Python side: file formats, the
seeded workload generator and ctypes bindings of the C-ABI.
"""
from . import formats, gen  # noqa: F401

__all__ = ["formats", "gen"]
