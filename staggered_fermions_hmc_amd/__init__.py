"""MI355X (gfx950) staggered-fermion HMC hot path.

libsfh.so (HIP, built from csrc/) exposes the C ABI of include/sfh.h;
``device`` wraps it in Python objects.  The reference's field-level
Fermions.h / Lattice.h calls are the C++ drop-in of compat/.
"""
from . import _abi  # noqa: F401

__all__ = ["device"]
