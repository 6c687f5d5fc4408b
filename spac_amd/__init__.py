"""spac_amd — MI355X-native Feynman–Kac path solver and PDDSparse row assembly.

The product is the native library spac_amd/lib/libspac_fk.so (HIP kernels for
gfx950 behind the C ABI of include/spac_fk.h, plus the C++ plugin surface under
spac_amd/cpp).  This Python package only loads that library for tests and
benchmarks.
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
