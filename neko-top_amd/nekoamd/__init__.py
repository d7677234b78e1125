"""nekoamd: Python front end of the MI355X spectral-element operator core.

The compute path is libnekoamd.so (hand-written gfx950 HIP kernels behind the
C-ABI of include/nekoamd.h).  This package only binds it with ctypes for the
tests and the benchmark; it never computes a result itself.
"""
from ._lib import NekoAmdError, lib  # noqa: F401
from . import device  # noqa: F401
