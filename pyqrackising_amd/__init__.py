"""pyqrackising_amd: MI355X (gfx950) cut evaluation and bit-flip local search.

Drop-in for PyQrackIsing's maxcut_tfim / maxcut_tfim_sparse /
spin_glass_solver / spin_glass_solver_sparse (pyqrackising/__init__.py:5-11),
computed by hand-written HIP kernels in libpqi.so through a ctypes C ABI.
"""
from .solvers import (  # noqa: F401
    maxcut_tfim,
    maxcut_tfim_sparse,
    spin_glass_solver,
    spin_glass_solver_sparse,
)
from ._lib import DeviceGraph, PqiError, device_count, pack_bits  # noqa: F401

__all__ = [
    "maxcut_tfim",
    "maxcut_tfim_sparse",
    "spin_glass_solver",
    "spin_glass_solver_sparse",
    "DeviceGraph",
    "PqiError",
    "device_count",
    "pack_bits",
]
