"""pyqrackising_amd: MI355X (gfx950) cut evaluation and bit-flip local search.

Drop-in for PyQrackIsing's maxcut_tfim / maxcut_tfim_sparse /
spin_glass_solver / spin_glass_solver_sparse (pyqrackising/__init__.py:5-11)
and the branch-and-bound exact solvers solve_maxcut_bnb(_sparse),
computed by hand-written HIP kernels in libpqi.so through a ctypes C ABI.
"""
from .solvers import (  # noqa: F401
    maxcut_tfim,
    maxcut_tfim_sparse,
    spin_glass_solver,
    spin_glass_solver_sparse,
)
from .bnb import solve_maxcut_bnb, solve_maxcut_bnb_sparse  # noqa: F401
from ._lib import DeviceGraph, PqiError, device_count, pack_bits  # noqa: F401

__all__ = [
    "maxcut_tfim",
    "maxcut_tfim_sparse",
    "spin_glass_solver",
    "spin_glass_solver_sparse",
    "solve_maxcut_bnb",
    "solve_maxcut_bnb_sparse",
    "DeviceGraph",
    "PqiError",
    "device_count",
    "pack_bits",
]
