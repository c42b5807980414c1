"""pdcs_amd: MI355X-native (gfx950) PDHG conic solver, a drop-in for the PDCS GPU path.

The compute path is libpdcs_hip.so (hand-written HIP kernels + C++ host
control code behind the C ABI of include/pdcs.h).  This package is the Python host
side: ctypes bindings (``_lib``), the reference-shaped solve entry point
(``solver.rpdhg_gpu_solve``) and the synthetic workload generators
(``instances``).
"""
from . import instances  # noqa: F401

__all__ = ["instances", "load", "solver"]


def load():
    """Load the HIP library; raises when it has not been built."""
    from . import _lib

    return _lib.load()


def __getattr__(name):
    if name == "solver":
        from . import solver

        return solver
    raise AttributeError(name)
