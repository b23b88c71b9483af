"""MI355X (gfx950) multispin Metropolis sweep and parallel-tempering exchange.

The compute path is the HIP library ``libising_gpu.so`` behind the C ABI in
``include/ising_gpu.h``; ``_lib`` binds it with ctypes and raises if it is
missing.  ``instances`` builds the synthetic workloads.
"""
from . import instances  # noqa: F401

__all__ = ["instances", "load"]


def load():
    """Load the native library (raises ImportError if it has not been built)."""
    from . import _lib
    return _lib.load()
