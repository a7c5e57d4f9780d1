"""MI355X-native density-fitted Coulomb/exchange (DF-J/K) Fock build.

A drop-in for the J/K contractions of maxnus/hafnium's SCF
(hafnium/scf/coulomb_exchange.py:53-190): hand-written gfx950 HIP kernels
behind the C ABI of ``include/hfjk.h`` (``libhfjk.so``), driven from Python
with PyTorch-ROCm owning the device buffers.
"""

__version__ = "0.1.0"

from hafnium_amd.errors import (  # noqa: F401
    BackendError,
    ConvergenceError,
    HafniumError,
    InputError,
    StructureError,
)
