"""One-electron and density-fitting integrals of a molecule (``libhfjk.so``, host and GPU).

The producer of the SCF inputs: S, T, V (the reference's
``IntegralEngine.overlap/kinetic/nuclear_attraction``,
hafnium/integrals/engine.py:108-377) and the metric ``(P|Q)`` and
three-centre ``(P|uv)`` of ``DensityFitting.build``
(hafnium/integrals/density_fitting.py:58-85), and the exact four-centre
``(uv|wx)`` of small systems (the spherical atoms of the density guess),
computed by the library's McMurchie-Davidson engine: on host threads, or
for ``(P|Q)`` and ``(P|uv)`` also on the GPU (``metric_device``,
``three_center_device``), where the density fit then runs without a host
round trip.
"""

from __future__ import annotations

import ctypes

import numpy as np

from hafnium_amd import _lib
from hafnium_amd.molecule import BasisOnAtoms, Molecule


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def one_electron(basis: BasisOnAtoms, molecule: Molecule):
    """``(S, T, V)`` ``[n, n]``."""
    lib = _lib.load()
    n = basis.n_functions
    s, t, v = (np.empty((n, n)) for _ in range(3))
    charges = np.asarray(molecule.charges, dtype=np.float64)
    coords = np.ascontiguousarray(molecule.positions, dtype=np.float64)
    b = basis.c_struct()
    _lib.check(lib.hfjk_int1e(ctypes.byref(b), len(charges), _ptr(charges), _ptr(coords), _ptr(s), _ptr(t), _ptr(v)))
    return s, t, v


def metric(aux: BasisOnAtoms):
    """``(P|Q)`` ``[naux, naux]``."""
    lib = _lib.load()
    m = np.empty((aux.n_functions, aux.n_functions))
    b = aux.c_struct()
    _lib.check(lib.hfjk_int2c2e(ctypes.byref(b), _ptr(m)))
    return m


def three_center(aux: BasisOnAtoms, orbital: BasisOnAtoms):
    """``(P|uv)`` ``[naux, n, n]``."""
    lib = _lib.load()
    out = np.empty((aux.n_functions, orbital.n_functions, orbital.n_functions))
    a, o = aux.c_struct(), orbital.c_struct()
    _lib.check(lib.hfjk_int3c2e(ctypes.byref(a), ctypes.byref(o), _ptr(out)))
    return out


def _device_context(device):
    import torch

    from hafnium_amd.coulomb_exchange import _Context

    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    return dev, _Context(dev.index or 0)


def metric_device(aux: BasisOnAtoms, *, device=None):
    """``(P|Q)`` ``[naux, naux]`` computed on the GPU, as a device tensor."""
    import torch

    dev, ctx = _device_context(device)
    out = torch.empty((aux.n_functions, aux.n_functions), dtype=torch.float64, device=dev)
    b = aux.c_struct()
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(ctx.lib.hfjk_int2c2e_device(ctx.handle, ctypes.byref(b), ctypes.c_void_p(out.data_ptr()), stream))
    return out


def three_center_device(aux: BasisOnAtoms, orbital: BasisOnAtoms, *, device=None):
    """``(P|uv)`` ``[naux, n, n]`` computed on the GPU, as a device tensor."""
    import torch

    dev, ctx = _device_context(device)
    n = orbital.n_functions
    out = torch.empty((aux.n_functions, n, n), dtype=torch.float64, device=dev)
    a, o = aux.c_struct(), orbital.c_struct()
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(ctx.lib.hfjk_int3c2e_device(ctx.handle, ctypes.byref(a), ctypes.byref(o),
                                           ctypes.c_void_p(out.data_ptr()), stream))
    return out


def four_center(basis: BasisOnAtoms):
    """``(uv|wx)`` ``[n, n, n, n]`` in chemists' notation (the reference's ``two_electron.eri``,
    hafnium/integrals/two_electron.py:26-48, without screening). ``n^4`` doubles: for atoms and
    small molecules."""
    lib = _lib.load()
    n = basis.n_functions
    out = np.empty((n, n, n, n))
    b = basis.c_struct()
    _lib.check(lib.hfjk_int2e(ctypes.byref(b), _ptr(out)))
    return out


def water_from_fixture_positions(positions_bohr) -> Molecule:
    return Molecule((8, 1, 1), np.asarray(positions_bohr, dtype=np.float64))
