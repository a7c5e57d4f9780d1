"""k-point density-fitted Coulomb and exchange on an MI355X.

Drop-ins for the reference's periodic builders (hafnium/scf/coulomb_exchange.py:193-394):

* :class:`KPointMI355XCoulombExchange` for ``KPointDensityFittedCoulombExchange``:
  the fitted tensor ``[kpt, kpt, aux, ao, ao]`` of momentum transfer ``k2 - k1``
  with its signature ``[kpt, kpt, aux]``, the k-point weights, the overlaps
  ``S(k)`` and the Madelung constant;
  ``J(k) = sum_P s[k, k, P] B[k, k, P] sum_k' w_k' <B[k', k', P], D(k')>`` and
  ``K(k) = sum_k' w_k' sum_P s[k, k', P] B[k, k', P] D(k') B[k, k', P]^H
  + xi S(k) D(k) S(k)``;
* :class:`KPointMI355XCoulomb` for ``KPointDensityFittedCoulomb``: the
  zero-transfer blocks ``[kpt, aux, ao, ao]`` only, J as above, and exchange
  refused with ``InputError``.

The tensors live in HBM as complex128; the contractions run in the HIP
kernels of ``libhfjk.so`` (``hfjk_kpoint_coulomb`` / ``hfjk_kpoint_exchange``):
the Coulomb charges and J as streaming complex reductions, the exchange of
each (k, k') block through its real form on the FP64 direct-LDS
half-transform and the signed SYRK. FP64 only.
"""

from __future__ import annotations

import ctypes
from typing import Any

import numpy as np

from hafnium_amd import _lib
from hafnium_amd.errors import InputError


def _torch():
    import torch

    return torch


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _device_tensor(x, dtype, device):
    torch = _torch()
    if isinstance(x, np.ndarray):
        x = torch.from_numpy(np.ascontiguousarray(x))
    return x.to(device=device, dtype=dtype).contiguous()


class _KPointBase:
    def _setup(self, device) -> None:
        from hafnium_amd.coulomb_exchange import _Context

        torch = _torch()
        if not torch.cuda.is_available():
            from hafnium_amd.errors import BackendError

            raise BackendError("no GPU is visible to PyTorch-ROCm")
        self._device = torch.device(device) if device is not None else torch.device("cuda", 0)
        self._ctx = _Context(self._device.index or 0)
        self._workspace = None

    def _stream(self) -> ctypes.c_void_p:
        return ctypes.c_void_p(_torch().cuda.current_stream(self._device).cuda_stream)

    def _workspace_ptr(self):
        torch = _torch()
        size = ctypes.c_size_t()
        _lib.check(self._ctx.lib.hfjk_kpoint_workspace_bytes(self._ctx.handle, self.n_kpt, self.n_aux, self.n_ao,
                                                             ctypes.byref(size)))
        need = max(int(size.value), 256)
        if self._workspace is None or self._workspace.numel() < need + 256:
            self._workspace = None
            self._workspace = torch.empty(need + 256, dtype=torch.uint8, device=self._device)
        base = self._workspace.data_ptr()
        aligned = (base + 255) & ~255
        return ctypes.c_void_p(aligned), self._workspace.numel() - (aligned - base)

    def _density_in(self, density):
        torch = _torch()
        shape = (self.n_kpt, self.n_ao, self.n_ao)
        if tuple(density.shape) != shape:
            raise InputError(f"the density has shape {tuple(density.shape)}, expected {list(shape)}")
        return _device_tensor(density, torch.complex128, self._device)

    @staticmethod
    def _out_like(value, template):
        return value.cpu().numpy() if isinstance(template, np.ndarray) else value

    @property
    def device(self):
        return self._device

    def _coulomb(self, density, zero_transfer: int):
        torch = _torch()
        d = self._density_in(density)
        out = torch.empty_like(d)
        ws, nbytes = self._workspace_ptr()
        _lib.check(self._ctx.lib.hfjk_kpoint_coulomb(
            self._ctx.handle, self.n_kpt, self.n_aux, self.n_ao, zero_transfer, _ptr(self._b), _ptr(self._sig),
            _ptr(self._w), _ptr(d), _ptr(out), ws, nbytes, self._stream()))
        return self._out_like(out, density)


class KPointMI355XCoulombExchange(_KPointBase):
    """Contractions with the fitted tensor over k-points and the exchange correction
    (the reference's ``KPointDensityFittedCoulombExchange``)."""

    def __init__(self, df_tensor: Any, signature: Any, weights: Any, overlap: Any, madelung: float, *,
                 device: Any = None) -> None:
        torch = _torch()
        shape = tuple(df_tensor.shape)
        n_kpt = shape[0] if len(shape) == 5 else -1
        n_ao = shape[3] if len(shape) == 5 else -1
        if len(shape) != 5 or shape[1] != n_kpt or shape[4] != n_ao:
            raise InputError("the fitted tensor has shape [kpt, kpt, aux, ao, ao]")
        if tuple(signature.shape) != shape[:3]:
            raise InputError("the signature has shape [kpt, kpt, aux]")
        if tuple(weights.shape) != (n_kpt,) or tuple(overlap.shape) != (n_kpt, n_ao, n_ao):
            raise InputError("weights have shape [kpt] and the overlap [kpt, ao, ao]")
        self._setup(device)
        self._b = _device_tensor(df_tensor, torch.complex128, self._device)
        self._sig = _device_tensor(signature, torch.float64, self._device)
        self._w = _device_tensor(weights, torch.float64, self._device)
        self._s = _device_tensor(overlap, torch.complex128, self._device)
        self.madelung = float(madelung)
        self._shape = shape

    @property
    def n_kpt(self) -> int:
        return int(self._shape[0])

    @property
    def n_aux(self) -> int:
        return int(self._shape[2])

    @property
    def n_ao(self) -> int:
        return int(self._shape[3])

    def coulomb(self, density: Any):
        """``J(k)`` for the total ``density[kpt, ao, ao]`` over the full k set (coulomb_exchange.py:279-288)."""
        return self._coulomb(density, 0)

    def exchange(self, density: Any):
        """``K(k)`` for one spin block ``density[kpt, ao, ao]`` (Hermitian per k) over the full k set
        (coulomb_exchange.py:290-297)."""
        torch = _torch()
        d = self._density_in(density)
        out = torch.empty_like(d)
        ws, nbytes = self._workspace_ptr()
        _lib.check(self._ctx.lib.hfjk_kpoint_exchange(
            self._ctx.handle, self.n_kpt, self.n_aux, self.n_ao, _ptr(self._b), _ptr(self._sig), _ptr(self._w),
            _ptr(self._s), self.madelung, _ptr(d), _ptr(out), ws, nbytes, self._stream()))
        return self._out_like(out, density)


class KPointMI355XCoulomb(_KPointBase):
    """The Coulomb contraction alone from the zero-transfer fitted tensor (the reference's
    ``KPointDensityFittedCoulomb``)."""

    def __init__(self, df_tensor: Any, signature: Any, weights: Any, *, device: Any = None) -> None:
        torch = _torch()
        shape = tuple(df_tensor.shape)
        if len(shape) != 4 or shape[2] != shape[3]:
            raise InputError("the fitted tensor has shape [kpt, aux, ao, ao]")
        if tuple(signature.shape) != (shape[1],):
            raise InputError("the signature has shape [aux]")
        if tuple(weights.shape) != (shape[0],):
            raise InputError("weights have shape [kpt]")
        self._setup(device)
        self._b = _device_tensor(df_tensor, torch.complex128, self._device)
        self._sig = _device_tensor(signature, torch.float64, self._device)
        self._w = _device_tensor(weights, torch.float64, self._device)
        self._shape = shape

    @property
    def n_kpt(self) -> int:
        return int(self._shape[0])

    @property
    def n_aux(self) -> int:
        return int(self._shape[1])

    @property
    def n_ao(self) -> int:
        return int(self._shape[2])

    def coulomb(self, density: Any):
        """``J(k)`` for the total ``density[kpt, ao, ao]`` (coulomb_exchange.py:375-383)."""
        return self._coulomb(density, 1)

    def exchange(self, density: Any):
        """Not available: the fitting holds the blocks of zero transfer only."""
        del density
        raise InputError(
            "this fitting holds the zero-transfer blocks for the Coulomb term alone; "
            "exchange needs the full k-point tensor (KPointMI355XCoulombExchange)"
        )
