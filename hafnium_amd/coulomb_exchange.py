"""Density-fitted Coulomb and exchange contractions on an MI355X.

:class:`MI355XCoulombExchange` satisfies the reference's ``CoulombExchange``
protocol (hafnium/scf/coulomb_exchange.py:53-71): ``n_ao``, ``coulomb(D)``
and ``exchange(D)`` on ``[ao, ao]`` arrays. It is the drop-in for
``DensityFittedCoulombExchange`` (coulomb_exchange.py:124-190) with the
fitted tensor held in HBM in tiled packed storage and every contraction
running in the HIP kernels of ``libhfjk.so``. PyTorch owns the buffers and
supplies the stream; no arithmetic of the path runs in PyTorch or on the
host.

Besides the protocol it offers ``exchange_occupied(C)``, the occupied-orbital
factorisation ``K = sum_P (B_P C)(B_P C)^T`` (coulomb_exchange.py:129-133)
that the package's own SCF driver calls with the occupied orbitals, and
``precision`` ("fp64", "fp32" or "bf16x3") selecting the arithmetic of the
contractions.

The assembly of the reference's ``coulomb_matrix`` / ``exchange_matrix``
(coulomb_exchange.py:397-518) for a molecular density in any of its spin
structures is :func:`coulomb_matrix`, :func:`exchange_matrix` and
:func:`coulomb_exchange_matrices` on a :class:`SpinDensity` (restricted:
J of the spin-traced ``2 D_a``, K of ``D_a``; unrestricted: J of
``D_a + D_b`` on both blocks, K per block; generalized: J of
``D_aa + D_bb`` on both diagonal spin blocks, K of each of the four spin
sub-blocks); :func:`rhf_coulomb_exchange` is the restricted case on a bare
block.
"""

from __future__ import annotations

import ctypes
from typing import Any

import numpy as np

from hafnium_amd import _lib
from hafnium_amd.errors import BackendError, InputError

__all__ = ["MI355XCoulombExchange", "rhf_coulomb_exchange", "SpinDensity", "coulomb_matrix", "exchange_matrix",
           "coulomb_exchange_matrices", "SYMMETRY_TOLERANCE"]

SYMMETRY_TOLERANCE = 1e-12
"""Largest ``max|B - B^T| / max|B|`` accepted for a fitted tensor."""


def _torch():
    try:
        import torch
    except ImportError as error:  # pragma: no cover - torch is part of the image
        raise BackendError("PyTorch-ROCm is required to own the device buffers") from error
    return torch


class _Context:
    """One ``hfjk_ctx`` per device, destroyed with the Python object."""

    def __init__(self, device_index: int) -> None:
        lib = _lib.load()
        handle = ctypes.c_void_p()
        _lib.check(lib.hfjk_create(ctypes.byref(handle), device_index))
        self.handle = handle
        self.lib = lib

    def __del__(self) -> None:
        handle = getattr(self, "handle", None)
        if handle is not None and handle.value:
            self.lib.hfjk_destroy(handle)
            self.handle = None


def _stream(torch, device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(tensor.data_ptr())


class MI355XCoulombExchange:
    """Contractions with a fitted tensor ``B[aux, ao, ao]`` resident in HBM.

    Parameters
    ----------
    df_tensor
        ``[aux, ao, ao]`` float64, symmetric in the two AO indices: a torch
        tensor (on a GPU or the host) or a NumPy array. It is packed into the
        tiled storage once; the caller's array is not kept.
    precision
        Arithmetic of the contractions: "fp64" (default), "fp32" or "bf16x3".
    device
        Torch device of the GPU; defaults to the tensor's GPU or ``cuda:0``.
    """

    def __init__(self, df_tensor: Any, *, precision: str = "fp64", device: Any = None) -> None:
        torch = _torch()
        shape = tuple(df_tensor.shape)
        if len(shape) != 3 or shape[1] != shape[2]:
            raise InputError("the fitted tensor has shape [aux, ao, ao]")
        if shape[0] < 1 or shape[1] < 1:
            raise InputError("the fitted tensor has shape [aux, ao, ao] with aux, ao > 0")
        if isinstance(df_tensor, np.ndarray):
            if np.iscomplexobj(df_tensor):
                raise BackendError("complex fitted tensors (k-point classes) are not implemented on MI355X")
            df_tensor = torch.from_numpy(np.ascontiguousarray(df_tensor, dtype=np.float64))
        if df_tensor.is_complex():
            raise BackendError("complex fitted tensors (k-point classes) are not implemented on MI355X")
        self._device = torch.device(device) if device is not None else (
            df_tensor.device if df_tensor.is_cuda else torch.device("cuda", 0)
        )
        if self._device.type != "cuda":
            raise BackendError("the DF-J/K path runs on a GPU device only")
        if not torch.cuda.is_available():
            raise BackendError("no GPU is visible to PyTorch-ROCm")
        self._init_common(shape[0], shape[1], precision)
        full = df_tensor.to(device=self._device, dtype=torch.float64).contiguous()
        stats = torch.zeros(2, dtype=torch.float64, device=self._device)
        lib = self._ctx.lib
        _lib.check(lib.hfjk_pack(self._ctx.handle, _ptr(full), self._naux, self._nao, _ptr(self._b64),
                                 _ptr(stats), self._stream()))
        asym, mag = (float(v) for v in stats.cpu())
        del full
        if asym > SYMMETRY_TOLERANCE * max(mag, 1e-300):
            raise InputError(
                f"the fitted tensor is not symmetric in its AO indices (max |B - B^T| = {asym:.3e})"
            )
        self._bind()

    @classmethod
    def synthetic(cls, seed: int, naux: int, nao: int, *, precision: str = "fp64", device: Any = None,
                  q0: int = 0) -> "MI355XCoulombExchange":
        """The seeded synthetic tensor of ``hafnium_amd.synthetic``, generated in HBM.

        Aux functions ``q0 .. q0 + naux - 1`` of the seed's tensor.
        """
        torch = _torch()
        self = cls.__new__(cls)
        self._device = torch.device(device) if device is not None else torch.device("cuda", 0)
        if not torch.cuda.is_available():
            raise BackendError("no GPU is visible to PyTorch-ROCm")
        if naux < 1 or nao < 1:
            raise InputError("naux and nao are positive")
        self._init_common(naux, nao, precision)
        _lib.check(self._ctx.lib.hfjk_generate(self._ctx.handle, seed, q0, naux, nao, _ptr(self._b64),
                                               self._stream()))
        self._bind()
        return self

    def _init_common(self, naux: int, nao: int, precision: str) -> None:
        torch = _torch()
        self._naux = int(naux)
        self._nao = int(nao)
        self._precision = _lib.precision_code(precision)
        self._ctx = _Context(self._device.index if self._device.index is not None else 0)
        self._packed = int(self._ctx.lib.hfjk_packed_elems(self._nao))
        self._b64 = torch.empty(self._naux * self._packed, dtype=torch.float64, device=self._device)
        self._b32 = None
        self._b16 = None
        self._full = {}  # precision code -> full row-major copy of B
        self._workspace = None

    def _stream(self) -> ctypes.c_void_p:
        return _stream(_torch(), self._device)

    def _bind(self) -> None:
        lib = self._ctx.lib
        _lib.check(lib.hfjk_bind(
            self._ctx.handle, self._naux, self._nao, _ptr(self._b64),
            _ptr(self._b32) if self._b32 is not None else None,
            _ptr(self._b16) if self._b16 is not None else None,
        ))
        for code, full in self._full.items():
            _lib.check(lib.hfjk_bind_full(self._ctx.handle, code, _ptr(full)))

    def expand_full(self, precision: str = "fp64") -> None:
        """Keep a full row-major copy of B in ``precision`` ("fp64", "fp32" or "bf16x3"; ``naux npad^2``
        elements, npad = 32 ceil(nao/32); for "bf16x3" two bf16 parts per element, each 32-column
        block of a row stored as the 32 hi parts, then the 32 lo parts).

        The half-transform of that precision then streams both of its operands
        straight into LDS (13 % faster in FP64, 7 % in FP32 and, on 256 x 256
        tiles past 128 occupied columns, 17 % in BF16x3 at the large
        configuration), at twice the memory of the packed copy. The
        non-symmetric exchange reads its rows from the FP64 copy too.
        ``drop_full`` releases it.
        """
        torch = _torch()
        code = _lib.precision_code(precision)
        if code in self._full:
            return
        npad = -(-self._nao // 32) * 32
        count = self._naux * npad * npad
        if code == _lib.BF16X3:
            full = torch.empty(2 * count, dtype=torch.int16, device=self._device)
        else:
            dtype = torch.float64 if code == _lib.FP64 else torch.float32
            full = torch.empty(count, dtype=dtype, device=self._device)
        _lib.check(self._ctx.lib.hfjk_expand_full(self._ctx.handle, _ptr(self._b64), self._naux, self._nao, code,
                                                  _ptr(full), self._stream()))
        self._full[code] = full
        _lib.check(self._ctx.lib.hfjk_bind_full(self._ctx.handle, code, _ptr(full)))

    def has_full(self, precision: str = "fp64") -> bool:
        """Whether :meth:`expand_full` holds a full copy of B in ``precision``."""
        return _lib.precision_code(precision) in self._full

    def drop_full(self, precision: str = "fp64") -> None:
        """Release the full copy of :meth:`expand_full`."""
        code = _lib.precision_code(precision)
        _lib.check(self._ctx.lib.hfjk_bind_full(self._ctx.handle, code, None))
        self._full.pop(code, None)

    def _ensure_copy(self, precision: int) -> None:
        """Derive the FP32 or BF16 hi/lo copy of B the first time a mode needs it."""
        torch = _torch()
        count = self._naux * self._packed
        lib = self._ctx.lib
        if precision == _lib.FP32 and self._b32 is None:
            self._b32 = torch.empty(count, dtype=torch.float32, device=self._device)
            _lib.check(lib.hfjk_split(self._ctx.handle, _ptr(self._b64), count, _lib.FP32, _ptr(self._b32),
                                      self._stream()))
            self._bind()
        if precision == _lib.BF16X3 and self._b16 is None:
            self._b16 = torch.empty(2 * count, dtype=torch.int16, device=self._device)
            _lib.check(lib.hfjk_split(self._ctx.handle, _ptr(self._b64), count, _lib.BF16X3, _ptr(self._b16),
                                      self._stream()))
            self._bind()

    def _workspace_for(self, nocc: int, precision: int):
        torch = _torch()
        size = ctypes.c_size_t()
        _lib.check(self._ctx.lib.hfjk_workspace_bytes(self._ctx.handle, nocc, precision, ctypes.byref(size)))
        need = max(int(size.value), 256)
        if self._workspace is None or self._workspace.numel() < need:
            self._workspace = None
            self._workspace = torch.empty(need + 256, dtype=torch.uint8, device=self._device)
        base = self._workspace.data_ptr()
        aligned = (base + 255) & ~255
        return ctypes.c_void_p(aligned), self._workspace.numel() - (aligned - base)

    # -------------------------------------------------------------- protocol

    @property
    def n_ao(self) -> int:
        """Number of atomic orbitals."""
        return self._nao

    @property
    def n_aux(self) -> int:
        """Number of auxiliary functions."""
        return self._naux

    @property
    def precision(self) -> str:
        return {v: k for k, v in _lib.PRECISIONS.items()}[self._precision]

    @precision.setter
    def precision(self, value: str) -> None:
        self._precision = _lib.precision_code(value)

    @property
    def device(self):
        return self._device

    def _density_parts(self, array: Any):
        """``(real, imaginary)`` parts of a density as device float64 ``[ao, ao]`` tensors; the
        imaginary part is None for a real density."""
        torch = _torch()
        n = self._nao
        if isinstance(array, np.ndarray):
            array = torch.from_numpy(np.ascontiguousarray(
                array, dtype=np.complex128 if np.iscomplexobj(array) else np.float64))
        if tuple(array.shape) != (n, n):
            raise InputError(f"the density has shape {tuple(array.shape)}, the integrals act on {n} atomic orbitals")
        if array.is_complex():
            z = array.to(device=self._device, dtype=torch.complex128)
            return z.real.contiguous(), z.imag.contiguous()
        return array.to(device=self._device, dtype=torch.float64).contiguous(), None

    @staticmethod
    def _out_like(value, template):
        if isinstance(template, np.ndarray):
            return value.cpu().numpy()
        return value

    def _coulomb_real(self, d, prec: int):
        torch = _torch()
        out = torch.empty((self._nao, self._nao), dtype=torch.float64, device=self._device)
        ws, nbytes = self._workspace_for(1, prec)
        _lib.check(self._ctx.lib.hfjk_coulomb(self._ctx.handle, _ptr(d), _ptr(out), prec, ws, nbytes,
                                              self._stream()))
        return out

    def coulomb(self, density: Any, *, precision: str | None = None):
        """``J = sum_P B[P] <B[P], D>`` (coulomb_exchange.py:176-182).

        Any real or complex density: J depends on the symmetric part of D
        only (B[P] is symmetric), and a complex density gives
        ``J[Re D] + i J[Im D]``, real for a Hermitian one up to rounding.
        """
        torch = _torch()
        prec = self._precision if precision is None else _lib.precision_code(precision)
        re, im = self._density_parts(density)
        self._ensure_copy(prec)
        out = self._coulomb_real(re, prec)
        if im is not None:
            out = torch.complex(out, self._coulomb_real(im, prec))
        return self._out_like(out, density)

    def coulomb_from_charges(self, charges, *, precision: str | None = None):
        """``J = sum_P charges[P] B[P]`` for a device float64 vector of naux charges (no reference
        counterpart; the Coulomb matrix of a predicted fitted density)."""
        torch = _torch()
        prec = self._precision if precision is None else _lib.precision_code(precision)
        if not (isinstance(charges, torch.Tensor) and charges.dtype == torch.float64
                and charges.numel() == self._naux and charges.device == self._device):
            raise InputError(f"charges are a float64 tensor of {self._naux} values on {self._device}")
        q = charges.contiguous()
        self._ensure_copy(prec)
        out = torch.empty((self._nao, self._nao), dtype=torch.float64, device=self._device)
        ws, nbytes = self._workspace_for(1, prec)
        _lib.check(self._ctx.lib.hfjk_coulomb_charges(self._ctx.handle, _ptr(q), _ptr(out), prec, ws, nbytes,
                                                      self._stream()))
        return out

    GENERAL_WORKSPACE_LIMIT = 8 << 30
    """Bytes of workspace the non-symmetric exchange may use (more aux functions per pass)."""

    def _general_workspace(self, prec: int):
        """A workspace for the non-symmetric exchange: all aux functions in one pass when that fits
        ``GENERAL_WORKSPACE_LIMIT``, else as many as fit (at least one)."""
        torch = _torch()
        lib = self._ctx.lib
        size = ctypes.c_size_t()
        _lib.check(lib.hfjk_general_workspace_bytes(self._ctx.handle, prec, 0, ctypes.byref(size)))
        need = int(size.value)
        if need > self.GENERAL_WORKSPACE_LIMIT:
            _lib.check(lib.hfjk_general_workspace_bytes(self._ctx.handle, prec, 1, ctypes.byref(size)))
            need = max(int(size.value), self.GENERAL_WORKSPACE_LIMIT)
        if self._workspace is None or self._workspace.numel() < need + 256:
            self._workspace = None
            self._workspace = torch.empty(need + 256, dtype=torch.uint8, device=self._device)
        base = self._workspace.data_ptr()
        aligned = (base + 255) & ~255
        return ctypes.c_void_p(aligned), self._workspace.numel() - (aligned - base)

    def _exchange_real(self, d, prec: int):
        """K of one real density: the signed factor of a symmetric D, then the occupied-form
        exchange with a workspace sized for the retained rank; the FP64 density form otherwise."""
        torch = _torch()
        n = self._nao
        lib = self._ctx.lib
        factor = torch.empty(n * n, dtype=torch.float64, device=self._device)
        signs = torch.empty(n, dtype=torch.float64, device=self._device)
        rank, negative = ctypes.c_int(), ctypes.c_int()
        ws, nbytes = self._workspace_for(1, _lib.FP64)
        _lib.check(lib.hfjk_density_factor(self._ctx.handle, _ptr(d), _ptr(factor), _ptr(signs), ctypes.byref(rank),
                                           ctypes.byref(negative), ws, nbytes, self._stream()))
        r = int(rank.value)
        out = torch.empty((n, n), dtype=torch.float64, device=self._device)
        if r < 0:
            ws, nbytes = self._general_workspace(prec)
            _lib.check(lib.hfjk_exchange_general(self._ctx.handle, _ptr(d), _ptr(out), prec, ws, nbytes,
                                                 self._stream()))
            return out
        if r == 0:
            return out.zero_()
        ws, nbytes = self._workspace_for(r, prec)
        # All signs +1 (a positive semidefinite density): the unsigned kernels.
        _lib.check(lib.hfjk_exchange_occ(self._ctx.handle, _ptr(factor), r, _ptr(signs) if negative.value else None,
                                         _ptr(out), prec, ws, nbytes, self._stream()))
        return out

    def exchange(self, density: Any, *, precision: str | None = None):
        """``K = sum_P B[P] D B[P]`` (coulomb_exchange.py:184-190) for any real or complex density.

        A symmetric density is factorised on the GPU (``D = C_+ C_+^T -
        C_- C_-^T``) and the occupied-form kernels run on the factors in the
        chosen precision, so indefinite and fractionally occupied densities
        give the density-form result. A non-symmetric real density D = S + A
        goes through the Hermitian matrix S + i A: its eigenvectors give one
        half-transform and two SYRKs (B S B and B A B) in the chosen
        precision. A complex density gives
        ``K[Re D] + i K[Im D]``: for a Hermitian one the real part is
        symmetric and the imaginary part antisymmetric.
        """
        torch = _torch()
        prec = self._precision if precision is None else _lib.precision_code(precision)
        re, im = self._density_parts(density)
        self._ensure_copy(prec)
        out = self._exchange_real(re, prec)
        if im is not None:
            out = torch.complex(out, self._exchange_real(im, prec))
        return self._out_like(out, density)

    def _orbitals_in(self, orbitals: Any):
        """Real ``[ao, occ]`` orbitals as a device float64 tensor."""
        torch = _torch()
        if isinstance(orbitals, np.ndarray):
            if np.iscomplexobj(orbitals):
                raise BackendError("complex orbitals are not implemented on MI355X; pass the density to exchange()")
            orbitals = torch.from_numpy(np.ascontiguousarray(orbitals, dtype=np.float64))
        if orbitals.is_complex():
            raise BackendError("complex orbitals are not implemented on MI355X; pass the density to exchange()")
        if orbitals.ndim != 2 or orbitals.shape[0] != self._nao:
            raise InputError(f"the orbitals have shape {tuple(orbitals.shape)}, expected [{self._nao}, occ]")
        return orbitals.to(device=self._device, dtype=torch.float64).contiguous()

    def exchange_occupied(self, orbitals: Any, signs: Any = None, *, precision: str | None = None):
        """``K = sum_P (B[P] C) diag(signs) (B[P] C)^T`` for ``orbitals`` ``[ao, occ]``."""
        torch = _torch()
        prec = self._precision if precision is None else _lib.precision_code(precision)
        template = orbitals
        c = self._orbitals_in(orbitals)
        nocc = int(c.shape[1])
        s = None
        if signs is not None:
            s = torch.as_tensor(np.asarray(signs, dtype=np.float64) if isinstance(signs, (list, np.ndarray))
                                else signs).to(device=self._device, dtype=torch.float64).contiguous()
            if s.shape != (nocc,):
                raise InputError("one sign per orbital column")
        self._ensure_copy(prec)
        out = torch.empty((self._nao, self._nao), dtype=torch.float64, device=self._device)
        ws, nbytes = self._workspace_for(max(nocc, 1), prec)
        _lib.check(self._ctx.lib.hfjk_exchange_occ(
            self._ctx.handle, _ptr(c), nocc, _ptr(s) if s is not None else None, _ptr(out), prec, ws, nbytes,
            self._stream()))
        return self._out_like(out, template)

    def jk_occupied(self, orbitals: Any, *, precision: str | None = None):
        """``(J, K)`` of ``D = C C^T`` for ``orbitals`` ``[ao, occ]`` in one call.

        ``coulomb_exchange_matrices`` (coulomb_exchange.py:397-404) for an
        aufbau density: equal to ``(coulomb(C C^T), exchange_occupied(C))``,
        with the Coulomb charges formed from the half-transformed ``X = B C``
        so the fitted tensor is read once less.
        """
        torch = _torch()
        prec = self._precision if precision is None else _lib.precision_code(precision)
        template = orbitals
        c = self._orbitals_in(orbitals)
        nocc = int(c.shape[1])
        self._ensure_copy(prec)
        j = torch.empty((self._nao, self._nao), dtype=torch.float64, device=self._device)
        k = torch.empty((self._nao, self._nao), dtype=torch.float64, device=self._device)
        ws, nbytes = self._workspace_for(max(nocc, 1), prec)
        _lib.check(self._ctx.lib.hfjk_jk_occ(self._ctx.handle, _ptr(c), nocc, _ptr(j), _ptr(k), prec, ws, nbytes,
                                             self._stream()))
        return self._out_like(j, template), self._out_like(k, template)

    # --------------------------------------------------------------- tuning

    def set_tuning(self, param: int, value: int) -> None:
        """Set a kernel tuning parameter (``_lib.TUNE_*``; 0 restores the default)."""
        _lib.check(self._ctx.lib.hfjk_set_tuning(self._ctx.handle, int(param), int(value)))

    def get_tuning(self, param: int) -> int:
        out = ctypes.c_int()
        _lib.check(self._ctx.lib.hfjk_get_tuning(self._ctx.handle, int(param), ctypes.byref(out)))
        return int(out.value)

    # --------------------------------------------------------------- timing

    def profile(self, enable: bool = True) -> None:
        """Record HIP events around every kernel of later calls."""
        _lib.check(self._ctx.lib.hfjk_profile(self._ctx.handle, 1 if enable else 0))

    def stage_times(self) -> list[tuple[str, float]]:
        """``(stage, milliseconds)`` of every kernel stage since the last read.

        Read from the HIP events the library records on the call's stream
        while profiling is on; reading synchronises on them and clears them.
        """
        cap = 4096
        names = (ctypes.c_char_p * cap)()
        ms = (ctypes.c_float * cap)()
        count = self._ctx.lib.hfjk_profile_read(self._ctx.handle, names, ms, cap)
        if count < 0:
            _lib.check(-count)
        return [(names[i].decode(), float(ms[i])) for i in range(min(count, cap))]

    def packed_tensor(self, precision: str = "fp64"):
        """The resident tiled packed tensor copy (for tests and diagnostics)."""
        prec = _lib.precision_code(precision)
        self._ensure_copy(prec)
        return {_lib.FP64: self._b64, _lib.FP32: self._b32, _lib.BF16X3: self._b16}[prec]


def rhf_coulomb_exchange(two_body, density_alpha):
    """``(J[2 D_a], K[D_a])`` of a real restricted density block.

    The molecular RHF branch of the reference's ``coulomb_matrix`` (J of the
    spin trace ``2 D_a``, coulomb_exchange.py:407-427) and ``exchange_matrix``
    (K of the stored block, one spin sub-block: :444-462, :504-509).
    """
    return two_body.coulomb(2.0 * density_alpha), two_body.exchange(density_alpha)


RESTRICTED = "restricted"
UNRESTRICTED = "unrestricted"
GENERALIZED = "generalized"
_BLOCKS = {RESTRICTED: 1, UNRESTRICTED: 2, GENERALIZED: 1}


class SpinDensity:
    """A Hermitian molecular density in one of the three collinear/spinor structures.

    The molecular (no k axis, no irrep sectors) cases of the reference's
    ``BlockMatrix`` densities (structured/matrix.py): ``kind`` "restricted"
    holds ``(D_a,)`` (real, each spatial orbital doubly occupied, so the spin
    trace is ``2 D_a``); "unrestricted" holds ``(D_a, D_b)``; "generalized"
    holds one ``[2n, 2n]`` spinor block ordered (alpha, beta). Blocks are NumPy
    arrays or torch tensors, real or complex.
    """

    def __init__(self, kind: str, blocks) -> None:
        if kind not in _BLOCKS:
            raise InputError("the spin structure is restricted, unrestricted or generalized")
        blocks = tuple(blocks)
        if len(blocks) != _BLOCKS[kind]:
            raise InputError(f"a {kind} density holds {_BLOCKS[kind]} block(s), got {len(blocks)}")
        shapes = {tuple(b.shape) for b in blocks}
        if len(shapes) != 1 or len(next(iter(shapes))) != 2 or blocks[0].shape[0] != blocks[0].shape[1]:
            raise InputError("density blocks are square and of one shape")
        if kind == GENERALIZED and blocks[0].shape[0] % 2:
            raise InputError("a spinor block has an even dimension (alpha and beta)")
        self.kind = kind
        self.blocks = blocks

    @property
    def n_spatial(self) -> int:
        n = int(self.blocks[0].shape[0])
        return n // 2 if self.kind == GENERALIZED else n

    def spin_trace(self):
        """The total spatial density (structured/matrix.py:535-567 for the molecular classes)."""
        n = self.n_spatial
        if self.kind == RESTRICTED:
            return 2.0 * self.blocks[0]
        if self.kind == UNRESTRICTED:
            return self.blocks[0] + self.blocks[1]
        d = self.blocks[0]
        return d[:n, :n] + d[n:, n:]


def _real(x):
    if isinstance(x, np.ndarray):
        return np.ascontiguousarray(np.real(x), dtype=np.float64)
    torch = _torch()
    return (x.real if x.is_complex() else x).to(torch.float64).contiguous()


def _check_basis(two_body, density: SpinDensity) -> None:
    if density.n_spatial != two_body.n_ao:
        raise InputError(f"the density has {density.n_spatial} atomic orbitals, the integrals {two_body.n_ao}")


def _spinor_diagonal(j):
    """``[[J, 0], [0, J]]`` in the array type of ``j``."""
    if isinstance(j, np.ndarray):
        return np.kron(np.eye(2), j)
    torch = _torch()
    return torch.kron(torch.eye(2, dtype=j.dtype, device=j.device), j)


def coulomb_matrix(two_body, density: SpinDensity):
    """The Coulomb matrix in the structure of ``density`` (coulomb_exchange.py:407-427).

    J of the real part of the spin trace (the imaginary part of a Hermitian
    density is antisymmetric and drops out), placed on every stored block:
    the one restricted block, both collinear blocks, or both diagonal spin
    blocks of the spinor block.
    """
    _check_basis(two_body, density)
    j = two_body.coulomb(_real(density.spin_trace()))
    if density.kind == RESTRICTED:
        return SpinDensity(RESTRICTED, (j,))
    if density.kind == UNRESTRICTED:
        return SpinDensity(UNRESTRICTED, (j, j))
    return SpinDensity(GENERALIZED, (_spinor_diagonal(j),))


def _exchange_block(two_body, block, n: int):
    """K of one stored block, spin sub-block by spin sub-block (coulomb_exchange.py:504-518)."""
    m = int(block.shape[-1]) // n
    if m == 1:
        return two_body.exchange(block)
    rows = [[two_body.exchange(block[s * n:(s + 1) * n, t * n:(t + 1) * n]) for t in range(m)] for s in range(m)]
    if isinstance(block, np.ndarray):
        return np.block(rows)
    torch = _torch()
    return torch.cat([torch.cat(r, dim=1) for r in rows], dim=0)


def exchange_matrix(two_body, density: SpinDensity):
    """The exchange matrix in the structure of ``density`` (coulomb_exchange.py:444-462): K of
    each stored block, each spin sub-block of a spinor block contracted on its own."""
    _check_basis(two_body, density)
    n = density.n_spatial
    return SpinDensity(density.kind, tuple(_exchange_block(two_body, b, n) for b in density.blocks))


def coulomb_exchange_matrices(two_body, density: SpinDensity):
    """``(J, K)`` in the structure of ``density`` (coulomb_exchange.py:397-404)."""
    return coulomb_matrix(two_body, density), exchange_matrix(two_body, density)
