"""ctypes binding of ``libhfjk.so`` (the C ABI of ``include/hfjk.h``).

The library is loaded from this package's directory only. If it is missing
the import of any GPU entry point raises :class:`BackendError`: there is no
CPU path behind this package.
"""

from __future__ import annotations

import ctypes
import os
import threading

from hafnium_amd.errors import BackendError, InputError

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libhfjk.so")

FP64 = 0
FP32 = 1
BF16X3 = 2
PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16x3": BF16X3}

TUNE_X_BN = 1
TUNE_X16_BN = 2
TUNE_SYRK_SPLITS = 3
TUNE_X_BM = 4
TUNE_SYRK_KIND = 5
TUNE_X_FULL = 6
TUNE_GENERAL_CHUNK = 7
TUNE_SIDE_GROUPS = 8
TUNE_COULOMB_SERIAL = 9

OK = 0
ERR_INPUT = 1
ERR_BACKEND = 2
ERR_HIP = 3
ERR_WORKSPACE = 4
ERR_SOLVER = 5

_c_int = ctypes.c_int
_c_ll = ctypes.c_longlong
_c_size = ctypes.c_size_t
_c_u64 = ctypes.c_uint64
_c_u32 = ctypes.c_uint32
_vp = ctypes.c_void_p

# (name, restype, argtypes)
SIGNATURES = [
    ("hfjk_version", _c_int, []),
    ("hfjk_last_error", ctypes.c_char_p, []),
    ("hfjk_create", _c_int, [ctypes.POINTER(_vp), _c_int]),
    ("hfjk_destroy", _c_int, [_vp]),
    ("hfjk_packed_elems", _c_ll, [_c_int]),
    ("hfjk_pack", _c_int, [_vp, _vp, _c_int, _c_int, _vp, _vp, _vp]),
    ("hfjk_generate", _c_int, [_vp, _c_u64, _c_int, _c_int, _c_int, _vp, _vp]),
    ("hfjk_normal_matrix", _c_int, [_vp, _c_u64, _c_int, _c_int, _c_u32, _vp, _vp]),
    ("hfjk_split", _c_int, [_vp, _vp, _c_ll, _c_int, _vp, _vp]),
    ("hfjk_bind", _c_int, [_vp, _c_int, _c_int, _vp, _vp, _vp]),
    ("hfjk_workspace_bytes", _c_int, [_vp, _c_int, _c_int, ctypes.POINTER(_c_size)]),
    ("hfjk_coulomb", _c_int, [_vp, _vp, _vp, _c_int, _vp, _c_size, _vp]),
    ("hfjk_coulomb_charges", _c_int, [_vp, _vp, _vp, _c_int, _vp, _c_size, _vp]),
    ("hfjk_exchange_occ", _c_int, [_vp, _vp, _c_int, _vp, _vp, _c_int, _vp, _c_size, _vp]),
    ("hfjk_exchange_occ_noalloc", _c_int, [_vp, _vp, _c_int, _vp, _vp, _c_int, _vp]),
    ("hfjk_exchange_density", _c_int, [_vp, _vp, _vp, _c_int, _vp, _c_size, _vp]),
    ("hfjk_density_factor", _c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_size, _vp]),
    ("hfjk_exchange_general", _c_int, [_vp, _vp, _vp, _c_int, _vp, _c_size, _vp]),
    ("hfjk_general_workspace_bytes", _c_int, [_vp, _c_int, _c_int, ctypes.POINTER(_c_size)]),
    ("hfjk_jk_occ", _c_int, [_vp, _vp, _c_int, _vp, _vp, _c_int, _vp, _c_size, _vp]),
    ("hfjk_expand_full", _c_int, [_vp, _vp, _c_int, _c_int, _c_int, _vp, _vp]),
    ("hfjk_bind_full", _c_int, [_vp, _c_int, _vp]),
    ("hfjk_basis_nfunc", _c_int, [_vp, ctypes.POINTER(_c_int)]),
    ("hfjk_int1e", _c_int, [_vp, _c_int, _vp, _vp, _vp, _vp, _vp]),
    ("hfjk_int2c2e", _c_int, [_vp, _vp]),
    ("hfjk_int3c2e", _c_int, [_vp, _vp, _vp]),
    ("hfjk_int2c2e_device", _c_int, [_vp, _vp, _vp, _vp]),
    ("hfjk_int3c2e_device", _c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("hfjk_int2e", _c_int, [_vp, _vp]),
    ("hfjk_batch_create", _c_int, [_vp, _vp, _c_int, _c_int, _c_int, ctypes.POINTER(_vp)]),
    ("hfjk_batch_workspace_bytes", _c_int, [_vp, ctypes.POINTER(_c_size)]),
    ("hfjk_batch_sizes", _c_int, [_vp, ctypes.POINTER(_c_ll), ctypes.POINTER(_c_ll), ctypes.POINTER(_c_ll)]),
    ("hfjk_batch_bind", _c_int, [_vp, _vp, _c_size, _vp]),
    ("hfjk_batch_run", _c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("hfjk_batch_coulomb_charges", _c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("hfjk_batch_destroy", _c_int, [_vp]),
    ("hfjk_dgemm", _c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, ctypes.c_double, _vp, _c_int, _vp,
                            _c_int, ctypes.c_double, _vp, _c_int, _vp]),
    ("hfjk_dsyevd", _c_int, [_vp, _c_int, _vp, _vp, _vp, _vp, _vp]),
    ("hfjk_dpotrf", _c_int, [_vp, _c_int, _vp, _vp, _vp]),
    ("hfjk_dtrsm_lower", _c_int, [_vp, _c_int, _c_int, _vp, _vp, _vp]),
    ("hfjk_combine", _c_int, [_vp, _c_ll, ctypes.c_double, _vp, ctypes.c_double, _vp, ctypes.c_double, _vp, _vp,
                              _vp]),
    ("hfjk_antisymmetrize", _c_int, [_vp, _c_int, _vp, _vp, _vp]),
    ("hfjk_dot_workspace", _c_ll, []),
    ("hfjk_ddot", _c_int, [_vp, _c_ll, _vp, _vp, _vp, _vp, _vp]),
    ("hfjk_set_tuning", _c_int, [_vp, _c_int, _c_int]),
    ("hfjk_get_tuning", _c_int, [_vp, _c_int, ctypes.POINTER(_c_int)]),
    ("hfjk_profile", _c_int, [_vp, _c_int]),
    ("hfjk_profile_read", _c_int, [_vp, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_float), _c_int]),
    ("hfjk_dgemm_batched", _c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, ctypes.c_double, _vp, _c_int, _c_ll,
                                    _vp, _c_int, _c_ll, ctypes.c_double, _vp, _c_int, _c_ll, _c_int, _vp]),
    ("hfjk_dsyev_batched_work", _c_ll, [_c_int, _c_int]),
    ("hfjk_dsyev_batched", _c_int, [_vp, _c_int, _c_int, _vp, _vp, _vp, _vp, _c_int, _vp]),
    ("hfjk_eigh_lowest_batched_work", _c_ll, [_c_int, _c_int, _c_int]),
    ("hfjk_eigh_lowest_batched", _c_int, [_vp, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp]),
    ("hfjk_ddot_batched", _c_int, [_vp, _c_int, _c_ll, _vp, _c_ll, _vp, _c_ll, _vp, _vp]),
    ("hfjk_history_dots_batched", _c_int, [_vp, _c_int, _c_ll, _c_int, _vp, _c_int, _vp, _vp]),
    ("hfjk_diis_batched", _c_int, [_vp, _c_int, _c_int, _vp, _vp, _c_int, ctypes.POINTER(_c_int), _c_int, _vp, _vp]),
    ("hfjk_diis_combine_batched", _c_int, [_vp, _c_int, _c_ll, _c_int, _vp, _vp, _vp, _vp]),
    ("hfjk_antisymmetrize_batched", _c_int, [_vp, _c_int, _c_int, _vp, _vp, _vp]),
    ("hfjk_symmetrize_batched", _c_int, [_vp, _c_int, _c_ll, _vp, _vp]),
    ("hfjk_orthogonalizer_batched", _c_int, [_vp, _c_int, _c_int, _vp, _vp, ctypes.c_double, _vp, _vp, _vp]),
    ("hfjk_shift_discarded_batched", _c_int, [_vp, _c_int, _c_int, _vp, ctypes.c_double, _vp, _vp]),
    ("hfjk_mask_columns_batched", _c_int, [_vp, _c_int, _c_int, _c_int, _vp, _vp, _vp]),
    ("hfjk_select_batched", _c_int, [_vp, _c_int, _c_ll, _vp, _vp, _vp, _vp, _vp]),
    ("hfjk_copy_batched", _c_int, [_vp, _c_int, _vp, _vp, _vp, _c_ll, _vp, _c_int, _c_ll, _vp, _c_int, _vp]),
    ("hfjk_mlp_gemm", _c_int, [_vp, _c_int, _c_int, _c_int, _vp, _c_int, _c_int, _c_int, _vp, _vp, _c_int, _c_int,
                               _c_int, _vp, _vp, _c_int, _c_int, _vp, _vp]),
    ("hfjk_mlp_silu_backward", _c_int, [_vp, _c_ll, _vp, _vp, _vp, _vp]),
    ("hfjk_mlp_column_sum", _c_int, [_vp, _c_int, _c_int, _vp, _c_int, _vp, _vp]),
    ("hfjk_mlp_mse", _c_int, [_vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _vp, ctypes.c_float, _vp, _vp,
                              _vp]),
    ("hfjk_mlp_adam", _c_int, [_vp, _c_ll, _vp, _vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _vp]),
    ("hfjk_mlp_to_bf16", _c_int, [_vp, _c_ll, _vp, _vp, _vp]),
    ("hfjk_kpoint_workspace_bytes", _c_int, [_vp, _c_int, _c_int, _c_int, ctypes.POINTER(_c_size)]),
    ("hfjk_kpoint_coulomb", _c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_size,
                                     _vp]),
    ("hfjk_kpoint_exchange", _c_int, [_vp, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, ctypes.c_double, _vp, _vp,
                                      _vp, _c_size, _vp]),
]

_lock = threading.Lock()
_lib = None


def load() -> ctypes.CDLL:
    """Load ``libhfjk.so`` once; raise :class:`BackendError` when it is absent."""
    global _lib
    with _lock:
        if _lib is None:
            # One HIP runtime per process: PyTorch-ROCm ships its own
            # libamdhip64.so.7 / librocblas.so.5 / librocsolver.so.0 and
            # loads them by file name. Importing torch first registers those
            # sonames, so the library's dependencies bind to the same runtime;
            # loading the library first would bring in a second HIP runtime
            # and torch would then see no GPU.
            import torch  # noqa: F401

            if not os.path.exists(LIB_PATH):
                raise BackendError(
                    f"{LIB_PATH} is not built: run `python -m hafnium_amd._build` "
                    "(the DF-J/K path has no CPU fallback)"
                )
            lib = ctypes.CDLL(LIB_PATH)
            for name, restype, argtypes in SIGNATURES:
                fn = getattr(lib, name)
                fn.restype = restype
                fn.argtypes = argtypes
            _lib = lib
    return _lib


def check(status: int) -> None:
    """Map a status of the C ABI to the package's exceptions."""
    if status == OK:
        return
    message = load().hfjk_last_error().decode(errors="replace")
    if status == ERR_INPUT:
        raise InputError(message)
    if status == ERR_BACKEND:
        raise BackendError(message)
    raise BackendError(f"hfjk error {status}: {message}")


def precision_code(precision: str | int) -> int:
    if isinstance(precision, int):
        if precision not in PRECISIONS.values():
            raise InputError(f"unknown precision {precision}")
        return precision
    try:
        return PRECISIONS[precision]
    except KeyError:
        raise InputError(f"precision is one of {sorted(PRECISIONS)}, got {precision!r}") from None
