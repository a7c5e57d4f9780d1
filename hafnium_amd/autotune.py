"""Autotuner of the exchange kernels' tile schedule.

For a bound tensor, an occupied width and a precision it times every
candidate tile of the half-transform (column tile, and for FP64/FP32 the
64- or 128-row tile) and every candidate SYRK split count with the
library's HIP-event stage timers, keeps the fastest of each
(the two kernels are independent: the split count only moves the SYRK, the
tile only the half-transform) and leaves them set on the context. Results are
cached in a JSON file keyed by device, precision and shape, so a later run
(or another process) reuses them without timing.

No counterpart in the reference (its J/K has no tunable kernels).
"""

from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass, field

from hafnium_amd import _lib

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "hafnium_amd", "tuning.json")
TILE_CANDIDATES = (32, 64, 128, 256)


@dataclass
class TuningResult:
    key: str
    x_bn: int
    syrk_splits: int
    x_ms: dict = field(default_factory=dict)
    syrk_ms: dict = field(default_factory=dict)
    seconds: float = 0.0
    from_cache: bool = False
    x_bm: int = 64


def _device_name(jk) -> str:
    """Architecture and compute-unit count (the marketing name reads differently under a profiler)."""
    import torch

    props = torch.cuda.get_device_properties(jk.device)
    arch = (getattr(props, "gcnArchName", "") or "gpu").split(":")[0]
    return f"{arch}x{props.multi_processor_count}"


def tuning_key(jk, nocc: int, precision: str) -> str:
    """Device, precision, shape and the storage of B the half-transform reads ("|full" when a full
    copy of that precision is bound: its kernels differ from the packed tile's)."""
    has_full = getattr(jk, "has_full", None)
    full = "|full" if has_full is not None and has_full(precision) else ""
    return f"{_device_name(jk)}|{precision}|nao={jk.n_ao}|naux={jk.n_aux}|nocc={nocc}{full}"


def _split_candidates(jk) -> list[int]:
    nb = (-(-jk.n_ao // 64) * 64 + 127) // 128
    pairs = nb * (nb + 1) // 2
    out = set()
    for target in (512, 1024, 2048, 4096):
        s = max(1, min(-(-target // pairs), jk.n_aux))
        out.add(-(-s // 8) * 8)
    return sorted(out)


def _stage_ms(stages, name):
    vals = [ms for n, ms in stages if n == name]
    return min(vals) if vals else float("inf")


def autotune(jk, orbitals, precision: str = "fp64", *, repeats: int = 2, cache_path: str | None = DEFAULT_CACHE,
             use_cache: bool = True) -> TuningResult:
    """Choose and set the fastest half-transform tile and SYRK split count.

    ``orbitals`` is a representative ``[nao, nocc]`` matrix (values do not
    affect the timing). Both are timed inside the combined J + K build
    (``jk_occupied``), where the Coulomb half runs beside the SYRK; among
    split counts within 1 % of the fastest the smallest is kept (fewer
    partials to reduce). The chosen parameters stay set on ``jk``.
    """
    import torch

    nocc = int(orbitals.shape[1])
    key = tuning_key(jk, nocc, precision)
    tile_param = _lib.TUNE_X16_BN if precision == "bf16x3" else _lib.TUNE_X_BN
    if use_cache and cache_path and os.path.exists(cache_path):
        with open(cache_path) as fh:
            cached = json.load(fh).get(key)
        if cached:
            jk.set_tuning(tile_param, cached["x_bn"])
            jk.set_tuning(_lib.TUNE_SYRK_SPLITS, cached["syrk_splits"])
            bm = int(cached.get("x_bm", 64))
            if precision != "bf16x3":
                jk.set_tuning(_lib.TUNE_X_BM, bm)
            return TuningResult(key, cached["x_bn"], cached["syrk_splits"], cached.get("x_ms", {}),
                                cached.get("syrk_ms", {}), 0.0, True, bm)

    start = time.perf_counter()
    ldo = -(-nocc // 32) * 32
    tiles = [(t, 64) for t in TILE_CANDIDATES if t // 2 < ldo or t == 32]
    if precision != "bf16x3":
        tiles += [(t, 128) for t, _ in tiles if t <= 128]
    splits = _split_candidates(jk)
    x_ms: dict = {}
    syrk_ms: dict[int, float] = {}
    jk.profile(True)
    try:
        jk.stage_times()
        # Half-transform tile, with the default split count.
        jk.set_tuning(_lib.TUNE_SYRK_SPLITS, 0)
        for tile in tiles:
            jk.set_tuning(tile_param, tile[0])
            if precision != "bf16x3":
                jk.set_tuning(_lib.TUNE_X_BM, tile[1])
            jk.jk_occupied(orbitals, precision=precision)  # warm-up
            torch.cuda.synchronize(jk.device)
            jk.stage_times()
            for _ in range(repeats):
                jk.jk_occupied(orbitals, precision=precision)
            x_ms[tile] = _stage_ms(jk.stage_times(), "exchange_half_transform")
        best_tile = min(x_ms, key=x_ms.get)
        jk.set_tuning(tile_param, best_tile[0])
        if precision != "bf16x3":
            jk.set_tuning(_lib.TUNE_X_BM, best_tile[1])
        for split in splits:
            jk.set_tuning(_lib.TUNE_SYRK_SPLITS, split)
            jk.jk_occupied(orbitals, precision=precision)
            torch.cuda.synchronize(jk.device)
            jk.stage_times()
            for _ in range(repeats):
                jk.jk_occupied(orbitals, precision=precision)
            stages = jk.stage_times()
            syrk_ms[split] = _stage_ms(stages, "exchange_syrk") + _stage_ms(stages, "exchange_reduce")
        fastest = min(syrk_ms.values())
        best_split = min(k for k, v in syrk_ms.items() if v <= 1.01 * fastest)
        jk.set_tuning(_lib.TUNE_SYRK_SPLITS, best_split)
    finally:
        jk.profile(False)
    # Tile labels: "BN" for 64-row tiles, "BNx128" for 128-row tiles.
    label = lambda t: str(t[0]) if t[1] == 64 else f"{t[0]}x{t[1]}"  # noqa: E731
    result = TuningResult(key, best_tile[0], best_split, {label(k): v for k, v in x_ms.items()},
                          {str(k): v for k, v in syrk_ms.items()}, time.perf_counter() - start, False, best_tile[1])
    if cache_path:
        _store(cache_path, key, result)
    return result


def _store(path: str, key: str, result: TuningResult) -> None:
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    data = {}
    if os.path.exists(path):
        try:
            with open(path) as fh:
                data = json.load(fh)
        except (OSError, ValueError):
            data = {}
    data[key] = {"x_bn": result.x_bn, "x_bm": result.x_bm, "syrk_splits": result.syrk_splits,
                 "x_ms": result.x_ms, "syrk_ms": result.syrk_ms}
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1)
    os.replace(tmp, path)
