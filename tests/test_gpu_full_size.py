"""Parity at the full sizes of the benchmark workloads, through the C ABI, on an MI355X.

* Large single build (nao 1500, naux 5000, nocc 250, seed 1): the timed path
  (full copies of B, the stored tile table, the combined J + K call) on an
  8-aux slice against the oracle; the FP32 and BF16x3 builds of all 5000 aux
  functions (from their full copies, as timed) against the FP64 build; the
  FP32 exchange error flat over the SYRK split counts.
* The 512-instance batch (nao ~ U{40..400}, naux = 3.5 nao): every instance
  in FP64 and BF16x3 against a single-instance build of the same seed, 16
  instances (the 8 largest and 8 seeded picks) against the oracle.
* A 12-water cluster (cc-pVDZ, def2-universal-jkfit): the FP64 GPU SCF
  against the oracle SCF on the same integrals, and the batched driver
  against the serial one.
"""

from __future__ import annotations

import json
import os

import numpy as np
import pytest

from hafnium_amd import synthetic
from oracle import jk_ref, layout_ref, scf_ref
from tests.tolerances import BF16X3_REL_TOL, FP32_REL_TOL, FP64_REL_TOL

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TUNING = os.path.join(ROOT, "hafnium_amd", "data", "tuning_large_gfx950.json")
LARGE = {"seed": 1, "nao": 1500, "naux": 5000, "nocc": 250}
TOL = {"fp64": FP64_REL_TOL, "fp32": FP32_REL_TOL, "bf16x3": BF16X3_REL_TOL}
# FP32-MFMA at the large configuration: the SYRK keeps every FP32 accumulator
# to at most 4096 products (partials reduced in FP64), so K and E_J - E_K stay
# within 1e-6 of FP64 whatever the split count.
FP32_LARGE_TOL = 1e-6


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available(), "the gpu tests need an MI355X"
    return torch


def rel(a, b):
    return jk_ref.max_relative_error(np.asarray(a), np.asarray(b))


def apply_stored_tuning(jk, precision):
    """The tile schedule the benchmark's large-configuration entry stores for ``precision``."""
    from hafnium_amd import _lib

    full = "|full" if jk.has_full(precision) else ""
    with open(TUNING) as fh:
        entry = json.load(fh)[f"gfx950x256|{precision}|nao=1500|naux=5000|nocc=250{full}"]
    jk.set_tuning(_lib.TUNE_X16_BN if precision == "bf16x3" else _lib.TUNE_X_BN, entry["x_bn"])
    if precision != "bf16x3":
        jk.set_tuning(_lib.TUNE_X_BM, entry["x_bm"])
    jk.set_tuning(_lib.TUNE_SYRK_SPLITS, entry["syrk_splits"])


def test_timed_large_path_on_an_aux_slice_matches_the_oracle(torch):
    """The benchmark's kernels (direct-LDS half-transform from the full copy of B of each
    precision, tuned tiles, combined call with the Coulomb half beside the SYRK) at nao 1500 /
    nocc 250."""
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    n, o, q0, naux = LARGE["nao"], LARGE["nocc"], 2711, 8
    print(f"large slice: nao={n} nocc={o} aux {q0}..{q0 + naux - 1}", flush=True)
    b = synthetic.tensor_slice_full(LARGE["seed"], n, q0, q0 + naux)
    c = synthetic.occupied_orbitals(LARGE["seed"], n, o)
    d = c @ c.T
    j_ref = jk_ref.coulomb(b, d)
    k_ref = jk_ref.exchange_occupied(b, c)
    jk = MI355XCoulombExchange.synthetic(LARGE["seed"], naux, n, q0=q0)
    for precision in ("fp64", "fp32", "bf16x3"):
        jk.expand_full(precision)
        apply_stored_tuning(jk, precision)
        j, k = jk.jk_occupied(torch.from_numpy(c).cuda(), precision=precision)
        err_j, err_k = rel(j.cpu(), j_ref), rel(k.cpu(), k_ref)
        print(f"  {precision}: J {err_j:.2e} K {err_k:.2e}", flush=True)
        assert err_j < TOL[precision]
        assert err_k < TOL[precision]
    # The density-form exchange (the reference's K) at the same shape, in FP64.
    assert rel(jk.exchange(d, precision="fp64"), jk_ref.exchange(b, d)) < FP64_REL_TOL


def test_full_512_instance_batch(torch):
    """Every instance of the batched workload against a single-instance build, 16 against the oracle."""
    from hafnium_amd.batch import BatchedCoulombExchange
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    shapes = synthetic.batch_shapes(512)
    print(f"batch: {len(shapes)} instances, {sum(q for _, _, q, _ in shapes)} aux functions", flush=True)
    bt = BatchedCoulombExchange([(n, q, o) for _, n, q, o in shapes])
    bt.generate([s for s, _, _, _ in shapes])
    cs = [synthetic.occupied_orbitals(s, n, o) for s, n, _, o in shapes]
    ds = [x @ x.T for x in cs]
    c, d = bt.pack_orbitals(cs), bt.pack_matrices(ds)
    results = {}
    for precision in ("fp64", "bf16x3"):
        j, k = bt.run(c, d, precision=precision)
        results[precision] = (j.cpu().numpy(), k.cpu().numpy())
        print(f"  batched {precision} done", flush=True)
        del j, k
        bt.release(precision)  # its workspace (and a derived tensor copy)
        torch.cuda.empty_cache()
    del bt, c, d
    torch.cuda.empty_cache()

    def instance(flat, i):
        n = shapes[i][1]
        return flat[cum[i]:cum[i] + n * n].reshape(n, n)

    cum = np.cumsum([0] + [n * n for _, n, _, _ in shapes])
    order = sorted(range(len(shapes)), key=lambda i: -shapes[i][1])
    rng = np.random.default_rng(20261203)
    checked = set(order[:8]) | set(int(i) for i in rng.choice(order[8:], 8, replace=False))
    worst = {"fp64": 0.0, "bf16x3": 0.0}
    for i, (seed, n, q, o) in enumerate(shapes):
        single = MI355XCoulombExchange.synthetic(seed, q, n)
        j1, k1 = single.jk_occupied(cs[i], precision="fp64")
        for precision, (jf, kf) in results.items():
            err = max(rel(instance(jf, i), j1), rel(instance(kf, i), k1))
            worst[precision] = max(worst[precision], err)
            assert err < (1e-12 if precision == "fp64" else BF16X3_REL_TOL), (i, seed, n, q, o, precision)
        if i in checked:
            b = layout_ref.unpack(single.packed_tensor("fp64").view(q, -1).cpu().numpy(), n)
            assert rel(instance(results["fp64"][0], i), jk_ref.coulomb(b, ds[i])) < FP64_REL_TOL
            assert rel(instance(results["fp64"][1], i), jk_ref.exchange_occupied(b, cs[i])) < FP64_REL_TOL
            assert rel(instance(results["bf16x3"][1], i), jk_ref.exchange_occupied(b, cs[i])) < BF16X3_REL_TOL
            if i == order[0]:
                assert rel(instance(results["fp64"][1], i), jk_ref.exchange(b, ds[i])) < FP64_REL_TOL
        del single
    print(f"  worst per-instance difference against single builds: {worst}", flush=True)


@pytest.fixture(scope="module")
def large(torch):
    """The whole large configuration in HBM, its FP64 build the comparator of the reduced modes
    (after the batch test, which needs most of the HBM)."""
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    n, q, o = LARGE["nao"], LARGE["naux"], LARGE["nocc"]
    print(f"large: generating nao={n} naux={q}", flush=True)
    jk = MI355XCoulombExchange.synthetic(LARGE["seed"], q, n)
    c = torch.from_numpy(synthetic.occupied_orbitals(LARGE["seed"], n, o)).cuda()
    d = (c @ c.T).cpu().numpy()
    j, k = jk.jk_occupied(c, precision="fp64")
    yield jk, c, d, j.cpu().numpy(), k.cpu().numpy()
    del jk
    torch.cuda.empty_cache()


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_reduced_modes_against_fp64_at_the_large_configuration(torch, large, precision):
    jk, c, d, j64, k64 = large
    jk.expand_full(precision)  # as the benchmark runs it
    apply_stored_tuning(jk, precision)
    j, k = jk.jk_occupied(c, precision=precision)
    jk.drop_full(precision)
    j, k = j.cpu().numpy(), k.cpu().numpy()
    ej, ek = jk_ref.energies(d, j, k)
    ej0, ek0 = jk_ref.energies(d, j64, k64)
    err = {"J": rel(j, j64), "K": rel(k, k64), "E": abs((ej - ek) - (ej0 - ek0)) / abs(ej0 - ek0)}
    print(f"large {precision}: {err}", flush=True)
    tol = FP32_LARGE_TOL if precision == "fp32" else BF16X3_REL_TOL
    assert max(err.values()) < tol


def test_fp32_exchange_error_is_flat_over_split_counts(torch):
    """The FP32 accumulators sum a fixed number of products, so the split count the autotuner
    picks moves the speed, not the error."""
    from hafnium_amd import _lib
    from hafnium_amd.coulomb_exchange import MI355XCoulombExchange

    n, q, o = 1500, 1200, 250
    jk = MI355XCoulombExchange.synthetic(7, q, n)
    c = torch.from_numpy(synthetic.occupied_orbitals(7, n, o)).cuda()
    k64 = jk.exchange_occupied(c, precision="fp64").cpu().numpy()
    errs = {}
    for splits in (8, 16, 32, 56):
        jk.set_tuning(_lib.TUNE_SYRK_SPLITS, splits)
        errs[splits] = rel(jk.exchange_occupied(c, precision="fp32").cpu().numpy(), k64)
    print(f"fp32 K error by split count: {errs}", flush=True)
    assert max(errs.values()) < FP32_LARGE_TOL
    assert max(errs.values()) < 2.0 * min(errs.values())


@pytest.fixture(scope="module")
def twelve_waters():
    from hafnium_amd.clusters import build_system
    from hafnium_amd.molecule import water_cluster

    return build_system(water_cluster(12, seed=112))


def test_twelve_water_fp64_scf_matches_the_oracle(twelve_waters):
    from hafnium_amd import integrals
    from hafnium_amd.clusters import run_cluster
    from hafnium_amd.molecule import BasisOnAtoms

    result, _ = run_cluster(12, 112, system=twelve_waters, energy_tolerance=1e-10)
    mol = twelve_waters.molecule
    orb = BasisOnAtoms("cc-pvdz", mol)
    aux = BasisOnAtoms("def2-universal-jkfit", mol)
    s, t, v = integrals.one_electron(orb, mol)
    b = scf_ref.df_tensor(integrals.metric(aux), integrals.three_center(aux, orb), "cholesky")
    print(f"12 waters: nao={b.shape[1]} naux={b.shape[0]}", flush=True)
    j_of, k_of = scf_ref.fitted_operators(b)
    ref = scf_ref.rhf(s, t + v, j_of, k_of, mol.n_electrons // 2, mol.nuclear_repulsion())
    assert ref.converged
    print(f"  gpu {result.energy:.12f} oracle {ref.energy:.12f}", flush=True)
    assert result.energy == pytest.approx(ref.energy, abs=1e-9)


def test_twelve_water_batched_matches_serial(twelve_waters):
    from hafnium_amd.batch_scf import BatchedRhf, atomic_start, molecule_inputs
    from hafnium_amd.clusters import atomic_density_guess, build_system, run_cluster
    from hafnium_amd.molecule import water_cluster
    from hafnium_amd.scf import ScfOptions

    specs = [(12, 112), (5, 113), (9, 114)]
    cache: dict = {}
    serial, inputs = [], []
    for n, seed in specs:
        mol = water_cluster(n, seed)
        system = twelve_waters if (n, seed) == (12, 112) else build_system(mol)
        result, _ = run_cluster(n, seed, system=system, guess=atomic_density_guess(mol, cache=cache))
        serial.append(result)
        inputs.append(molecule_inputs(mol))
    batch = BatchedRhf(inputs)
    got = batch.run(atomic_start(batch, cache), ScfOptions(energy_tolerance=1e-8, gradient_tolerance=1e-6))
    for (n, _), g, ref in zip(specs, got, serial):
        print(f"  {n} waters: batched {g.energy:.10f} serial {ref.energy:.10f}", flush=True)
        assert g.converged
        assert g.energy == pytest.approx(ref.energy, abs=1e-8)
