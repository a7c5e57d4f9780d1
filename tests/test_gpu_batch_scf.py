"""Batched RHF on the GPU against one-at-a-time SCF: energies, iterations, precision, starts."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SAME_STATE_TOL = 1e-8  # Eh: batched and serial runs converge to the same FP64 state


@pytest.fixture(scope="module")
def setup():
    from hafnium_amd.batch_scf import BatchedRhf, molecule_inputs
    from hafnium_amd.clusters import atomic_density_guess, build_system, run_cluster
    from hafnium_amd.molecule import water_cluster

    specs = [(2, 41), (3, 42), (1, 43), (2, 44)]
    cache: dict = {}
    serial = []
    inputs = []
    for n, seed in specs:
        mol = water_cluster(n, seed)
        system = build_system(mol)
        guess = atomic_density_guess(mol, cache=cache)
        result, _ = run_cluster(n, seed, system=system, guess=guess)
        serial.append(result)
        inputs.append(molecule_inputs(mol))
    return BatchedRhf(inputs), serial, cache, inputs


@pytest.mark.parametrize("eigensolver", ["lowest", "jacobi", "divide"])
def test_batched_fp64_matches_serial(setup, eigensolver):
    from hafnium_amd.batch_scf import BatchedRhf, atomic_start
    from hafnium_amd.scf import ScfOptions

    batch, serial, cache, inputs = setup
    batch = BatchedRhf(inputs, eigensolver=eigensolver)
    opts = ScfOptions(energy_tolerance=1e-8, gradient_tolerance=1e-6)
    results = batch.run(atomic_start(batch, cache), opts)
    for got, ref in zip(results, serial):
        assert got.converged
        assert got.energy == pytest.approx(ref.energy, abs=SAME_STATE_TOL)
        assert abs(got.iterations - ref.n_iterations) <= 2


def test_molecular_start_reaches_the_same_state_in_fewer_iterations(setup):
    """The superposition of water densities converges to the atomic start's FP64 state, and
    in no more iterations on any molecule."""
    from hafnium_amd.batch_scf import atomic_start, molecular_start
    from hafnium_amd.learned_guess import water_orbitals
    from hafnium_amd.scf import ScfOptions

    batch, serial, cache, _ = setup
    opts = ScfOptions(energy_tolerance=1e-8, gradient_tolerance=1e-6)
    atomic = batch.run(atomic_start(batch, cache), opts)
    molecular = batch.run(molecular_start(batch, water_orbitals(cache=cache)), opts)
    for got, base, ref in zip(molecular, atomic, serial):
        assert got.converged
        assert got.energy == pytest.approx(ref.energy, abs=SAME_STATE_TOL)
        assert got.iterations <= base.iterations
    assert sum(r.iterations for r in molecular) < sum(r.iterations for r in atomic)


def test_batched_mixed_precision_reaches_the_fp64_state(setup):
    from hafnium_amd.batch_scf import atomic_start
    from hafnium_amd.scf import ScfOptions

    batch, serial, cache, _ = setup
    opts = ScfOptions(energy_tolerance=1e-8, gradient_tolerance=1e-6, precision="bf16x3")
    results = batch.run(atomic_start(batch, cache, "bf16x3"), opts)
    for got, ref in zip(results, serial):
        assert got.converged
        assert got.energy == pytest.approx(ref.energy, abs=SAME_STATE_TOL)
        assert got.history[0][2] == "bf16x3" and got.history[got.iterations - 1][2] == "fp64"


def test_coulomb_from_charges_matches_the_density_build(setup):
    import torch

    from hafnium_amd import learned_guess as lg
    from hafnium_amd import _lib

    batch, _, cache, inputs = setup
    rng = np.random.default_rng(3)
    dens, charges = [], []
    for x, f in zip(inputs, batch.factors):
        c = rng.standard_normal((x.nao, x.n_alpha)) * 0.1
        d = c @ c.T
        dens.append((d, c))
        metric, three_center = x.metric.cpu().numpy(), x.three_center.cpu().numpy()
        coef = lg.fitted_coefficients(metric, three_center, d)  # fits 2 D
        v = torch.from_numpy(metric @ coef).to(batch.device).view(-1, 1).contiguous()
        _lib.check(batch.lib.hfjk_dtrsm_lower(batch._ctx.handle, x.naux, 1, ctypes_ptr(f), ctypes_ptr(v),
                                              batch._stream()))
        charges.append(v.view(-1))
    j_charges = batch.jk.coulomb_from_charges(torch.cat(charges))
    d_flat = batch.jk.pack_matrices([d for d, _ in dens])
    c_flat = batch.jk.pack_orbitals([c for _, c in dens])
    j_run, _ = batch.jk.run(c_flat, d_flat)
    torch.cuda.synchronize()
    a, b = j_charges.cpu().numpy(), 2.0 * j_run.cpu().numpy()
    assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()


def ctypes_ptr(t):
    import ctypes

    return ctypes.c_void_p(t.data_ptr())


def test_learned_start_falls_back_per_molecule(setup):
    from hafnium_amd import learned_guess as lg
    from hafnium_amd.batch_scf import atomic_start, learned_start
    from hafnium_amd.mlp import Mlp
    from hafnium_amd.scf import ScfOptions

    batch, serial, cache, _ = setup
    aux = "def2-universal-jkfit"
    # An untrained network with its outputs scaled up: its densities are worse than the atoms'.
    elements = {}
    for z in lg.ELEMENTS:
        n_coef = lg.shell_blocks(aux, z)[1]
        probe = Mlp([lg.N_FEATURES, *lg.HIDDEN, n_coef], seed=z)
        elements[z] = lg.ElementModel(z, np.zeros(lg.N_FEATURES), np.ones(lg.N_FEATURES), np.zeros(n_coef),
                                      np.linalg.cholesky(lg.one_centre_metric(aux, z)).T, 1e3,
                                      probe.params.cpu().numpy())
    model = lg.LearnedGuessModel(elements, aux)
    sad = atomic_start(batch, cache)
    start = learned_start(batch, model, sad)
    assert start.info["fallback"] == batch.count and start.info["learned"] == 0
    results = batch.run(start, ScfOptions(energy_tolerance=1e-8, gradient_tolerance=1e-6))
    for got, ref in zip(results, serial):
        assert got.energy == pytest.approx(ref.energy, abs=SAME_STATE_TOL)


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_pipelined_batches_match_serial(setup, precision):
    """Two batches on their own streams, iterations interleaved (one batch's eigensolver runs while
    the other's J/K build does), reach the serial FP64 states; each group as a single batch gives
    the same energies and iteration counts."""
    from hafnium_amd.batch_scf import BatchedRhf, atomic_start, balanced_groups, run_pipelined
    from hafnium_amd.scf import ScfOptions

    _, serial, cache, inputs = setup
    groups = balanced_groups([x.nao for x in inputs], 2)
    assert sorted(i for g in groups for i in g) == list(range(len(inputs)))
    batches = [BatchedRhf([inputs[i] for i in g]) for g in groups]
    opts = ScfOptions(precision=precision, energy_tolerance=1e-8, gradient_tolerance=1e-6)
    outs = run_pipelined([(b, atomic_start(b, cache, precision)) for b in batches], opts)
    alone = [b.run(atomic_start(b, cache, precision), opts) for b in batches]
    for g, out, ref_out in zip(groups, outs, alone):
        for i, got, ref in zip(g, out, ref_out):
            assert got.converged
            assert got.energy == pytest.approx(serial[i].energy, abs=SAME_STATE_TOL)
            assert got.energy == pytest.approx(ref.energy, abs=1e-10)
            assert got.iterations == ref.iterations

