"""Learned guess on the GPU: data from converged FP64 SCF, training, the guess and its fallback."""

from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SAME_STATE_TOL = 1e-8  # Eh: every start must reach the FP64 energy of the atomic-guess SCF


@pytest.fixture(scope="module")
def tiny_model(tmp_path_factory):
    from hafnium_amd import learned_guess as lg

    path = str(tmp_path_factory.mktemp("lg") / "data.npz")
    data = lg.generate_dataset([501, 502, 503, 504, 505, 506], [2, 3, 2, 3, 2, 3], path, log=lambda *a, **k: None)
    model, report = lg.train(data, epochs=40, batch=32, lr=2e-3, validation_fraction=0.34, log=lambda *a, **k: None)
    return model, report


def test_dataset_holds_per_atom_records(tmp_path):
    from hafnium_amd import learned_guess as lg

    path = str(tmp_path / "d.npz")
    data = lg.generate_dataset([601], [2], path, log=lambda *a, **k: None)
    assert data["x_8"].shape == (2, lg.N_FEATURES) and data["x_1"].shape == (4, lg.N_FEATURES)
    assert data["y_8"].shape == (2, 77) and data["y_1"].shape == (4, 18)
    # The fitted density integrates to the electron count: sum_P d_P int chi_P = 20 for (H2O)_2.
    with np.load(path, allow_pickle=False) as f:
        assert f["meta"].shape == (1, 4)


def test_training_beats_the_mean_density(tiny_model):
    _, report = tiny_model
    for z, r in report.items():
        assert np.isfinite(r["best_val"])
        assert r["best_val"] < r["mean_density_val"], (z, r)


def test_learned_guess_starts_a_converging_scf(tiny_model, tmp_path):
    from hafnium_amd import learned_guess as lg
    from hafnium_amd.clusters import atomic_density_guess, build_system, run_cluster
    from hafnium_amd.molecule import water_cluster

    model, _ = tiny_model
    path = str(tmp_path / "m.npz")
    model.save(path)
    loaded = lg.LearnedGuessModel.load(path)
    system = build_system(water_cluster(3, seed=777), keep_integrals=True)
    np.testing.assert_allclose(loaded.predict(system.molecule), model.predict(system.molecule), rtol=1e-6,
                               atol=1e-6)
    sad = atomic_density_guess(system.molecule)
    start, info = lg.learned_guess(system, loaded, sad)
    s = system.hamiltonian.overlap
    d = start.density
    if info["used"] == "learned":
        # The learned start is an aufbau density: tr(D S) = n_alpha and D S D = D.
        assert float((d @ s).trace()) == pytest.approx(system.hamiltonian.n_alpha, abs=1e-8)
        assert float((d @ s @ d - d).abs().max()) < 1e-8
    else:
        np.testing.assert_allclose(d.cpu().numpy(), sad, atol=1e-14)
    # The handed-on first build is the Fock matrix of that density.
    fock, _ = system.hamiltonian.fock(start.density, start.orbitals)
    assert float((fock - start.first_build[0]).abs().max()) < 1e-12
    ref, _ = run_cluster(3, 777, system=system, guess=sad)
    got, _ = run_cluster(3, 777, system=system, guess=start)
    assert got.energy == pytest.approx(ref.energy, abs=SAME_STATE_TOL)


def test_fallback_to_the_atomic_guess(tiny_model):
    from hafnium_amd import learned_guess as lg
    from hafnium_amd.clusters import atomic_density_guess, build_system
    from hafnium_amd.molecule import water_cluster

    model, _ = tiny_model
    # Force a useless prediction from the host: scale the whitened outputs up a thousandfold.
    broken = lg.LearnedGuessModel({z: lg.ElementModel(em.z, em.feature_mean, em.feature_scale, em.target_mean,
                                                      em.whitening, em.target_scale * 1e3, em.params)
                                   for z, em in model.elements.items()}, model.aux_name)
    system = build_system(water_cluster(2, seed=778), keep_integrals=True)
    sad = atomic_density_guess(system.molecule)
    start, info = lg.learned_guess(system, broken, sad)
    assert info["used"] == "sad"
    assert info["energy_learned"] > info["energy_sad"]
    # The fallback is the atomic guess itself, with its own Fock build as the first iteration.
    np.testing.assert_allclose(start.density.cpu().numpy(), sad, atol=1e-14)
    assert start.orbitals is None
    fock, _ = system.hamiltonian.fock(start.density)
    assert float((fock - start.first_build[0]).abs().max()) < 1e-12


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "hafnium_amd", "data",
                                                    "learned_guess_water.npz")),
                    reason="no trained model stored in hafnium_amd/data")
def test_stored_model_improves_the_start_on_held_out_clusters():
    from hafnium_amd import learned_guess as lg
    from hafnium_amd.clusters import atomic_density_guess, build_system, run_cluster
    from hafnium_amd.molecule import water_cluster

    model = lg.LearnedGuessModel.load()
    cache: dict = {}
    sad_its, fallback_its, learned_its, used, gaps = [], [], [], [], []
    for n, seed in [(4, 900001), (6, 900002), (8, 900003), (10, 900004), (12, 900005), (14, 900006)]:
        system = build_system(water_cluster(n, seed), keep_integrals=True)
        sad = atomic_density_guess(system.molecule, cache=cache)
        ref, _ = run_cluster(n, seed, system=system, guess=sad)
        fb, _ = run_cluster(n, seed, system=system, guess=lg.atomic_iterate_start(system, sad))
        start, info = lg.learned_guess(system, model, sad)
        got, _ = run_cluster(n, seed, system=system, guess=start)
        for r in (fb, got):
            assert r.energy == pytest.approx(ref.energy, abs=SAME_STATE_TOL)
        sad_its.append(ref.n_iterations)
        fallback_its.append(fb.n_iterations)
        learned_its.append(got.n_iterations)
        used.append(info["used"])
        gaps.append((info["energy_learned"] - ref.energy, info["energy_sad"] - ref.energy))
    # What the network adds over the method without it: a start of lower energy than the atomic
    # guess on most clusters, and fewer iterations than the SCF from the atomic density; against the
    # atomic guess's first iterate alone the iteration counts are on par.
    assert used.count("learned") >= len(used) - 1, (used, gaps)
    assert np.mean(learned_its) < np.mean(sad_its), (learned_its, sad_its)
    assert np.mean(learned_its) <= np.mean(fallback_its) + 0.5, (learned_its, fallback_its)
