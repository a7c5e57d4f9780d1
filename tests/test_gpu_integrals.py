"""The GPU two- and three-centre integrals against the reference's stored PySCF values and the
(separately pinned) host engine."""

from __future__ import annotations

import time

import numpy as np
import pytest

from hafnium_amd import integrals, molecule
from tests.test_integrals import INTEGRAL_TOL

pytestmark = pytest.mark.gpu


def test_stored_metric_and_three_centre_of_water(water):
    mol = integrals.water_from_fixture_positions(water["positions_bohr"])
    orb = molecule.BasisOnAtoms("sto-3g", mol)
    aux = molecule.BasisOnAtoms("def2-universal-jkfit", mol)
    m = integrals.metric_device(aux).cpu().numpy()
    tc = integrals.three_center_device(aux, orb).cpu().numpy()
    np.testing.assert_allclose(m, water["metric"], atol=INTEGRAL_TOL * 1000, rtol=1e-12)
    np.testing.assert_allclose(tc, water["three_center"], atol=INTEGRAL_TOL * 10)


@pytest.mark.parametrize("n_waters,seed", [(1, 5), (3, 7)])
def test_device_integrals_equal_the_host_engine(n_waters, seed):
    """cc-pVDZ (general contractions, d) with def2-universal-jkfit (up to g), several centres."""
    mol = molecule.water_cluster(n_waters, seed)
    orb = molecule.BasisOnAtoms("cc-pvdz", mol)
    aux = molecule.BasisOnAtoms("def2-universal-jkfit", mol)
    m_host = integrals.metric(aux)
    tc_host = integrals.three_center(aux, orb)
    m = integrals.metric_device(aux).cpu().numpy()
    tc = integrals.three_center_device(aux, orb).cpu().numpy()
    np.testing.assert_allclose(m, m_host, atol=1e-12 * np.abs(m_host).max())
    np.testing.assert_allclose(tc, tc_host, atol=1e-12 * np.abs(tc_host).max())


def test_twelve_waters_on_the_gpu():
    """The largest cluster of the SCF workload: device against host, and the device time."""
    import torch

    mol = molecule.water_cluster(12, 112)
    orb = molecule.BasisOnAtoms("cc-pvdz", mol)
    aux = molecule.BasisOnAtoms("def2-universal-jkfit", mol)
    integrals.metric_device(aux)  # first launch: code object and tables
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m = integrals.metric_device(aux)
    tc = integrals.three_center_device(aux, orb)
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    t0 = time.perf_counter()
    m_host = integrals.metric(aux)
    tc_host = integrals.three_center(aux, orb)
    host_seconds = time.perf_counter() - t0
    print(f"12 waters: device integrals {seconds:.4f} s, host {host_seconds:.4f} s", flush=True)
    np.testing.assert_allclose(m.cpu().numpy(), m_host, atol=1e-12 * np.abs(m_host).max())
    np.testing.assert_allclose(tc.cpu().numpy(), tc_host, atol=1e-12 * np.abs(tc_host).max())
