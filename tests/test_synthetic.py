"""The synthetic generator: golden sample, distribution, symmetry and envelope."""

from __future__ import annotations

import numpy as np

from hafnium_amd import synthetic


def test_golden_sample_is_reproduced_bit_for_bit(synthetic_c0):
    g = synthetic_c0
    b = synthetic.tensor_full(int(g["seed"]), int(g["naux"]), int(g["nao"]))
    np.testing.assert_array_equal(b[g["sample_q"], g["sample_m"], g["sample_n"]], g["sample_b"])


def test_philox_known_answer():
    # Random123 known-answer vectors for Philox4x32-10 (kat_vectors):
    # counter 0, key 0 and counter/key all ones.
    out = synthetic.philox4x32(0, 0, 0, 0, 0)
    assert [int(v) for v in out] == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    ones = 0xFFFFFFFF
    out = synthetic.philox4x32(ones, ones, ones, ones, (ones << 32) | ones)
    assert [int(v) for v in out] == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]


def test_tensor_is_symmetric_with_the_decay_envelope():
    b = synthetic.tensor_full(5, 4, 50)
    np.testing.assert_array_equal(b, np.transpose(b, (0, 2, 1)))
    env = synthetic.decay(50)
    np.testing.assert_allclose(env, np.exp(-np.arange(50) / 20.0), rtol=1e-14)
    z = b[:, np.arange(50), np.arange(50)]
    assert np.all(np.isfinite(z))


def test_deviates_are_standard_normal():
    z = synthetic.normal_from_counters(np.arange(200000), 1, 2, synthetic.TAG_TENSOR, 11)
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1.0) < 0.01
    # Tail mass beyond 3 sigma of a normal is 0.27 %.
    assert 0.0018 < np.mean(np.abs(z) > 3.0) < 0.0036


def test_own_transcendentals_track_libm():
    u = np.linspace(2.0**-60, 1.0, 10001)
    np.testing.assert_allclose(synthetic._log(u), np.log(u), rtol=1e-15, atol=1e-15)
    u = np.linspace(0.0, 1.0, 10001, endpoint=False)
    np.testing.assert_allclose(synthetic._cos_two_pi(u), np.cos(2 * np.pi * u), atol=1e-15)


def test_occupied_orbitals_are_orthonormal():
    c = synthetic.occupied_orbitals(1, 60, 12)
    np.testing.assert_allclose(c.T @ c, np.eye(12), atol=1e-13)
    d = c @ c.T
    np.testing.assert_allclose(d @ d, d, atol=1e-13)
    assert abs(np.trace(d) - 12) < 1e-12
