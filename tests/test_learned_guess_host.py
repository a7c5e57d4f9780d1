"""Host side of the learned guess: harmonics, frames, descriptors, coefficient rotation, fitted J."""

from __future__ import annotations

import numpy as np
import pytest

from hafnium_amd import learned_guess as lg
from hafnium_amd.molecule import Molecule, water_cluster

AUX = "def2-universal-jkfit"


def _rotation(seed):
    q = np.random.default_rng(seed).standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _moved(mol, seed):
    r = _rotation(seed)
    return Molecule(mol.charges, mol.positions @ r.T + np.array([0.7, -1.3, 2.1]))


def test_solid_harmonics_are_orthonormal_on_the_sphere():
    # Gauss-Legendre x uniform-phi product rule exact to degree 9 in each variable.
    x, wx = np.polynomial.legendre.leggauss(10)
    phi = np.linspace(0.0, 2 * np.pi, 20, endpoint=False)
    ct, ph = np.meshgrid(x, phi, indexing="ij")
    st = np.sqrt(1 - ct ** 2)
    pts = np.stack([st * np.cos(ph), st * np.sin(ph), ct], -1).reshape(-1, 3)
    w = np.repeat(wx, len(phi)) * (2 * np.pi / len(phi)) / (4 * np.pi)
    for l in range(5):
        y = lg.solid_harmonics(l, pts)
        np.testing.assert_allclose((w[:, None] * y).T @ y, np.eye(2 * l + 1), atol=1e-12)
    # Known forms: p = (x, y, z) up to the common factor, d_0 ~ (2z^2 - x^2 - y^2) / 2.
    v = np.array([[0.3, -0.5, 0.8]])
    p = lg.solid_harmonics(1, v)[0]
    np.testing.assert_allclose(p / p[0], v[0] / v[0, 0], rtol=1e-12)
    d = lg.solid_harmonics(2, v)[0]
    ratio = d[2] / (2 * v[0, 2] ** 2 - v[0, 0] ** 2 - v[0, 1] ** 2)
    d_other = lg.solid_harmonics(2, np.array([[1.0, 0.2, -0.4]]))[0]
    assert d_other[2] / (2 * 0.16 - 1.0 - 0.04) == pytest.approx(ratio, rel=1e-12)


def test_rotation_matrices_form_a_representation():
    r1, r2 = _rotation(1), _rotation(2)
    d1, d2, d12 = (lg.rotation_matrices(r, 4) for r in (r1, r2, r1 @ r2))
    for l in range(5):
        np.testing.assert_allclose(d12[l], d1[l] @ d2[l], atol=1e-11)
        np.testing.assert_allclose(d1[l] @ d1[l].T, np.eye(2 * l + 1), atol=1e-11)
    np.testing.assert_allclose(lg.rotation_matrices(np.eye(3), 4)[3], np.eye(7), atol=1e-12)


def test_frames_and_descriptors_are_invariant():
    mol = water_cluster(4, seed=21)
    rots = lg.water_frames(mol)
    for r in rots:
        np.testing.assert_allclose(r.T @ r, np.eye(3), atol=1e-12)
        assert np.linalg.det(r) == pytest.approx(1.0)
    feats = lg.descriptors(mol)
    assert feats.shape == (12, lg.N_FEATURES)
    np.testing.assert_allclose(lg.descriptors(_moved(mol, 5)), feats, atol=1e-12)
    # A lone water sees only its own two partners: O has 2 H neighbours, no O.
    lone = lg.descriptors(water_cluster(1, seed=0))
    per = lone.reshape(3, 2, lg.N_RADIAL, 9)
    assert np.all(per[0, 1] == 0.0) and np.any(per[0, 0] != 0.0)
    with pytest.raises(Exception):
        lg.water_frames(Molecule((1, 8, 1), np.zeros((3, 3))))


def test_local_coefficients_follow_the_integral_engine_under_rotation():
    from hafnium_amd import integrals
    from hafnium_amd.molecule import BasisOnAtoms

    mol = water_cluster(2, seed=3)

    def fitted(m):
        orb, aux = BasisOnAtoms("sto-3g", m), BasisOnAtoms(AUX, m)
        s, t, v = integrals.one_electron(orb, m)
        w, u = np.linalg.eigh(s)
        x = u / np.sqrt(w)
        _, c = np.linalg.eigh(x.T @ (t + v) @ x)
        c = x @ c[:, : m.n_electrons // 2]
        return lg.fitted_coefficients(integrals.metric(aux), integrals.three_center(aux, orb), c @ c.T)

    c1, c2 = fitted(mol), fitted(_moved(mol, 9))
    l1, l2 = lg.to_local(mol, c1, AUX), lg.to_local(_moved(mol, 9), c2, AUX)
    for z in (1, 8):
        scale = np.abs(l1[z]).max()
        assert np.abs(l1[z] - l2[z]).max() < 1e-9 * scale
    np.testing.assert_allclose(lg.to_lab(mol, l1, AUX), c1, atol=1e-13)
    # A wrong convention (forward instead of inverse rotation) would not be invariant.
    assert len(lg.shell_blocks(AUX, 8)[0]) > 0 and l1[8].shape[1] == 77 and l1[1].shape[1] == 18


def test_coulomb_of_fitted_coefficients_is_the_fitted_coulomb():
    from oracle import scf_ref

    rng = np.random.default_rng(0)
    naux, n = 30, 8
    a = rng.standard_normal((naux, naux))
    metric = a @ a.T + naux * np.eye(naux)
    tc = rng.standard_normal((naux, n, n))
    tc = tc + tc.transpose(0, 2, 1)
    d = rng.standard_normal((n, n))
    d = d + d.T
    b = scf_ref.df_tensor(metric, tc, "cholesky")
    j_of, _ = scf_ref.fitted_operators(b)
    got = lg.coulomb_of_coefficients(tc, lg.fitted_coefficients(metric, tc, d))
    np.testing.assert_allclose(got, j_of(2.0 * d), rtol=1e-10, atol=1e-10)


def test_whitening_measures_the_coulomb_self_energy():
    g = lg.one_centre_metric(AUX, 1)
    np.testing.assert_allclose(g, g.T, atol=1e-12)
    u = np.linalg.cholesky(g).T
    rng = np.random.default_rng(1)
    mean = rng.standard_normal(g.shape[0])
    em = lg.ElementModel(1, np.zeros(3), np.ones(3), mean, u, 0.5, np.zeros(0))
    c = rng.standard_normal((4, g.shape[0]))
    y = em.whiten(c)
    np.testing.assert_allclose(em.unwhiten(y), c, atol=1e-10)
    delta = c - mean
    np.testing.assert_allclose(np.sum(y ** 2, axis=1), np.einsum("ai,ij,aj->a", delta, g, delta) / 0.25,
                               rtol=1e-10)


def test_molecular_factors_rotate_orbitals_with_the_integral_engine():
    """Rotating a water's orbitals into its frame preserves their overlaps: for any C,
    C_k^T S_kk C_k equals C^T S_ref C (S_kk = water k's own overlap block). An unrotated
    copy does not, and the reference water is mapped onto itself."""
    from hafnium_amd import integrals
    from hafnium_amd.molecule import BasisOnAtoms

    ref = lg.reference_water()
    s_ref = integrals.one_electron(BasisOnAtoms("cc-pvdz", ref), ref)[0]
    c = np.random.default_rng(1).standard_normal((s_ref.shape[0], 5))
    gram = c.T @ s_ref @ c
    mol = water_cluster(3, seed=11)
    s = integrals.one_electron(BasisOnAtoms("cc-pvdz", mol), mol)[0]
    dens, fac = lg.molecular_factors(mol, c)
    n = s_ref.shape[0]
    assert fac.shape == (3 * n, 15)
    np.testing.assert_allclose(dens, fac @ fac.T)
    for k in range(3):
        blk = fac[k * n:(k + 1) * n, 5 * k:5 * k + 5]
        skk = s[k * n:(k + 1) * n, k * n:(k + 1) * n]
        assert np.abs(blk.T @ skk @ blk - gram).max() < 1e-10 * np.abs(gram).max()
        assert np.abs(fac[k * n:(k + 1) * n, :5 * k]).max(initial=0.0) == 0.0
    unrotated = s[n:2 * n, n:2 * n]
    assert np.abs(c.T @ unrotated @ c - gram).max() > 1e-3 * np.abs(gram).max()
    _, self_fac = lg.molecular_factors(ref, c)
    np.testing.assert_allclose(self_fac, c, atol=1e-12)


def test_datasets_merge_and_recompute_their_descriptors():
    """Chunks of a dataset stored without descriptors merge into one, with cluster indices
    renumbered, and the descriptors recomputed from (seed, size) equal the direct ones."""
    from hafnium_amd import learned_guess as lg
    from hafnium_amd.molecule import water_cluster

    specs = [(2, 11), (3, 12), (1, 13)]

    def part(rows, first):
        out = {"meta": np.array([[seed, n, -1.0, 10] for n, seed in rows], dtype=np.float64)}
        for z in lg.ELEMENTS:
            ys, cl = [], []
            for k, (n, seed) in enumerate(rows):
                count = int(np.sum(np.asarray(water_cluster(n, seed).charges) == z))
                ys.append(np.full((count, 3), first + k, dtype=np.float64))
                cl += [k] * count
            out[f"y_{z}"] = np.concatenate(ys)
            out[f"cluster_{z}"] = np.asarray(cl, dtype=np.int32)
        return out

    merged = lg.with_features(lg.merge_datasets([part(specs[:2], 0), part(specs[2:], 2)]))
    assert merged["meta"].shape[0] == 3
    for z in lg.ELEMENTS:
        np.testing.assert_array_equal(merged[f"y_{z}"][:, 0], merged[f"cluster_{z}"])
        direct = []
        for n, seed in specs:
            mol = water_cluster(n, seed)
            direct.append(lg.descriptors(mol)[np.asarray(mol.charges) == z].astype(np.float32))
        np.testing.assert_array_equal(merged[f"x_{z}"], np.concatenate(direct))


def test_batched_descriptors_and_rotations_match_one_molecule_at_a_time():
    mols = [water_cluster(n, 4000 + n) for n in (1, 2, 5, 3)]
    many = lg.descriptors_many(mols)
    for m, f in zip(mols, many):
        np.testing.assert_array_equal(f, lg.descriptors(m))
    rng = np.random.default_rng(5)
    local = [{z: rng.standard_normal((int(np.sum(np.asarray(m.charges) == z)), lg.shell_blocks(AUX, z)[1]))
              for z in lg.ELEMENTS} for m in mols]
    for m, l, lab in zip(mols, local, lg.to_lab_many(mols, local, AUX)):
        np.testing.assert_allclose(lab, lg.to_lab(m, l, AUX), atol=1e-14)
        # and back: the local coefficients are recovered from the laboratory frame.
        back = lg.to_local(m, lab, AUX)
        for z in lg.ELEMENTS:
            np.testing.assert_allclose(back[z], l[z], atol=1e-12)
