"""CPU oracle of the DF-J/K Fock build: TEST INFRASTRUCTURE ONLY.

This package restates the reference algorithms in NumPy so that tests, the
smoke check and the CPU-baseline leg of the benchmark can check and time the
HIP path against them. Nothing in ``hafnium_amd`` imports it, and no product
path may call it: the GPU path fails loudly when its HIP library is missing
instead of falling back here.

Modules
-------
jk_ref
    Density-form and occupied-form J/K, energies, error metrics.
layout_ref
    The tiled packed storage of a symmetric ``[aux, ao, ao]`` tensor.
scf_ref
    Restricted Hartree-Fock with a fitted or exact two-electron operator,
    DIIS, canonical orthogonalisation and the fitted tensor.

Pinning (every function here is checked by a test that a plausible mistake
fails; see tests/test_oracle.py):
* J/K density form: the looped four-index definitions on a tiny fitted ERI,
  the exact J/K from the stored four-centre integrals of water/STO-3G to the
  fitting tolerance, and the stored PySCF DF-RHF energy of water/STO-3G.
* Occupied form: equality with the density form for D = C C^T.
* Fitted tensor: sum_P B B against (uv|P) M^-1 (Q|wx) from an explicit
  solve, eigh and Cholesky variants agreeing, and the stored DF-RHF energy.
* SCF, DIIS, orthogonaliser, aufbau: the stored exact RHF energy and orbital
  energies of water/STO-3G from the stored four-centre integrals, the DIIS
  linear-model solution, idempotency and electron count. The averaged
  occupation (spherical atoms): the oxygen atom's (1, 1, 2/3, 2/3, 2/3)
  occupations, electron count and isotropic p blocks, which whole-level
  filling breaks.
* Tiled packed layout: round trip, and dot products against the full form.
"""
