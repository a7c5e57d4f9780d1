"""Batched persistent J/K against the oracle, per instance, and under hipGraph replay."""

from __future__ import annotations

import numpy as np
import pytest

from hafnium_amd import synthetic
from oracle import jk_ref
from tests.tolerances import BF16X3_REL_TOL, FP32_REL_TOL, FP64_REL_TOL

pytestmark = pytest.mark.gpu

TOL = {"fp64": FP64_REL_TOL, "fp32": FP32_REL_TOL, "bf16x3": BF16X3_REL_TOL}

# Uneven instances: tiny, ragged, several row blocks, wide and narrow occupied sets.
SHAPES = [(1, 3, 1), (33, 40, 5), (40, 140, 8), (70, 245, 21), (97, 60, 30), (130, 455, 26), (64, 224, 64),
          (150, 30, 90), (45, 158, 12), (200, 120, 40)]


def build(torch, seeds):
    from hafnium_amd.batch import BatchedCoulombExchange

    bt = BatchedCoulombExchange(SHAPES)
    bt.generate(seeds)
    cs = [synthetic.occupied_orbitals(s, n, o) for s, (n, q, o) in zip(seeds, SHAPES)]
    ds = [c @ c.T for c in cs]
    return bt, cs, ds


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available()
    return torch


@pytest.mark.parametrize("precision", ["fp64", "bf16x3", "fp32"])
@pytest.mark.parametrize("tile", [0, 32, 64, 128, 256])
def test_every_instance_matches_the_oracle(torch, precision, tile):
    seeds = list(range(1000, 1000 + len(SHAPES)))
    bt, cs, ds = build(torch, seeds)
    j, k = bt.run(bt.pack_orbitals(cs), bt.pack_matrices(ds), precision=precision, tile=tile)
    for i, ((n, q, o), seed) in enumerate(zip(SHAPES, seeds)):
        b = synthetic.tensor_full(seed, q, n)
        assert jk_ref.max_relative_error(bt.matrix(j, i).cpu().numpy(), jk_ref.coulomb(b, ds[i])) < TOL[precision]
        assert jk_ref.max_relative_error(bt.matrix(k, i).cpu().numpy(), jk_ref.exchange(b, ds[i])) < TOL[precision]


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
def test_charges_from_x_match_the_density_build(torch, precision):
    """Without densities (D = C C^T) the charges come from the half-transformed X: J matches the
    oracle and the density-form batched J, K is the same build."""
    seeds = list(range(1500, 1500 + len(SHAPES)))
    bt, cs, ds = build(torch, seeds)
    c = bt.pack_orbitals(cs)
    j_d, k_d = bt.run(c, bt.pack_matrices(ds), precision=precision)
    j_d, k_d = j_d.clone(), k_d.clone()
    j_x, k_x = bt.run(c, precision=precision)
    assert torch.equal(k_x, k_d)
    for i, ((n, q, o), seed) in enumerate(zip(SHAPES, seeds)):
        b = synthetic.tensor_full(seed, q, n)
        got = bt.matrix(j_x, i).cpu().numpy()
        assert jk_ref.max_relative_error(got, jk_ref.coulomb(b, ds[i])) < TOL[precision]
        if precision == "fp64":
            assert jk_ref.max_relative_error(got, bt.matrix(j_d, i).cpu().numpy()) < 1e-12
    graph, (j, k) = bt.capture(c, precision=precision)
    for _ in range(2):
        j.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(j, j_x) and torch.equal(k, k_x)
        bt.run(c, precision=precision)
        torch.cuda.synchronize()


def test_graph_replays_match_eager_runs(torch):
    seeds = list(range(2000, 2000 + len(SHAPES)))
    bt, cs, ds = build(torch, seeds)
    c, d = bt.pack_orbitals(cs), bt.pack_matrices(ds)
    j0, k0 = bt.run(c, d, precision="fp64")
    j0, k0 = j0.clone(), k0.clone()
    graph, (j, k) = bt.capture(c, d, precision="fp64")
    for _ in range(3):
        j.zero_()
        k.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(j, j0) and torch.equal(k, k0)
        # Eager work between replays must not disturb the next replay.
        bt.run(c, d, precision="fp64")
        torch.cuda.synchronize()
