"""MLP kernels on the GPU against the CPU oracle: GEMM forms, forward, gradients, Adam, training."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_ref

pytestmark = pytest.mark.gpu

# GEMM: identical bf16 operands, FP32 (kernel) vs FP64 (oracle) sums of <= 300 terms.
GEMM_REL_TOL = 1e-5
# Whole network: intermediates rounded to bf16 can land on the other side of a
# rounding boundary when the FP32 and FP64 sums differ in the last bit, so a few
# operands differ by one bf16 ulp (2^-8 relative).
NET_REL_TOL = 2e-3


def _torch():
    import torch

    return torch


def _gemm(lib_mlp, a, b, ta, tb, rows=None, bias=None, silu=False, a16=False, b16=False):
    torch = _torch()
    dev = lib_mlp.device
    m = a.shape[1] if ta else (len(rows) if rows is not None else a.shape[0])
    k = a.shape[0] if ta else a.shape[1]
    if ta and rows is not None:
        k = len(rows)
    n = b.shape[0] if tb else b.shape[1]
    ad = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)
    bd = torch.from_numpy(np.ascontiguousarray(b, dtype=np.float32)).to(dev)
    if a16:
        ad = ad.to(torch.bfloat16)
    if b16:
        bd = bd.to(torch.bfloat16)
    c = torch.empty((m, n), dtype=torch.float32, device=dev)
    h = torch.empty_like(c) if silu else None
    rows_d = None if rows is None else (torch.tensor(rows, dtype=torch.int32, device=dev), 0)
    bias_d = None if bias is None else (torch.from_numpy(bias.astype(np.float32)).to(dev), 0)
    lib_mlp._gemm(m, n, k, ad, 0, a16, a.shape[1], ta, rows_d, bd, 0, b16, b.shape[1], tb, bias_d, c, h)
    torch.cuda.synchronize()
    return c.cpu().numpy().astype(np.float64), None if h is None else h.cpu().numpy().astype(np.float64)


@pytest.fixture(scope="module")
def net():
    from hafnium_amd.mlp import Mlp

    return Mlp([72, 96, 80, 13], seed=1)


@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_gemm_transpositions(net, ta, tb):
    rng = np.random.default_rng(10 + 2 * ta + tb)
    m, n, k = 131, 77, 300  # no dimension a multiple of the 64 x 64 x 32 tiles
    a = rng.standard_normal((k, m) if ta else (m, k))
    b = rng.standard_normal((n, k) if tb else (k, n))
    c, _ = _gemm(net, a, b, ta, tb)
    op_a = mlp_ref.bf16_round(a.T if ta else a)
    op_b = mlp_ref.bf16_round(b.T if tb else b)
    ref = op_a @ op_b
    assert np.abs(c - ref).max() <= GEMM_REL_TOL * np.abs(ref).max()


def test_gemm_gather_bias_silu_and_bf16_operands(net):
    rng = np.random.default_rng(20)
    x = rng.standard_normal((50, 40))
    w = rng.standard_normal((40, 70))
    bias = rng.standard_normal(70)
    rows = [49, 0, 7, 7, 31, 2, 18, 44, 5]
    c, h = _gemm(net, x, w, False, False, rows=rows, bias=bias, silu=True, b16=True)
    ref = mlp_ref.bf16_round(x[rows]) @ mlp_ref.bf16_round(w) + bias.astype(np.float32)  # bias stays FP32
    assert np.abs(c - ref).max() <= GEMM_REL_TOL * np.abs(ref).max()
    np.testing.assert_allclose(h, mlp_ref.silu(c), rtol=1e-5, atol=1e-6)
    # Transposed gathered operand: op(A) = x[rows]^T, the weight-gradient form.
    g = rng.standard_normal((len(rows), 70))
    c2, _ = _gemm(net, x, g, True, False, rows=rows)
    ref2 = mlp_ref.bf16_round(x[rows]).T @ mlp_ref.bf16_round(g)
    assert np.abs(c2 - ref2).max() <= GEMM_REL_TOL * np.abs(ref2).max()
    # bf16 storage of A gives the same result as rounding it on load.
    c3, _ = _gemm(net, x, w, False, False, a16=True)
    c4, _ = _gemm(net, x, w, False, False)
    assert np.abs(c3 - c4).max() <= GEMM_REL_TOL * np.abs(c4).max()


def test_forward_and_gradients_match_the_oracle(net):
    torch = _torch()
    rng = np.random.default_rng(30)
    x = rng.standard_normal((200, 72)).astype(np.float32)
    t = rng.standard_normal((200, 13)).astype(np.float32)
    wts = rng.uniform(0.5, 2.0, 13).astype(np.float32)
    layers = net.layers()
    xd = torch.from_numpy(x).to(net.device)
    td = torch.from_numpy(t).to(net.device)
    wd = torch.from_numpy(wts).to(net.device)
    rows = np.arange(200)[::-1].copy()
    rd = torch.from_numpy(rows.astype(np.int32)).to(net.device)
    y = net.forward(xd, (rd, 0, 200), keep=True)
    grad = torch.empty_like(y)
    loss = net.loss(y, td, (rd, 0, 200), wd, 1.0 / 200, grad)
    net.backward(grad)
    torch.cuda.synchronize()
    y_ref, _ = mlp_ref.forward(layers, x[rows], exact=False)
    assert np.abs(y.cpu().numpy() - y_ref).max() <= NET_REL_TOL * np.abs(y_ref).max()
    loss_ref, grads_ref = mlp_ref.loss_and_gradients(layers, x[rows], t[rows], wts, 1.0 / 200, exact=False)
    assert float(loss.item()) == pytest.approx(loss_ref, rel=NET_REL_TOL)
    for (dw, db), (dw_ref, db_ref) in zip(net.gradients(), grads_ref):
        assert np.abs(dw - dw_ref).max() <= NET_REL_TOL * np.abs(dw_ref).max()
        assert np.abs(db - db_ref).max() <= NET_REL_TOL * np.abs(db_ref).max()
    # And the exact (float64) network, to the size of bf16 rounding.
    _, grads_exact = mlp_ref.loss_and_gradients(layers, x[rows], t[rows], wts, 1.0 / 200, exact=True)
    for (dw, _), (dw_ex, _) in zip(net.gradients(), grads_exact):
        assert np.abs(dw - dw_ex).max() <= 5e-2 * np.abs(dw_ex).max()


def test_adam_matches_the_oracle():
    torch = _torch()
    from hafnium_amd.mlp import Mlp

    net = Mlp([8, 16, 3], seed=2)
    p0 = net.params.cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(40)
    m1 = np.zeros_like(p0)
    m2 = np.zeros_like(p0)
    p = p0.copy()
    for step in range(1, 4):
        g = rng.standard_normal(p.size)
        net.grads.copy_(torch.from_numpy(g.astype(np.float32)).to(net.device))
        net.adam(1e-2, weight_decay=1e-3)
        p, m1, m2 = mlp_ref.adam_step(p, g.astype(np.float32).astype(np.float64), m1, m2, step, 1e-2,
                                      weight_decay=1e-3)
    torch.cuda.synchronize()
    np.testing.assert_allclose(net.params.cpu().numpy(), p, rtol=1e-5, atol=1e-7)
    # The bf16 copy follows the master weights.
    np.testing.assert_array_equal(net.params16.float().cpu().numpy(),
                                  mlp_ref.bf16_round(net.params.cpu().numpy()).astype(np.float32))


def test_training_fits_a_teacher_network():
    torch = _torch()
    from hafnium_amd.mlp import Mlp

    rng = np.random.default_rng(50)
    teacher = [(rng.standard_normal((16, 32)) / 4, rng.standard_normal(32) * 0.1),
               (rng.standard_normal((32, 4)) / 6, rng.standard_normal(4) * 0.1)]
    x = rng.standard_normal((2048, 16)).astype(np.float32)
    t, _ = mlp_ref.forward(teacher, x, exact=True)
    net = Mlp([16, 64, 64, 4], seed=3)
    xd = torch.from_numpy(x).to(net.device)
    td = torch.from_numpy(t.astype(np.float32)).to(net.device)
    all_rows = torch.arange(2048, dtype=torch.int32, device=net.device)
    first = float(net.loss(net.forward(xd), td).item()) / 2048
    for epoch in range(30):
        perm = torch.from_numpy(rng.permutation(2048).astype(np.int32)).to(net.device)
        for lo in range(0, 2048, 128):
            net.train_step(xd, td, (perm, lo, lo + 128), lr=3e-3)
    last = float(net.loss(net.forward(xd, (all_rows, 0, 2048)), td).item()) / 2048
    assert last < 0.05 * first
