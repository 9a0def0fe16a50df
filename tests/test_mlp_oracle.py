"""The FP64 restatement of the learned warm-start network (oracle/mlp.py), pinned
on the CPU: its BF16 rounding against torch's float32 -> bfloat16 conversion, its
gradient against central finite differences, and closed forms."""
import numpy as np
import pytest
import torch

from oracle import mlp


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.normal(size=4000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, size=4000),
        # exact ties (low 16 bits 0x8000) round to even
        (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32),
        np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, 3.3895314e38, 1e-40, -1e-45], dtype=np.float32),
    ])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    got = mlp.bf16_round(x)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("h", [16, 32])
def test_gradient_matches_finite_differences(h):
    rng = np.random.default_rng(h)
    P = mlp.random_params(h, rng, scale=1.5).astype(np.float64)
    X = rng.normal(size=(32, mlp.NF))
    T = rng.normal(size=(32, mlp.NO))
    _, G = mlp.loss_grad(P, h, X, T, bf16=False)
    idx = []
    for l in range(4):
        W, b = mlp.woff(l, h), mlp.boff(l, h)
        nin = mlp.NF if l == 0 else h
        nout = mlp.NO if l == 3 else h
        for _ in range(12):  # real (non-padding) weights and biases of every layer
            idx.append(W + rng.integers(nin) * mlp.fan_out(l, h) + rng.integers(nout))
        for _ in range(4):
            idx.append(b + rng.integers(nout))
    eps = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += eps
        Pm[i] -= eps
        fd = (mlp.loss_grad(Pp, h, X, T, bf16=False)[0] - mlp.loss_grad(Pm, h, X, T, bf16=False)[0]) / (2 * eps)
        assert abs(fd - G[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, G[i])
    # padding gets no gradient
    W0 = mlp.unpack(G, h)[0]
    assert not W0[0][mlp.NF:].any() and not W0[3][:, mlp.NO:].any()


def test_closed_form_with_zero_weights():
    """All weights 0: y = b4, the loss and the bias-4 gradient in closed form,
    the hidden biases get no gradient (W4 = 0 blocks it)."""
    h = 16
    P = np.zeros(mlp.nparam(h))
    W, b = mlp.unpack(P, h)
    b[3][:4] = [0.5, -1.0, 2.0, 0.25]
    b[1][:] = 0.3
    rng = np.random.default_rng(1)
    X = rng.normal(size=(16, mlp.NF))
    T = rng.normal(size=(16, mlp.NO))
    loss, G = mlp.loss_grad(P, h, X, T, bf16=False)
    d = b[3][:4] - T
    assert np.isclose(loss, (d * d).sum() / 64.0, rtol=1e-14)
    gW, gb = mlp.unpack(G, h)
    assert np.allclose(gb[3][:4], 2.0 * d.sum(axis=0) / 64.0, rtol=1e-14)
    a2 = mlp.gelu(mlp.gelu(np.full(h, 0.3)) @ W[2] + b[2])  # hidden layer 3 activations
    assert np.allclose(gW[3][:, :4], np.outer(a2, 2.0 * d.sum(axis=0) / 64.0), rtol=1e-12)
    assert not gb[0].any() and not gb[1].any() and not gb[2].any()


def test_bf16_emulation_stays_close_to_exact():
    h = 64
    rng = np.random.default_rng(2)
    P = mlp.random_params(h, rng)
    X = rng.normal(size=(64, mlp.NF))
    T = rng.normal(size=(64, mlp.NO))
    y1, _ = mlp.forward(P, h, X, bf16=True)
    y0, _ = mlp.forward(P, h, X, bf16=False)
    assert np.linalg.norm(y1 - y0) <= 0.05 * np.linalg.norm(y0)
    _, g1 = mlp.loss_grad(P, h, X, T, bf16=True)
    _, g0 = mlp.loss_grad(P, h, X, T, bf16=False)
    assert np.linalg.norm(g1 - g0) <= 0.05 * np.linalg.norm(g0)
