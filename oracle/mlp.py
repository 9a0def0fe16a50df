"""ORACLE — test infrastructure only. FP64 restatement of the learned Newton
warm-start network (pippin_amd/csrc/mlp.hpp, mlp_dev.hpp, mlp.hip).

The network has no counterpart in the reference (its warm start is the previous
acceleration, crates/pippin/src/newton.rs:238-245), so this file restates the
design, not reference code:

    24 -> H -> H -> H -> 4, GELU (tanh form) after the hidden layers,
    loss = sum (y - t)^2 / (4 n) over a batch of n samples.

Parameters use the device's flat layout: W1 [32][H] b1 [H] W2 [H][H] b2
W3 [H][H] b3 W4 [H][16] b4 [16], row-major [in][out]; inputs 24..31 and outputs
4..15 are zero padding.

The device rounds to BF16 (round to nearest even) at fixed points (inputs,
weights, hidden activations, the output gradient and each pre-activation
gradient) and accumulates in FP32. With ``bf16=True`` this restatement rounds at
the same points and accumulates in FP64, so the two differ by FP32 accumulation
(and the rare BF16 rounding that it flips). With ``bf16=False`` it is the exact
FP64 function: tests pin its gradient by central finite differences, and
``bf16_round`` against torch's own float32 -> bfloat16 conversion.
"""
from __future__ import annotations

import numpy as np

NF, KIN, NO, NOP = 24, 32, 4, 16
C_GELU = 0.7978845608028654  # sqrt(2 / pi)


def fan_in(l, h):
    return KIN if l == 0 else h


def fan_out(l, h):
    return NOP if l == 3 else h


def woff(l, h):
    o = 0
    for k in range(l):
        o += fan_in(k, h) * fan_out(k, h) + fan_out(k, h)
    return o


def boff(l, h):
    return woff(l, h) + fan_in(l, h) * fan_out(l, h)


def nparam(h):
    return woff(4, h)


def unpack(P, h):
    """Views W[l] (in x out) and b[l] of the flat parameter vector."""
    W = [P[woff(l, h):boff(l, h)].reshape(fan_in(l, h), fan_out(l, h)) for l in range(4)]
    b = [P[boff(l, h):boff(l, h) + fan_out(l, h)] for l in range(4)]
    return W, b


def random_params(h, rng, scale=1.0):
    """Parameters with the padding at zero (weights U(+-scale/sqrt(fan_in)), biases U(+-0.1))."""
    P = np.zeros(nparam(h), dtype=np.float32)
    W, b = unpack(P, h)
    for l in range(4):
        fin = NF if l == 0 else h
        w = rng.uniform(-1, 1, size=W[l].shape) * scale / np.sqrt(fin)
        if l == 0:
            w[NF:, :] = 0
        if l == 3:
            w[:, NO:] = 0
        W[l][:] = w
        bb = rng.uniform(-0.1, 0.1, size=b[l].shape)
        if l == 3:
            bb[NO:] = 0
        b[l][:] = bb
    return P


def bf16_round(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r).astype(np.uint32) << 16
    return r.view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3.0 * 0.044715 * z * z)


def _f32(x):
    # values the device holds in FP32 registers before rounding to BF16
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def forward(P, h, X, bf16=True):
    """Outputs (n x 4) and the intermediates of the pass."""
    P = np.asarray(P, dtype=np.float64)
    W, b = unpack(P, h)
    rnd = bf16_round if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    n = X.shape[0]
    x0 = np.zeros((n, KIN))
    x0[:, :NF] = X
    a = rnd(x0)
    Wr = [rnd(w) for w in W]
    bb = [np.asarray(v, dtype=np.float64) for v in b]
    zs, acts = [], [a]
    for l in range(3):
        z = a @ Wr[l] + bb[l]
        if bf16:
            z = _f32(z)
        a = rnd(gelu(z))
        zs.append(z)
        acts.append(a)
    y = a @ Wr[3] + bb[3]
    return y[:, :NO], dict(zs=zs, acts=acts, W=Wr, y=y)


def loss_grad(P, h, X, T, bf16=True):
    """Mean squared error and its gradient (flat, the parameter layout)."""
    rnd = bf16_round if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    y4, c = forward(P, h, X, bf16)
    n = X.shape[0]
    d = y4 - T
    loss = float((d * d).sum() / (4.0 * n))
    dy = np.zeros((n, NOP))
    dy[:, :NO] = d * (2.0 / (4.0 * n))
    if bf16:
        dy = rnd(_f32(dy))
    G = np.zeros(nparam(h))
    gW, gb = unpack(G, h)
    dz = dy
    for l in range(3, -1, -1):
        gW[l][:] = c["acts"][l].T @ dz
        gb[l][:] = dz.sum(axis=0)
        if l == 0:
            break
        da = dz @ c["W"][l].T
        g = da * gelu_grad(c["zs"][l - 1])
        dz = rnd(_f32(g)) if bf16 else g
    return loss, G
