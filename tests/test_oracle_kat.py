"""The oracle's primitives against the reference's own known-answer tests and
against what the mathematics fixes (CPU only).

Reference tests restated here:
  newton.rs:479-485  impedance_curve
  math.rs:469-516    quat_mat_roundtrip, block_cholesky_matches_dense_on_block_diagonal, cholesky_solves
  simd.rs:78-88      dot_matches_scalar
Added pins: correctly rounded sin/cos against exact decimal arithmetic, Cholesky
against numpy, the strided-accumulator order of `dot` for n >= 8, quaternion
algebra against closed forms.
"""
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

from oracle import pyoracle as po


# ---- newton.rs:479-485 ----
def test_impedance_curve():
    s = [0.9, 0.95, 0.001, 0.5, 2.0]
    assert abs(po.impedance(s, 0.0) - 0.9) < 1e-12
    assert abs(po.impedance(s, 0.0005) - 0.925) < 1e-12
    assert abs(po.impedance(s, 0.01) - 0.95) < 1e-12


def test_impedance_shape():
    # midpoint-symmetric sigmoid: imp(w*mid) = d0 + mid*(d1-d0) for power p at x = mid
    s = [0.5, 0.9, 0.01, 0.3, 3.0]
    assert po.impedance(s, 0.003) == pytest.approx(0.5 + 0.4 * 0.3, abs=1e-14)
    # monotone between dmin and dmax, saturating beyond the width, symmetric in |pos|
    xs = np.linspace(0, 0.02, 201)
    v = np.array([po.impedance(s, x) for x in xs])
    assert np.all(np.diff(v) >= -1e-15) and v[0] == pytest.approx(0.5) and v[-1] == pytest.approx(0.9)
    assert po.impedance(s, -0.004) == po.impedance(s, 0.004)


# ---- math.rs:469-477 ----
def test_quat_mat_roundtrip():
    q = po.quat_from_axis_angle([1.0, 2.0, 3.0], 1.1)
    q2 = po.quat_from_mat(po.quat_to_mat(q))
    assert np.max(np.abs(q - q2)) < 1e-12


def test_quat_closed_forms():
    # axis-angle about z by theta: (cos t/2, 0, 0, sin t/2) and Rz(theta)
    for t in (0.0, 0.3, -1.7, 3.0):
        q = po.quat_from_axis_angle([0.0, 0.0, 2.0], t)
        assert np.allclose(q, [math.cos(t / 2), 0, 0, math.sin(t / 2)], atol=1e-15)
        m = po.quat_to_mat(q).reshape(3, 3)
        rz = np.array([[math.cos(t), -math.sin(t), 0], [math.sin(t), math.cos(t), 0], [0, 0, 1]])
        assert np.allclose(m, rz, atol=2e-16 * 8)
    # integrate_local: q * exp(w h / 2) for a body-frame angular velocity
    q = po.quat_from_axis_angle([1.0, 0.0, 0.0], 0.4)
    out = po.quat_integrate(q, [0.0, 0.0, 2.0], 0.05)
    # expected: rotation about x by 0.4, then about the body z by 0.1
    rx = po.quat_to_mat(q).reshape(3, 3)
    rz = po.quat_to_mat(po.quat_from_axis_angle([0, 0, 1.0], 0.1)).reshape(3, 3)
    assert np.allclose(po.quat_to_mat(out).reshape(3, 3), rx @ rz, atol=1e-15)
    assert abs(np.linalg.norm(out) - 1.0) < 1e-15


# ---- math.rs:479-500 ----
def test_block_cholesky_matches_dense_on_block_diagonal():
    n = 5
    a = np.zeros(n * n)
    blocks = [(0, 2), (2, 5)]
    for (b0, b1) in blocks:
        for i in range(b0, b1):
            for j in range(b0, b1):
                a[i * n + j] = 4.0 + i if i == j else 0.3 + 0.1 * (i + j)
    rhs = 1.0 + np.arange(n, dtype=np.float64)
    dense, x1 = a.copy(), rhs.copy()
    assert po.cholesky_block(dense, n, 0, n)
    po.cholesky_solve_block(dense, n, 0, n, x1)
    blk, x2 = a.copy(), rhs.copy()
    for (b0, b1) in blocks:
        assert po.cholesky_block(blk, n, b0, b1)
        po.cholesky_solve_block(blk, n, b0, b1, x2)
    assert np.max(np.abs(x1 - x2)) < 1e-14
    # and against the textbook solution
    assert np.allclose(a.reshape(n, n) @ x2, rhs, atol=1e-13)


# ---- math.rs:502-515 ----
def test_cholesky_solves():
    a = np.array([4.0, 2.0, 0.6, 2.0, 5.0, 1.0, 0.6, 1.0, 3.0])
    orig = a.copy()
    assert po.cholesky_block(a, 3, 0, 3)
    x = np.array([1.0, 2.0, 3.0])
    po.cholesky_solve_block(a, 3, 0, 3, x)
    assert np.max(np.abs(orig.reshape(3, 3) @ x - [1.0, 2.0, 3.0])) < 1e-12


@pytest.mark.parametrize("n", [1, 3, 8, 14, 15, 36])
def test_cholesky_matches_numpy(n):
    rng = np.random.default_rng(n)
    b = rng.normal(size=(n, n))
    a = b @ b.T + n * np.eye(n)
    l = a.reshape(-1).copy()
    assert po.cholesky_block(l, n, 0, n)
    L = np.tril(l.reshape(n, n))
    ref = np.linalg.cholesky(a)
    assert np.max(np.abs(L - ref)) < 1e-12 * np.max(np.abs(ref))
    x = rng.normal(size=n)
    y = x.copy()
    po.cholesky_solve_block(l, n, 0, n, y)
    assert np.allclose(y, np.linalg.solve(a, x), rtol=1e-11, atol=1e-13)


def test_cholesky_rejects_indefinite():
    a = np.array([1.0, 2.0, 2.0, 1.0])
    assert not po.cholesky_block(a, 2, 0, 2)


# ---- simd.rs:78-88 ----
def test_dot_matches_scalar():
    for n in range(40):
        a = np.sin(np.arange(n) * 0.37)
        b = np.cos(np.arange(n) * 0.91)
        s = math.fsum(a * b)
        assert abs(po.dot(a, b) - s) < 1e-12, n


def test_dot_uses_four_strided_accumulators():
    # simd.rs:47-63: for n >= 8 lanes k, k+4, ... accumulate separately and
    # combine as (s0+s1)+(s2+s3), then the tail. Values chosen so that the
    # sequential order gives 0 and the strided order gives 1 (exact in FP64).
    big = 2.0 ** 53
    # s0 = big + 0, s1 = 1 + 1 = 2, (s0 + s1) = big + 2 exactly;
    # a sequential sum gives ((big + 1) + 1) = big (ties to even)
    a = np.array([big, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0])
    assert po.dot(a, np.ones(8)) == big + 2.0
    assert ((big + 1.0) + 1.0) == big
    # n = 9: the strided part gives big + 2, the tail element is added last
    a9 = np.append(a, -big)
    assert po.dot(a9, np.ones(9)) == 2.0
    # short vectors (n < 8) are summed in order from +0.0 (simd.rs:17-22)
    assert math.copysign(1.0, po.dot(np.array([-0.0]), np.array([1.0]))) == 1.0
    v = np.array([big, 1.0, 1.0])
    assert po.dot(v, np.ones(3)) == big


# ---- correctly rounded sin/cos (shared by the oracle and the device) ----
def _dec_sincos(x: float):
    getcontext().prec = 80
    d = Decimal(x)
    # reduce by 2*pi in high precision
    pi = Decimal("3.14159265358979323846264338327950288419716939937510582097494459230781640628620899")
    k = (d / (2 * pi)).to_integral_value()
    r = d - k * 2 * pi
    def series(term, n):
        acc = Decimal(0)
        while term != 0 and (acc == 0 or abs(term) > abs(acc) * Decimal(10) ** -75):
            acc += term
            term = -term * r * r / ((n + 1) * (n + 2))
            n += 2
        return acc

    return float(series(r, 1)), float(series(Decimal(1), 0))


def test_cr_sincos_is_correctly_rounded():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-8, 8, 400), rng.uniform(-1e-2, 1e-2, 50), rng.uniform(-40, 40, 50),
                         [0.0, 1e-300, math.pi / 4, math.pi / 2, 1.0, -2.5, 3.0, 0.5 * 0.785398163397]])
    s, c = po.cr_sincos(xs)
    for x, si, ci in zip(xs, s, c):
        es, ec = _dec_sincos(float(x))
        assert si == es, (x, si, es)
        assert ci == ec, (x, ci, ec)
    s0, c0 = po.cr_sincos(np.array([-0.0]))
    assert math.copysign(1.0, s0[0]) == -1.0 and c0[0] == 1.0
