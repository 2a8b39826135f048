"""Learned-closure pieces of the oracle (CPU).

The collision with per-cell third/fourth-order rates: unit rates are the
reference exactly; mass and momentum gain do not depend on the rates; the
post-collision central moments (brute-force sums, no force) are (1 - s3) k3
and rho/9 + (1 - s4)(k4 - rho/9) while second order is untouched.  The MLP
gradients against central finite differences in FP64.  A planted wrong
GELU derivative must fail the gradient check.
"""

import numpy as np
import pytest


def _state(oracle, shape=(3, 2, 2), seed=0, amp=0.03):
    rng = np.random.default_rng(seed)
    f, *_ = oracle.init_feq(*shape, 0.04, -0.01)
    return f * (1 + amp * rng.standard_normal(f.shape))


def _cm(oracle, f, u, v, w, p, q, r):
    xi = [oracle.EX - u[..., None], oracle.EY - v[..., None], oracle.EZ - w[..., None]]
    return (f * xi[0] ** p * xi[1] ** q * xi[2] ** r).sum(-1)


def test_unit_rates_are_the_reference(oracle):
    f = _state(oracle)
    F = [np.full(f.shape[:3], 1e-3), np.full(f.shape[:3], -5e-4), np.zeros(f.shape[:3])]
    rho, u, v, w = oracle.macro_with_force(f, *F)
    a = oracle.collide(f, rho, u, v, w, *F, 1.7, 1.0)
    b = oracle.collide(f, rho, u, v, w, *F, 1.7, 1.0, rates=(1.0, 1.0, 1.0, 1.0))
    np.testing.assert_array_equal(a, b)


def test_rates_relax_the_higher_moments(oracle):
    f = _state(oracle, seed=1)
    z = np.zeros(f.shape[:3])
    rho, u, v, w = oracle.macro_with_force(f, z, z, z)
    rng = np.random.default_rng(2)
    rates = tuple(rng.uniform(0.3, 1.7, f.shape[:3]) for _ in range(4))
    g = oracle.collide(f, rho, u, v, w, z, z, z, 1.6, 1.0, rates=rates)
    ref = oracle.collide(f, rho, u, v, w, z, z, z, 1.6, 1.0)
    np.testing.assert_allclose(g.sum(-1), f.sum(-1), rtol=0, atol=1e-14)
    for e in (oracle.EX, oracle.EY, oracle.EZ):
        np.testing.assert_allclose((g * e).sum(-1), (ref * e).sum(-1), rtol=0, atol=1e-14)
    for pqr in [(1, 1, 0), (2, 0, 0), (0, 1, 1), (0, 0, 2)]:
        np.testing.assert_allclose(_cm(oracle, g, u, v, w, *pqr), _cm(oracle, ref, u, v, w, *pqr),
                                   rtol=0, atol=1e-14)
    s3xy, s3z, s4xy, s4z = rates
    for pqr, s in [((2, 1, 0), s3xy), ((1, 2, 0), s3xy), ((2, 0, 1), s3z), ((0, 2, 1), s3z),
                   ((1, 0, 2), s3z), ((0, 1, 2), s3z)]:
        np.testing.assert_allclose(_cm(oracle, g, u, v, w, *pqr),
                                   (1 - s) * _cm(oracle, f, u, v, w, *pqr), rtol=0, atol=1e-14)
    for pqr, s in [((2, 2, 0), s4xy), ((2, 0, 2), s4z), ((0, 2, 2), s4z)]:
        k4 = _cm(oracle, f, u, v, w, *pqr)
        np.testing.assert_allclose(_cm(oracle, g, u, v, w, *pqr), rho / 9 + (1 - s) * (k4 - rho / 9),
                                   rtol=0, atol=1e-14)


def test_features_of_rest_and_sheared_states(oracle):
    rho = 1.02
    f = np.broadcast_to(rho * oracle.W, (2, 1, 1, 19)).copy()
    z = np.zeros((2, 1, 1))
    X = oracle.central_features(f, z, z, z)
    want = np.zeros(19)            # at rest every central moment is at its equilibrium
    want[0] = rho - 1
    np.testing.assert_allclose(X[0, 0, 0], want, rtol=0, atol=1e-15)
    # a pure xy shear: f7, f10 up, f8, f9 down keeps rho and u, moves k110 only
    g = f.copy()
    g[..., 7] += 1e-3
    g[..., 10] += 1e-3
    g[..., 8] -= 1e-3
    g[..., 9] -= 1e-3
    Y = oracle.central_features(g, z, z, z)
    np.testing.assert_allclose(Y[..., 4] - X[..., 4], 4e-3 / rho, rtol=1e-12)
    np.testing.assert_allclose(np.delete(Y - X, 4, axis=-1), 0, atol=1e-15)


def _params(rng, scale=0.3):
    return [scale * rng.standard_normal(s) for s in [(19, 32), (32,), (32, 32), (32,), (32, 4), (4,)]]


def _fd_check(oracle, params, X, T, wgt, grads, eps=1e-6, n=40, rng=None):
    rng = rng or np.random.default_rng(5)
    worst = 0.0
    for _ in range(n):
        i = int(rng.integers(len(params)))
        idx = tuple(int(rng.integers(d)) for d in params[i].shape)
        p_plus = [p.copy() for p in params]
        p_minus = [p.copy() for p in params]
        p_plus[i][idx] += eps
        p_minus[i][idx] -= eps
        lp, _ = oracle.closure_loss(p_plus, X, T, wgt)
        lm, _ = oracle.closure_loss(p_minus, X, T, wgt)
        num = (lp - lm) / (2 * eps)
        worst = max(worst, abs(num - grads[i][idx]) / max(1e-8, abs(num)))
    return worst


@pytest.mark.parametrize("wshape", [(50,), (50, 4)])
def test_mlp_gradients_match_finite_differences(oracle, wshape):
    rng = np.random.default_rng(3)
    params = _params(rng)
    X = rng.standard_normal((50, 19))
    T = 0.1 * rng.standard_normal((50, 4))
    wgt = rng.uniform(0.5, 2.0, wshape)
    _, grads = oracle.closure_loss(params, X, T, wgt)
    assert _fd_check(oracle, params, X, T, wgt, grads) < 1e-6


def test_per_output_weights_reduce_to_per_sample_ones(oracle):
    """Equal weights across the four outputs give the per-sample loss, and a
    zero weight removes that output's error from the loss entirely."""
    rng = np.random.default_rng(6)
    params = _params(rng)
    X = rng.standard_normal((40, 19))
    T = 0.1 * rng.standard_normal((40, 4))
    w = rng.uniform(0.5, 2.0, 40)
    l1, g1 = oracle.closure_loss(params, X, T, w)
    l4, g4 = oracle.closure_loss(params, X, T, np.repeat(w[:, None], 4, 1) / 4)
    assert abs(l1 - 4 * l4) < 1e-12 * l1
    for a, b in zip(g1, g4):
        np.testing.assert_allclose(a, 4 * b, rtol=1e-12, atol=1e-15)
    w4 = np.ones((40, 4))
    w4[:, 2] = 0.0
    T2 = T.copy()
    T2[:, 2] += 5.0                      # an output with zero weight cannot matter
    assert oracle.closure_loss(params, X, T, w4)[0] == pytest.approx(
        oracle.closure_loss(params, X, T2, w4)[0], rel=1e-14)


def test_planted_gelu_derivative_is_caught(oracle, monkeypatch):
    rng = np.random.default_rng(4)
    params = _params(rng)
    X = rng.standard_normal((30, 19))
    T = 0.1 * rng.standard_normal((30, 4))
    wgt = np.ones(30)
    monkeypatch.setattr(oracle, "gelu_grad", lambda x: 0.5 * (1.0 + np.tanh(0.79788 * x)))
    _, grads = oracle.closure_loss(params, X, T, wgt)
    assert _fd_check(oracle, params, X, T, wgt, grads) > 1e-3


def test_fallback_to_plain_rates(oracle):
    s = oracle.closure_rates(np.array([[0.1, -0.2, 0.3, 0.0], [0.5, 1.2, 0.0, 0.0]]))
    np.testing.assert_allclose(s, [[1.1, 0.8, 1.3, 1.0], [1.0, 1.0, 1.0, 1.0]])
