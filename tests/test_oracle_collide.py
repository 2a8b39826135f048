"""Oracle pinning: the cascaded collision and its forcing.

Each property is computed by brute-force moment sums over the 19 directions,
independently of the cascade's own central->raw->M^-1 path, and each test
that matters is paired with a planted mistake in the oracle (on the CPU) that
the same check must reject.
"""

import numpy as np
import pytest


def _e(oracle):
    return np.stack([oracle.EX, oracle.EY, oracle.EZ], 1).astype(float)


def _state(oracle, seed, shape=(3, 2, 2), amp=0.03):
    rng = np.random.default_rng(seed)
    f, *_ = oracle.init_feq(*shape, 0.04, -0.02)
    f = f * (1 + amp * rng.standard_normal(f.shape))
    F = [1e-3 * rng.standard_normal(shape) for _ in range(3)]
    return f, F


def _central(oracle, f, u):
    """sum_q f_q (e_q - u)^p for every (p,q,r) of the raw-moment basis."""
    e = _e(oracle)
    xi = e[None, None, None, :, :] - np.stack(u, -1)[..., None, :]
    out = {}
    for (p, q, r) in oracle.PQR:
        out[(p, q, r)] = (f * xi[..., 0] ** p * xi[..., 1] ** q * xi[..., 2] ** r).sum(-1)
    return out


def _collide(oracle, f, F, s_v, s_e=1.0):
    rho, u, v, w = oracle.macro_with_force(f, *F)
    return oracle.collide(f, rho, u, v, w, *F, s_v, s_e), (rho, u, v, w)


def check_mass_momentum(oracle, s_v=1 / 0.56):
    f, F = _state(oracle, 1)
    fp, (rho, *_) = _collide(oracle, f, F, s_v)
    e = _e(oracle)
    m_in = f @ e
    m_out = fp @ e
    np.testing.assert_allclose(fp.sum(-1), f.sum(-1), rtol=0, atol=1e-14)
    gain = 1.5 - s_v / 2
    np.testing.assert_allclose(m_out - m_in, gain * np.stack(F, -1), rtol=0, atol=1e-14)


def test_mass_and_momentum_gain(oracle):
    """Mass is conserved; momentum gains (1.5 - s_v/2) F per collision: F/2
    from the half-force shift in u plus (1 - s_v/2) F from Guo's term."""
    for s_v in (1 / 0.53, 1 / 0.56, 1.0, 1.2):
        check_mass_momentum(oracle, s_v)


def test_relaxed_central_moments(oracle):
    """Without force, the post-collision central moments about u are: shear
    x(1 - s_v), trace relaxed by s_e toward rho, deviators x(1 - s_v),
    the six third-order basis moments 0, the three fourth-order rho/9."""
    f, F = _state(oracle, 2)
    Z = [np.zeros_like(F[0])] * 3
    for s_v, s_e in [(1 / 0.56, 1.0), (1.3, 0.8), (0.0, 0.0)]:
        fp, (rho, u, v, w) = _collide(oracle, f, Z, s_v, s_e)
        k = _central(oracle, f, (u, v, w))
        kp = _central(oracle, fp, (u, v, w))
        for key in [(1, 1, 0), (1, 0, 1), (0, 1, 1)]:
            np.testing.assert_allclose(kp[key], (1 - s_v) * k[key], atol=1e-14)
        tr = k[(2, 0, 0)] + k[(0, 2, 0)] + k[(0, 0, 2)]
        trp = kp[(2, 0, 0)] + kp[(0, 2, 0)] + kp[(0, 0, 2)]
        np.testing.assert_allclose(trp, tr - s_e * (tr - rho), atol=1e-14)
        np.testing.assert_allclose(kp[(2, 0, 0)] - kp[(0, 2, 0)],
                                   (1 - s_v) * (k[(2, 0, 0)] - k[(0, 2, 0)]), atol=1e-14)
        np.testing.assert_allclose(kp[(2, 0, 0)] - kp[(0, 0, 2)],
                                   (1 - s_v) * (k[(2, 0, 0)] - k[(0, 0, 2)]), atol=1e-14)
        for key in [(2, 1, 0), (1, 2, 0), (2, 0, 1), (0, 2, 1), (1, 0, 2), (0, 1, 2)]:
            np.testing.assert_allclose(kp[key], 0.0, atol=1e-14)
        for key in [(2, 2, 0), (2, 0, 2), (0, 2, 2)]:
            np.testing.assert_allclose(kp[key], rho / 9, atol=1e-14)


def test_rest_state_is_fixed_point(oracle):
    rho = 1.07
    f = np.broadcast_to(rho * oracle.W, (2, 2, 2, 19)).copy()
    Z = [np.zeros((2, 2, 2))] * 3
    fp, _ = _collide(oracle, f, Z, 1 / 0.56)
    np.testing.assert_allclose(fp, f, atol=1e-15)


def test_full_relaxation_forgets_nonequilibrium(oracle):
    """At s_v = s_e = 1 the result depends only on rho and momentum."""
    f, F = _state(oracle, 3)
    Z = [np.zeros_like(F[0])] * 3
    e = _e(oracle)
    # a non-equilibrium perturbation with zero mass and momentum
    rng = np.random.default_rng(4)
    d = rng.standard_normal(19)
    A = np.vstack([np.ones(19), e.T])
    d -= A.T @ np.linalg.solve(A @ A.T, A @ d)
    g = f + 1e-3 * d
    fa, _ = _collide(oracle, f, Z, 1.0, 1.0)
    fb, _ = _collide(oracle, g, Z, 1.0, 1.0)
    np.testing.assert_allclose(fa, fb, atol=1e-15)


def test_guo_term_moments(oracle):
    """With u held fixed, collide(F) - collide(0) is Guo's term G: sum G = 0,
    sum e G = (1 - s_v/2) F, sum e e G = (1 - s_v/2)(u F + F u)."""
    f, F = _state(oracle, 5)
    s_v = 1 / 0.56
    rho, u, v, w = oracle.macro_with_force(f, *F)
    Z = [np.zeros_like(F[0])] * 3
    G = (oracle.collide(f, rho, u, v, w, *F, s_v) - oracle.collide(f, rho, u, v, w, *Z, s_v))
    e = _e(oracle)
    c = 1 - s_v / 2
    Fv = np.stack(F, -1)
    U = np.stack([u, v, w], -1)
    np.testing.assert_allclose(G.sum(-1), 0, atol=1e-16)
    np.testing.assert_allclose(G @ e, c * Fv, atol=1e-16)
    P = np.einsum("...q,qa,qb->...ab", G, e, e)
    np.testing.assert_allclose(P, c * (U[..., :, None] * Fv[..., None, :] + Fv[..., :, None] * U[..., None, :]),
                               atol=1e-16)


# --- planted mistakes: the checks above must reject a wrong oracle ---------
def test_planted_minv_row_swap_is_caught(oracle, monkeypatch):
    good = oracle.m_raw_inv_f32()
    bad = good.copy()
    bad[[1, 2]] = bad[[2, 1]]            # two populations exchanged
    monkeypatch.setattr(oracle, "m_raw_inv_f32", lambda: bad)
    with pytest.raises(AssertionError):
        check_mass_momentum(oracle)


def test_planted_guo_weight_is_caught(oracle, monkeypatch):
    bad = oracle.W.copy()
    bad[7], bad[8] = bad[8] * 1.5, bad[7]  # a wrong weight in the forcing
    monkeypatch.setattr(oracle, "W", bad)
    with pytest.raises(AssertionError):
        test_guo_term_moments(oracle)


def test_planted_relaxation_sign_is_caught(oracle, monkeypatch):
    real = oracle.collide

    def wrong(f, rho, ux, uy, uz, fx, fy, fz, s_v, s_e=1.0):
        return real(f, rho, ux, uy, uz, fx, fy, fz, 2.0 - s_v, s_e)  # relaxes with 1 - (2 - s_v)
    monkeypatch.setattr(oracle, "collide", wrong)
    with pytest.raises(AssertionError):
        test_relaxed_central_moments(oracle)
