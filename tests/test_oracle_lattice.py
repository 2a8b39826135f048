"""Oracle pinning: lattice constants, the raw-moment matrix and the equilibrium.

Each check is fixed by mathematics, not by retyping the oracle's formulas:
exact inverse, dyadic entries, lattice isotropy and the moments of the
second-order equilibrium.
"""

import itertools

import numpy as np


def test_direction_set_is_d3q19(oracle):
    e = np.stack([oracle.EX, oracle.EY, oracle.EZ], 1)
    assert len({tuple(v) for v in e}) == 19
    norms = (e ** 2).sum(1)
    assert sorted(norms.tolist()) == [0] + [1] * 6 + [2] * 12
    # every direction has its opposite
    for v in e:
        assert any((w == -v).all() for w in e)
    # reference ordering (clbm_moving.py:21-23): face pairs, then xy, xz, yz edges
    assert (e[1] == [1, 0, 0]).all() and (e[2] == [-1, 0, 0]).all()
    assert (e[7] == [1, 1, 0]).all() and (e[8] == [-1, 1, 0]).all() and (e[9] == [1, -1, 0]).all()
    assert (e[11] == [1, 0, 1]).all() and (e[15] == [0, 1, 1]).all() and (e[18] == [0, -1, -1]).all()


def test_weights_isotropy(oracle):
    e = np.stack([oracle.EX, oracle.EY, oracle.EZ], 1).astype(float)
    w = oracle.W
    assert abs(w.sum() - 1) < 1e-15
    cs2 = 1.0 / 3.0
    for a, b in itertools.product(range(3), repeat=2):
        assert abs((w * e[:, a] * e[:, b]).sum() - cs2 * (a == b)) < 1e-15
    for a, b, c in itertools.product(range(3), repeat=3):
        assert abs((w * e[:, a] * e[:, b] * e[:, c]).sum()) < 1e-15
    for a, b, c, d in itertools.product(range(3), repeat=4):
        iso = cs2 ** 2 * ((a == b) * (c == d) + (a == c) * (b == d) + (a == d) * (b == c))
        assert abs((w * e[:, a] * e[:, b] * e[:, c] * e[:, d]).sum() - iso) < 1e-15


def test_m_raw_inverse_is_exact_and_dyadic(oracle):
    M = oracle.m_raw()
    Mi = oracle.m_raw_inv_f32().astype(np.float64)
    assert np.array_equal(M @ Mi, np.eye(19))
    assert np.array_equal(Mi @ M, np.eye(19))
    vals = set(np.unique(np.abs(Mi)).tolist())
    assert vals <= {0.0, 0.25, 0.5, 1.0}
    assert np.count_nonzero(Mi) == 91
    # raw-moment ordering of config_moving.py:62-68
    assert oracle.PQR[0] == (0, 0, 0) and oracle.PQR[10] == (2, 1, 0) and oracle.PQR[18] == (0, 2, 2)


def test_init_feq_moments(oracle):
    ux, uy = 0.05, -0.012
    f, rho, u, v, w = oracle.init_feq(3, 2, 2, ux, uy)
    g = f[0, 0, 0]
    e = np.stack([oracle.EX, oracle.EY, oracle.EZ], 1).astype(float)
    assert abs(g.sum() - 1) < 1e-15
    np.testing.assert_allclose(e.T @ g, [ux, uy, 0.0], atol=1e-16)
    P = (e[:, :, None] * e[:, None, :] * g[:, None, None]).sum(0)
    uvec = np.array([ux, uy, 0.0])
    np.testing.assert_allclose(P, np.eye(3) / 3 + np.outer(uvec, uvec), atol=1e-16)
    assert (rho == 1).all() and (u == ux).all() and (v == uy).all() and (w == 0).all()
