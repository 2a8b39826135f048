"""The closure's training-data pipeline (CPU, against the oracle).

Target extraction is pinned by construction: when the later state is the
oracle's collision with known per-cell rates followed by periodic streaming,
the extracted targets are those rates.  The box filter is the block mean;
folding the input standardisation into the first layer leaves the network's
output unchanged.
"""

import numpy as np

from conftest import load_sub


def test_targets_recover_known_rates(oracle):
    cl = load_sub("closure")
    rng = np.random.default_rng(0)
    shape = (6, 5, 4)
    f, *_ = oracle.init_feq(*shape, 0.04, 0.01)
    f = f * (1 + 0.05 * rng.standard_normal(f.shape))
    s_true = rng.uniform(0.2, 1.8, shape + (4,))
    z = np.zeros(shape)
    rho, u, v, w = oracle.macro_with_force(f, z, z, z)
    post = oracle.collide(f, rho, u, v, w, z, z, z, 1.9, 1.0,
                          rates=tuple(s_true[..., i] for i in range(4)))
    nxt = oracle.stream(post)
    X, T, W = cl.features_and_targets(f, nxt)
    np.testing.assert_allclose(T + 1.0, s_true.reshape(-1, 4), rtol=0, atol=1e-9)
    np.testing.assert_allclose(X, oracle.central_features(f, z, z, z).reshape(-1, 19), atol=1e-14)
    assert np.all(W > 0)


def test_targets_of_the_plain_cascade_are_zero(oracle):
    cl = load_sub("closure")
    rng = np.random.default_rng(1)
    shape = (5, 4, 3)
    f, *_ = oracle.init_feq(*shape, 0.03, 0.0)
    f = f * (1 + 0.05 * rng.standard_normal(f.shape))
    z = np.zeros(shape)
    rho, u, v, w = oracle.macro_with_force(f, z, z, z)
    nxt = oracle.stream(oracle.collide(f, rho, u, v, w, z, z, z, 1.5, 1.0))
    _, T, _ = cl.features_and_targets(f, nxt)
    np.testing.assert_allclose(T, 0.0, atol=1e-9)


def test_box_filter_is_the_block_mean():
    cl = load_sub("closure")
    a = np.random.default_rng(2).standard_normal((4, 6, 2, 19))
    b = cl.box_filter(a)
    assert b.shape == (2, 3, 1, 19)
    np.testing.assert_allclose(b[1, 2, 0], a[2:4, 4:6, 0:2].mean(axis=(0, 1, 2)))


def test_folded_standardisation_gives_the_same_outputs(oracle):
    cl = load_sub("closure")
    rng = np.random.default_rng(3)
    p = cl.init_params(4, out_scale=0.5)
    X = rng.standard_normal((100, 19)) * rng.uniform(0.01, 3, 19) + rng.standard_normal(19)
    mean, std = X.mean(0), X.std(0)
    y1, _ = oracle.mlp_forward(cl.split_params(p), (X - mean) / std)
    y2, _ = oracle.mlp_forward(cl.split_params(cl.fold_standardization(p, mean, std)), X)
    np.testing.assert_allclose(y1, y2, rtol=1e-10, atol=1e-12)


def _rest(shape):
    """Populations at rest, rho = 1: every non-equilibrium moment is zero."""
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    return np.broadcast_to(w, shape + (19,)).copy()


def test_diagnostic_energies_of_known_perturbations():
    """At rest every order carries zero energy.  Adding d to f(1,1,0) and
    f(-1,-1,0) and taking it from f(-1,1,0) and f(1,-1,0) changes only k110,
    by 4d: order-2 energy 16 d^2 in that cell.  Adding d to f(1,1,0) and
    f(-1,1,0) and taking 2d from f(0,1,0) changes k200, k210 and k220 by 2d
    each (mass and momentum unchanged): 4 d^2 in each of orders 2, 3, 4."""
    cl = load_sub("closure")
    shape = (4, 3, 2)
    f = _rest(shape)
    d = cl.diagnostics(f)
    assert d["cells"] == 24 and all(abs(v) < 1e-30 for v in d["energy"].values())
    dl = 1e-3
    g = f.copy()
    g[1, 1, 0, [7, 10]] += dl
    g[1, 1, 0, [8, 9]] -= dl
    e = cl.diagnostics(g)["energy"]
    np.testing.assert_allclose([e["2"] * 24, e["3"], e["4"]], [16 * dl ** 2, 0, 0], atol=1e-18,
                               rtol=1e-9)
    h = f.copy()
    h[2, 0, 1, [7, 8]] += dl
    h[2, 0, 1, 3] -= 2 * dl
    e = cl.diagnostics(h)["energy"]
    np.testing.assert_allclose([e["2"] * 24, e["3"] * 24, e["4"] * 24], [4 * dl ** 2] * 3,
                               rtol=1e-9)
    # the mask selects cells: only the perturbed cell, then none of it
    m = np.zeros(shape, bool)
    m[2, 0, 1] = True
    np.testing.assert_allclose(cl.diagnostics(h, mask=m)["energy"]["3"], 4 * dl ** 2, rtol=1e-9)
    m[:] = False
    m[0, 0, 0] = True
    assert cl.diagnostics(h, mask=m)["energy"]["3"] < 1e-30
