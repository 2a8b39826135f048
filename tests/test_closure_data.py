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
