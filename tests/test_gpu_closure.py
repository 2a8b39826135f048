"""GPU parity of the learned-closure kernels.

Inference against a NumPy forward that rounds weights, inputs and hidden
activations to bf16 as the kernels do (tolerance: a few bf16 ulps of the
output scale, from rounding ties of hidden activations that may fall either
way); gradients against the FP64 oracle (relative L2 error 3e-2 and cosine
0.999, the bf16 resolution); bitwise-deterministic gradients; training that
fits a teacher; the closure step with a zero last layer (s = 1) against the
plain step (2e-6); a closure step against the FP64 oracle collision with the
rates of the bf16 network (1e-6).
"""


import numpy as np
import pytest

from conftest import load_sub

pytestmark = pytest.mark.gpu


def _bf16(a):
    a = np.asarray(a, dtype=np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def _emulate(oracle, p, X):
    W1, b1, W2, b2, W3, b3 = load_sub("closure").split_params(np.asarray(p, np.float64))
    Xb = _bf16(X)
    z1 = Xb @ _bf16(W1) + b1
    h1 = _bf16(oracle.gelu(z1))
    z2 = h1 @ _bf16(W2) + b2
    h2 = _bf16(oracle.gelu(z2))
    return h2 @ _bf16(W3) + b3


def _params(seed, scale=1.0):
    cl = load_sub("closure")
    return cl.init_params(seed, out_scale=scale)


def test_infer_matches_bf16_forward(pkg, oracle, gpu):
    cl = load_sub("closure")
    p = _params(1)
    X = np.random.default_rng(2).standard_normal((1000, 19))
    net = cl.ClosureNet(p)
    Y = net.infer(X).cpu().numpy()
    ref = _emulate(oracle, p.astype(np.float32), X.astype(np.float32))
    scale = np.abs(ref).max()
    np.testing.assert_allclose(Y, ref, rtol=0, atol=2e-2 * scale)
    assert np.median(np.abs(Y - ref)) < 1e-3 * scale


@pytest.mark.parametrize("wshape", [(3000,), (3000, 4)])
def test_gradient_matches_fp64_oracle_and_is_deterministic(pkg, oracle, gpu, wshape):
    cl = load_sub("closure")
    rng = np.random.default_rng(3)
    X = rng.standard_normal((3000, 19))
    T = 0.2 * rng.standard_normal((3000, 4))
    W = rng.uniform(0.5, 1.5, wshape)
    tr = cl.Trainer(X, T, W, blocks=16, params=_params(4))
    g1 = tr.gradient().cpu().numpy().astype(np.float64)
    g2 = tr.gradient().cpu().numpy().astype(np.float64)
    np.testing.assert_array_equal(g1, g2)
    # the trainer feeds standardised inputs
    Xs = (X - tr.mean) / tr.std
    p = tr.net.numpy()
    loss, grads = oracle.closure_loss(cl.split_params(p), Xs, T, W)
    ref = np.concatenate([g.ravel() for g in grads])
    got = g1[:cl.NP]
    assert abs(g1[cl.NP] - loss) < 2e-2 * loss
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    cos = got @ ref / (np.linalg.norm(got) * np.linalg.norm(ref))
    assert rel < 3e-2 and cos > 0.999, (rel, cos)


def test_training_fits_a_teacher(pkg, oracle, gpu):
    cl = load_sub("closure")
    rng = np.random.default_rng(5)
    X = rng.standard_normal((20000, 19))
    teacher = cl.split_params(_params(6, scale=0.3))
    T, _ = oracle.mlp_forward(teacher, X)
    tr = cl.Trainer(X, T, np.ones(len(X)), blocks=64, lr=3e-3, params=_params(7, scale=0.1))
    l0 = tr.loss()
    hist = tr.fit(epochs=60, batch=4096)
    assert hist[-1] < 0.1 * l0, (l0, hist[-5:])
    assert all(b <= a * 1.05 for a, b in zip(hist, hist[1:]))   # no blow-up along the way


def _case(pkg, shape):
    return pkg.config.CaseConfig(name="cl", Nx=shape[0], Ny=shape[1], Nz=shape[2], U_inf=0.05,
                                 Re=200, chord=10, s_v=1 / 0.56, ux_in=0.05, uy_in=-0.007)


def _state(oracle, shape, seed, amp=0.03):
    rng = np.random.default_rng(seed)
    f, *_ = oracle.init_feq(*shape, 0.05, -0.007)
    return f * (1 + amp * rng.standard_normal(f.shape))


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).transpose(2, 1, 0))).cuda()


def test_zero_output_network_is_the_plain_step(pkg, oracle, gpu):
    cl = load_sub("closure")
    shape = (13, 7, 5)           # 455 cells: a partial last wavefront
    cfg = _case(pkg, shape)
    a = pkg.CLBMSolver(cfg)
    b = pkg.CLBMSolver(cfg)
    f = _state(oracle, shape, 8)
    a.f.from_numpy(f)
    b.f.from_numpy(f)
    rng = np.random.default_rng(9)
    F = [(1e-3 * rng.standard_normal(shape)).astype(np.float32) for _ in range(2)]
    F[0][-1] = F[0][-2]          # matching outlet forces on most rows
    F[0][-1, 0, 0] += 1e-4       # ... and one row that differs
    tF = [_dev(x) for x in F]
    b.closure = cl.ClosureNet(_params(10, scale=0.0))
    for _ in range(4):
        a.step(*tF)
        b.step(*tF)
    np.testing.assert_allclose(b.f.to_numpy(), a.f.to_numpy(), rtol=0, atol=2e-6)


def test_closure_step_matches_oracle_rates(pkg, oracle, gpu):
    cl = load_sub("closure")
    shape = (10, 6, 4)
    cfg = _case(pkg, shape)
    s = pkg.CLBMSolver(cfg, inlet=False)
    f = _state(oracle, shape, 11)
    s.f.from_numpy(f)
    p = _params(12, scale=0.05)
    s.closure = cl.ClosureNet(p)
    z = np.zeros(shape)
    s.step()
    X = oracle.central_features(f.astype(np.float32).astype(np.float64), z, z, z)
    Y = _emulate(oracle, p.astype(np.float32), X.reshape(-1, 19).astype(np.float32))
    rates = oracle.closure_rates(Y).reshape(shape + (4,))
    want, *_ = oracle.lbm_step(f.astype(np.float32).astype(np.float64), z, z, z, cfg.s_v, cfg.s_e,
                               cfg.ux_in, cfg.uy_in,
                               rates=tuple(rates[..., i] for i in range(4)), inlet=False)
    np.testing.assert_allclose(s.f.to_numpy(), want, rtol=0, atol=1e-6)
    # the rates did move the populations away from the plain cascade
    plain, *_ = oracle.lbm_step(f, z, z, z, cfg.s_v, cfg.s_e, cfg.ux_in, cfg.uy_in, inlet=False)
    assert np.abs(want - plain).max() > 1e-7


def test_learned_closure_beats_plain_when_the_data_has_one(pkg, oracle, gpu):
    """End to end: states advanced by the oracle with state-dependent
    third/fourth-order rates (a fixed smooth function of the features) give
    targets, the GPU trains on them, and the fused closure step then predicts
    held-out states with at most half the one-step error of the plain
    cascade (measured: 2.5x smaller)."""
    cl = load_sub("closure")
    rng = np.random.default_rng(21)
    shape = (40, 30, 16)
    cfg = pkg.config.CaseConfig(name="cl", Nx=shape[0], Ny=shape[1], Nz=shape[2], U_inf=0.05,
                                Re=200, chord=10, s_v=1.9, ux_in=0.04, uy_in=0.0)
    truth = cl.split_params(cl.init_params(22, out_scale=0.15))
    z = np.zeros(shape)

    def pair(seed):
        r = np.random.default_rng(seed)
        f, *_ = oracle.init_feq(*shape, 0.04, 0.01)
        f = f * (1 + 0.03 * r.standard_normal(f.shape))
        X = oracle.central_features(f, z, z, z).reshape(-1, 19)
        mean, std = X.mean(0), X.std(0) + 1e-30
        y, _ = oracle.mlp_forward(truth, (X - mean) / std)
        s = np.clip(1.0 + y, 0.05, 1.95).reshape(shape + (4,))
        rho, u, v, w = oracle.macro_with_force(f, z, z, z)
        nxt = oracle.stream(oracle.collide(f, rho, u, v, w, z, z, z, cfg.s_v, cfg.s_e,
                                           rates=tuple(s[..., i] for i in range(4))))
        return f, nxt

    Xs, Ts, Ws = [], [], []
    for seed in range(3):
        a, b = pair(seed)
        X, T, W = cl.features_and_targets(a, b)
        Xs.append(X)
        Ts.append(T)
        Ws.append(np.ones(len(X)))
    tr = cl.Trainer(np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ws), blocks=128)
    tr.fit(epochs=40, batch=8192)
    net = tr.export()
    a, b = pair(99)
    e_plain = cl.one_step_error(cfg, a, b, None)
    e_net = cl.one_step_error(cfg, a, b, net)
    assert e_net < 0.5 * e_plain, (e_plain, e_net)


def test_diagnostics_report_the_rates_the_network_gives(pkg, oracle, gpu):
    """A network with zero weights in its last layer and bias b gives s = 1 + b
    in every cell: b = 0 is the plain cascade (no correction), b = 0.3 moves
    every rate by 0.3 and changes the post-collision third/fourth-order
    moments by 0.3 times their pre-collision values (correction energy
    0.09 x the order-3 + order-4 energy), and b = 1.5 (s = 2.5, outside
    [0, 2]) sends every cell to the fallback."""
    cl = load_sub("closure")
    f = _state(oracle, (9, 6, 4), 3)
    for b, on, dev in ((0.0, 1.0, 0.0), (0.3, 1.0, 0.3), (1.5, 0.0, 0.0)):
        p = cl.init_params(4, out_scale=0.0)
        p[-4:] = b
        d = cl.diagnostics(f, cl.ClosureNet(p))
        assert d["cells"] == 216
        assert d["on_network"] == on
        np.testing.assert_allclose(list(d["mean_abs_s_minus_1"].values()), [dev] * 4, atol=1e-6)
        np.testing.assert_allclose(list(d["max_abs_s_minus_1"].values()), [dev] * 4, atol=1e-6)
        e = d["energy"]
        assert e["3"] > 0 and e["4"] > 0
        np.testing.assert_allclose(d["correction_energy"], dev ** 2 * (e["3"] + e["4"]),
                                   rtol=1e-5, atol=1e-30)
