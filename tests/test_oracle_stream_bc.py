"""Oracle pinning: streaming, inlet and outlet rules, the driver-level step order."""

import numpy as np
import pytest


# directions of clbm_moving.py:21-23, written out independently of the oracle
REF_E = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
         (1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1),
         (-1, 0, -1), (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1)]


def check_stream(oracle):
    """Brute force on a tiny grid: every value moves by exactly e_q (mod N)."""
    nx, ny, nz = 5, 4, 3
    vals = np.arange(nx * ny * nz * 19, dtype=np.float64).reshape(nx, ny, nz, 19)
    out = oracle.stream(vals)
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                for q, (a, b, c) in enumerate(REF_E):
                    assert out[i, j, k, q] == vals[(i - a) % nx, (j - b) % ny, (k - c) % nz, q]


def test_stream_is_the_shift_permutation(oracle):
    check_stream(oracle)


def test_planted_stream_index_mistake_is_caught(oracle, monkeypatch):
    bad = oracle.EY.copy()
    bad[[15, 16]] = bad[[16, 15]]
    monkeypatch.setattr(oracle, "EY", bad)
    with pytest.raises(AssertionError):
        check_stream(oracle)


def _inlet_state(oracle, seed):
    rng = np.random.default_rng(seed)
    f, *_ = oracle.init_feq(4, 3, 3, 0.05, 0.0)
    return f * (1 + 0.05 * rng.standard_normal(f.shape))


@pytest.mark.parametrize("u_in,v_in", [(0.05, 0.0), (0.0495, -0.00696), (0.1, 0.02)])
def test_inlet_rule_moment_identities(oracle, u_in, v_in):
    """The reference's inlet (clbm_moving.py:266-294) imposes, in closed form:
    rho_cell = rho_in (1 + u_in/3), x momentum (4/3) rho_in u_in, y momentum
    rho_in v_in + (f15 - f16 + f17 - f18), z momentum f15 + f16 - f17 - f18,
    with rho_in = (S_known + 2 S_out)/(1 - u_in); the known populations are untouched."""
    f = _inlet_state(oracle, 1)
    g0 = f[0].copy()
    known = [0, 2, 3, 4, 5, 6, 8, 10, 12, 14, 15, 16, 17, 18]
    s_known = g0[..., [0, 3, 4, 5, 6, 15, 16, 17, 18]].sum(-1)
    s_out = g0[..., [2, 8, 10, 12, 14]].sum(-1)
    rho_in = (s_known + 2 * s_out) / (1 - u_in)
    oracle.bc_inlet(f, u_in, v_in)
    g = f[0]
    np.testing.assert_array_equal(g[..., known], g0[..., known])
    e = np.stack([oracle.EX, oracle.EY, oracle.EZ], 1).astype(float)
    m = g @ e
    np.testing.assert_allclose(g.sum(-1), rho_in * (1 + u_in / 3), atol=1e-15)
    np.testing.assert_allclose(m[..., 0], 4 / 3 * rho_in * u_in, atol=1e-15)
    np.testing.assert_allclose(m[..., 1], rho_in * v_in + g[..., 15] - g[..., 16] + g[..., 17] - g[..., 18],
                               atol=1e-15)
    np.testing.assert_allclose(m[..., 2], g[..., 15] + g[..., 16] - g[..., 17] - g[..., 18], atol=1e-15)
    # other x planes untouched
    assert (f[1:] == _inlet_state(oracle, 1)[1:]).all()


def _hh_check(oracle, f0, u_in, v_in):
    """What the Hecht-Harting rule must impose at x = 0: the cell's density is
    rho_in and its momentum exactly rho_in (u_in, v_in, 0); f1 - f1^eq =
    f2 - f2^eq (bounce-back of the normal non-equilibrium part); known
    populations untouched."""
    f = f0.copy()
    g0 = f0[0]
    known = [0, 2, 3, 4, 5, 6, 8, 10, 12, 14, 15, 16, 17, 18]
    s_known = g0[..., [0, 3, 4, 5, 6, 15, 16, 17, 18]].sum(-1)
    s_out = g0[..., [2, 8, 10, 12, 14]].sum(-1)
    rho_in = (s_known + 2 * s_out) / (1 - u_in)
    oracle.bc_inlet_hecht_harting(f, u_in, v_in)
    g = f[0]
    np.testing.assert_array_equal(g[..., known], g0[..., known])
    e = np.stack([oracle.EX, oracle.EY, oracle.EZ], 1).astype(float)
    m = g @ e
    np.testing.assert_allclose(g.sum(-1), rho_in, atol=1e-15)
    np.testing.assert_allclose(m[..., 0], rho_in * u_in, atol=1e-15)
    np.testing.assert_allclose(m[..., 1], rho_in * v_in, atol=1e-15)
    np.testing.assert_allclose(m[..., 2], 0.0, atol=1e-15)
    # equilibrium difference of the normal pair: w (6 e.u) rho with w = 1/18
    np.testing.assert_allclose(g[..., 1] - g[..., 2], rho_in * u_in / 3, atol=1e-15)
    assert (f[1:] == f0[1:]).all()
    return f


@pytest.mark.parametrize("u_in,v_in", [(0.05, 0.0), (0.0495, -0.00696), (0.1, 0.02)])
def test_hecht_harting_inlet_imposes_the_velocity(oracle, u_in, v_in):
    f = _hh_check(oracle, _inlet_state(oracle, 4), u_in, v_in)
    g = f.copy()
    oracle.bc_inlet_hecht_harting(g, u_in, v_in)             # idempotent
    np.testing.assert_allclose(g, f, atol=1e-16)
    # at rest with the equilibrium of the inlet velocity, the rule changes nothing
    eq, *_ = oracle.init_feq(3, 4, 3, u_in, v_in)
    h = eq.copy()
    oracle.bc_inlet_hecht_harting(h, u_in, v_in)
    np.testing.assert_allclose(h, eq, atol=1e-16)


def test_planted_hecht_harting_mistake_is_caught(oracle, monkeypatch):
    """Dropping the f15..f18 terms of the transverse correction (what the
    reference's own rule does) breaks the imposed y momentum."""
    orig = oracle.bc_inlet_hecht_harting

    def planted(f, u_in, v_in):
        g = f[0]
        save = g[..., 15:19].copy()
        g[..., 15:19] = 0.0
        orig(f, u_in, v_in)
        g[..., 15:19] = save
        return f

    monkeypatch.setattr(oracle, "bc_inlet_hecht_harting", planted)
    with pytest.raises(AssertionError):
        _hh_check(oracle, _inlet_state(oracle, 5), 0.05, 0.01)


def test_inlet_rule_is_idempotent(oracle):
    f = _inlet_state(oracle, 2)
    oracle.bc_inlet(f, 0.05, 0.01)
    g = f.copy()
    oracle.bc_inlet(g, 0.05, 0.01)
    np.testing.assert_allclose(g, f, atol=1e-16)


def test_outlet_copies_last_interior_plane(oracle):
    f = _inlet_state(oracle, 3)
    g = oracle.bc_outlet(f.copy())
    assert (g[-1] == f[-2]).all() and (g[:-1] == f[:-1]).all()


def test_step_order_and_periodic_mass(oracle):
    """lbm_step = K2..K7 in order; with no inlet a periodic box conserves mass."""
    f = _inlet_state(oracle, 4)
    Z = np.zeros(f.shape[:3])
    rho, u, v, w = oracle.macro_with_force(f, Z, Z, Z)
    g = oracle.stream(oracle.collide(f, rho, u, v, w, Z, Z, Z, 1 / 0.56))
    assert abs(g.sum() - f.sum()) < 1e-12
    ref = oracle.bc_outlet(oracle.bc_inlet(g.copy(), 0.05, 0.0))
    h, rho2, *_ = oracle.lbm_step(f, Z, Z, Z, 1 / 0.56, 1.0, 0.05, 0.0)
    np.testing.assert_array_equal(h, ref)
    np.testing.assert_allclose(rho2, ref.sum(-1), atol=1e-15)
