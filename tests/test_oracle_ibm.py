"""Oracle pinning: Peskin kernel, marker kinematics, interpolation and spreading."""

import math

import numpy as np
import pytest


def test_peskin_identities(oracle):
    """Peskin (2002, Acta Numerica s.6): for every r, sum_j phi(r-j) = 1,
    sum_j (r-j) phi(r-j) = 0, sum_j phi(r-j)^2 = 3/8; phi(0) = 1/2,
    phi(+-1) = 1/4, phi(+-2) = 0; support |r| < 2."""
    r = np.random.default_rng(0).uniform(0, 1, 1000)
    j = np.arange(-3, 5)
    ph = oracle.phi1d(r[:, None] - j[None, :])
    np.testing.assert_allclose(ph.sum(1), 1.0, atol=1e-14)
    np.testing.assert_allclose(((r[:, None] - j) * ph).sum(1), 0.0, atol=1e-14)
    np.testing.assert_allclose((ph ** 2).sum(1), 3 / 8, atol=1e-14)
    np.testing.assert_allclose(oracle.phi1d(np.array([0.0, 1.0, -1.0, 2.0, -2.0, 2.5])),
                               [0.5, 0.25, 0.25, 0.0, 0.0, 0.0], atol=1e-16)
    # continuous at the branch points
    for x in (1.0, 2.0):
        a, b = oracle.phi1d(np.array([x - 1e-9, x + 1e-9]))
        assert abs(a - b) < 1e-6


def test_update_markers_rigid_motion(oracle):
    rng = np.random.default_rng(1)
    X0, Y0 = rng.uniform(-20, 60, 50).astype(np.float32), rng.uniform(-5, 5, 50).astype(np.float32)
    Z0 = rng.uniform(0, 10, 50).astype(np.float32)
    xc, yc, th, vx, vy, om = 170.0, 100.0, -0.3, 0.001, 0.004, -2e-4
    X, Y, Z, VX, VY = oracle.update_markers(X0, Y0, Z0, xc, yc, th, vx, vy, om)
    # distances to the pivot are preserved, the rotation is counter-clockwise by theta
    d0 = np.hypot(X0.astype(float), Y0.astype(float))
    d1 = np.hypot(X - xc, Y - yc)
    np.testing.assert_allclose(d1, d0, atol=2e-4)
    ang = np.arctan2(Y - yc, X - xc) - np.arctan2(Y0, X0)
    np.testing.assert_allclose(np.angle(np.exp(1j * ang)), th, atol=2e-5)
    assert (Z == Z0).all()
    np.testing.assert_allclose(VX, vx - om * (Y - yc), rtol=1e-6)
    np.testing.assert_allclose(VY, vy + om * (X - xc), rtol=1e-6)
    assert X.dtype == np.float32


def _brute_stencil(x, n, clamp, quirk):
    """Independent enumeration: the 4 cells floor(x)-1 .. floor(x)+2 with the
    reference's index rule and the phi argument the rule uses."""
    b = math.floor(x) - 1
    cells, args = [], []
    for d in range(4):
        iu = b + d
        ii = min(max(iu, 0), n - 1) if clamp else iu % n
        cells.append(ii)
        args.append((ii if quirk else iu) - x)
    return cells, args


@pytest.mark.parametrize("quirk", [True, False])
def test_stencil_against_brute_force(oracle, quirk):
    nx, ny, nz = 9, 7, 5
    pts = np.array([[0.3, 0.2, 0.0], [8.7, 6.9, 4.99], [4.0, 3.0, 2.0], [-0.5, 3.5, 1.5],
                    [4.5, 0.01, 4.2]], dtype=np.float64)
    base, ii, jj, kk, phx, phy, phz = oracle.stencil(pts[:, 0], pts[:, 1], pts[:, 2], nx, ny, nz, quirk)
    for m, (x, y, z) in enumerate(pts):
        for arr, ph, v, n, cl in ((ii, phx, x, nx, True), (jj, phy, y, ny, False), (kk, phz, z, nz, False)):
            cells, args = _brute_stencil(v, n, cl, quirk)
            assert arr[m].tolist() == cells
            np.testing.assert_allclose(ph[m], oracle.phi1d(np.array(args)), atol=1e-16)
        assert base[m].tolist() == [math.floor(x) - 1, math.floor(y) - 1, math.floor(z) - 1]


def test_seam_layer_weight_quirk(oracle):
    """Markers on the seam layers z = 0 and z = Nz-1 keep only 3/4 of their z
    weight with the reference rule (the wrapped index enters phi: at z = 0
    the cell Nz-1 gets phi(Nz-1) = 0, at z = Nz-1 the cell 0 gets phi(1-Nz)
    = 0); without the quirk the weights sum to 1."""
    nz = 8
    z = np.array([0.0, nz - 1.0, 3.0])
    x = np.array([4.5, 4.5, 4.5])
    *_, phz = oracle.stencil(x, x, z, 10, 10, nz, True)
    np.testing.assert_allclose(phz.sum(1), [0.75, 0.75, 1.0], atol=1e-15)
    *_, phz = oracle.stencil(x, x, z, 10, 10, nz, False)
    np.testing.assert_allclose(phz.sum(1), [1.0, 1.0, 1.0], atol=1e-15)


def test_interpolation_reproduces_constant_and_linear_fields(oracle):
    nx, ny, nz = 16, 14, 10
    rng = np.random.default_rng(2)
    X = rng.uniform(4, 11, 30)
    Y = rng.uniform(4, 9, 30)
    Z = rng.uniform(3, 6, 30)
    c = np.full((nx, ny, nz), 0.037)
    U, V = oracle.interpolate(c, c, X, Y, Z)
    np.testing.assert_allclose(U, 0.037, atol=1e-15)
    lin = np.broadcast_to(np.arange(nx, dtype=float)[:, None, None], (nx, ny, nz))
    U, _ = oracle.interpolate(lin, c, X, Y, Z)
    np.testing.assert_allclose(U, X, atol=1e-12)


def test_interpolation_and_spreading_are_adjoint(oracle):
    """<S F, u> = <F, I u> with the same weights, incl. seam layers and x clamp."""
    nx, ny, nz = 12, 10, 6
    rng = np.random.default_rng(3)
    n = 40
    X = np.concatenate([rng.uniform(-1, 13, n - 4), [0.2, 11.8, 5.0, 5.0]])
    Y = rng.uniform(-1, 11, n)
    Z = np.concatenate([rng.uniform(0, 6, n - 2), [0.0, 5.0]])
    u = rng.standard_normal((nx, ny, nz))
    F = rng.standard_normal(n)
    for quirk in (True, False):
        Iu, _ = oracle.interpolate(u, u, X, Y, Z, quirk)
        SF, _ = oracle.spread(X, Y, Z, F, F, (nx, ny, nz), quirk)
        assert abs((SF * u).sum() - (F * Iu).sum()) < 1e-12


def test_marker_force_clamp(oracle):
    U = np.array([0.0, 0.2, -0.2, 0.03])
    V = np.zeros(4)
    FX, FY = oracle.marker_forces(U, V, np.zeros(4), np.zeros(4), 0.05)
    np.testing.assert_allclose(FX, [0.0, -0.05, 0.05, -0.03])


def test_motion_state_rates_are_derivatives(oracle):
    args = dict(x0=170.0, y0=100.0, A_heave=20.0, A_pitch=-math.radians(10), phi_pitch=math.pi / 2,
                AoA0_rad=-math.radians(8), w_motion=6.25e-4)
    for step in (1, 777, 5000):
        xc, yc, th, vx, vy, om = oracle.motion_state(step, **args)
        h = 1e-2
        a = oracle.motion_state(step + h, **args)
        b = oracle.motion_state(step - h, **args)
        assert abs((a[1] - b[1]) / (2 * h) - vy) < 1e-9
        assert abs((a[2] - b[2]) / (2 * h) - om) < 1e-11
        assert xc == 170.0 and vx == 0.0


def test_force_sum_is_exact(oracle):
    FX = np.array([1e-3, 2e16, -2e16, 3e-3], dtype=np.float64)
    s = oracle.force_sum_f64(FX, FX)
    assert s == (4e-3, 4e-3)
