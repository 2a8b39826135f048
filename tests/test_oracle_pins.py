"""Oracle pinning: K7 (macro_post), K13 (clip_forces), the rate and inflow
constants, the float32 trig convention and the order of the coupled loop.

Each check is a closed form, a brute-force definition or a hand composition
of oracle pieces, and each non-trivial one is shown to catch a planted
mistake.
"""

import math

import numpy as np
import pytest

# directions and weights of clbm_moving.py:21-30, written out independently of the oracle
REF_E = np.array([(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
                  (1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1),
                  (1, 0, -1), (-1, 0, -1), (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1)],
                 dtype=np.float64)
REF_W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def _equilibrium(rho, u):
    """Second-order equilibrium w_q rho (1 + 3 e.u + 9/2 (e.u)^2 - 3/2 u^2): its
    zeroth and first moments are rho and rho u exactly (D3Q19 weight identities)."""
    eu = u @ REF_E.T
    uu = (u * u).sum(-1, keepdims=True)
    return REF_W * rho[..., None] * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * uu)


# ---------------------------------------------------------------------------
# K7 macro_post (clbm_moving.py:305-321)
# ---------------------------------------------------------------------------
def check_macro_post(oracle):
    rng = np.random.default_rng(7)
    shape = (5, 4, 3)
    rho = 1 + 0.1 * rng.standard_normal(shape)
    u = 0.08 * rng.standard_normal(shape + (3,))
    f = _equilibrium(rho, u)
    r, ux, uy, uz = oracle.macro_post(f)
    np.testing.assert_allclose(r, rho, rtol=2e-15, atol=0)
    for a, comp in enumerate((ux, uy, uz)):
        np.testing.assert_allclose(comp, u[..., a], rtol=0, atol=1e-15)
    # away from equilibrium: the definition u = sum_q e_q f_q / sum_q f_q, cell by cell
    g = rng.uniform(0.01, 0.2, shape + (19,))
    r, ux, uy, uz = oracle.macro_post(g)
    for idx in np.ndindex(*shape):
        fq = g[idx]
        m = math.fsum(fq)
        assert abs(r[idx] - m) <= 1e-15 * m
        for a, comp in enumerate((ux, uy, uz)):
            want = math.fsum(REF_E[q, a] * fq[q] for q in range(19)) / m
            assert abs(comp[idx] - want) <= 1e-15


def test_macro_post_recovers_rho_and_u(oracle):
    check_macro_post(oracle)


@pytest.mark.parametrize("plant", ["swap_xy", "drop_q"])
def test_planted_macro_post_mistakes_are_caught(oracle, monkeypatch, plant):
    if plant == "swap_xy":
        ex, ey = oracle.EX.copy(), oracle.EY.copy()
        monkeypatch.setattr(oracle, "EX", ey)
        monkeypatch.setattr(oracle, "EY", ex)
    else:
        orig = oracle._moments_0_1

        def dropped(f):
            g = f.copy()
            g[..., 18] = 0.0
            return orig(g)

        monkeypatch.setattr(oracle, "_moments_0_1", dropped)
    with pytest.raises(AssertionError):
        check_macro_post(oracle)


# ---------------------------------------------------------------------------
# K13 clip_forces (airfoil_moving.py:219-225)
# ---------------------------------------------------------------------------
CLIP_TABLE = [
    # input (float64 accumulator), expected float32 output at f_max = 0.05
    (-3.0, np.float32(-0.05)),
    (-0.2, np.float32(-0.05)),
    (-0.05, np.float32(-0.05)),
    (-0.0499, np.float32(-0.0499)),
    (-1e-30, np.float32(-1e-30)),
    (0.0, np.float32(0.0)),
    (2.5e-3, np.float32(2.5e-3)),
    (0.05, np.float32(0.05)),
    (0.0500000001, np.float32(0.05)),
    (7.0, np.float32(0.05)),
]


def check_clip(oracle):
    vin = np.array([a for a, _ in CLIP_TABLE]).reshape(2, 5, 1)
    want = np.array([b for _, b in CLIP_TABLE], dtype=np.float32).reshape(2, 5, 1)
    fx, fy, fz = oracle.clip_forces(vin, -vin[::-1], 0.05)
    assert fx.dtype == np.float32 and fy.dtype == np.float32 and fz.dtype == np.float32
    np.testing.assert_array_equal(fx, want)
    np.testing.assert_array_equal(fy, -want[::-1])
    assert fz.shape == fx.shape and not np.any(fz)
    # a larger clamp leaves in-range values alone
    gx, _, _ = oracle.clip_forces(vin, vin, 10.0)
    np.testing.assert_array_equal(gx, np.clip(vin, -10, 10).astype(np.float32))


def test_clip_forces_table(oracle):
    check_clip(oracle)


def test_planted_one_sided_clip_is_caught(oracle, monkeypatch):
    def one_sided(fx, fy, f_max=0.05, dtype=np.float32):
        a = np.minimum(fx, f_max).astype(dtype)
        b = np.minimum(fy, f_max).astype(dtype)
        return a, b, np.zeros_like(a)

    monkeypatch.setattr(oracle, "clip_forces", one_sided)
    with pytest.raises(AssertionError):
        check_clip(oracle)


# ---------------------------------------------------------------------------
# relaxation_rates / inlet_velocity (config_moving.py:37-43, clbm_moving.py:33-35)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("U,chord,Re,nu,tau", [
    (0.05, 40.0, 200.0, 0.01, 0.53),      # 300 x 100 x 24 case
    (0.05, 80.0, 200.0, 0.02, 0.56),      # 600 x 200 x 50 case and the reference airplane
    (0.05, 160.0, 200.0, 0.04, 0.62),     # 1200 x 400 x 100 case
    (0.05, 120.0, 1000.0, 0.006, 0.518),  # config_highRe.py:43-44
    (0.05, 80.0, 5000.0, 0.0008, 0.5024),  # the Re 5000 coarse grid
])
def test_relaxation_rates_closed_form(oracle, U, chord, Re, nu, tau):
    r = oracle.relaxation_rates(U, chord, Re)
    assert abs(r["nu"] - nu) <= 1e-17
    assert abs(r["tau"] - tau) <= 2e-16
    assert abs(r["s_v"] - 1.0 / tau) <= 4e-16
    assert abs(r["s_v"] * r["tau"] - 1.0) <= 2e-16
    # s_e, s_q, s_pi are full relaxation (config_moving.py:41-43; hard-coded at
    # clbm_moving.py:188-196)
    assert (r["s_e"], r["s_q"], r["s_pi"]) == (1.0, 1.0, 1.0)


def test_relaxation_rate_known_digits(oracle):
    assert abs(oracle.relaxation_rates(0.05, 40, 200)["s_v"] - 1.8867924528301887) < 1e-15
    assert abs(oracle.relaxation_rates(0.05, 80, 200)["s_v"] - 1.7857142857142856) < 1e-15


def test_inlet_velocity_tilt(oracle):
    # cos 8 deg and sin 8 deg to 17 digits (tables), the tilt pointing below +x
    ux, uy = oracle.inlet_velocity(0.05, 8.0)
    assert abs(ux - 0.05 * 0.99026806874157031) < 1e-17
    assert abs(uy + 0.05 * 0.13917310096006544) < 1e-17
    assert uy < 0 < ux
    assert abs(math.hypot(ux, uy) - 0.05) < 1e-17
    assert abs(math.degrees(math.atan2(-uy, ux)) - 8.0) < 1e-12
    assert oracle.inlet_velocity(0.05, 0.0) == (0.05, 0.0)


# ---------------------------------------------------------------------------
# trig_f32: cos/sin of the float32 angle, FP64 then one rounding
# ---------------------------------------------------------------------------
def _nearest_f32(x64):
    """True when no float32 neighbour of float32(x64) is closer to x64."""
    c = np.float32(x64)
    lo = np.nextafter(c, np.float32(-np.inf))
    hi = np.nextafter(c, np.float32(np.inf))
    d = abs(float(c) - x64)
    return d <= abs(float(lo) - x64) and d <= abs(float(hi) - x64)


def check_trig(oracle):
    rng = np.random.default_rng(3)
    th = np.concatenate([rng.uniform(-math.pi, math.pi, 20000).astype(np.float32),
                         np.float32([0.0, 1e-8, -1e-8, math.pi / 2, -math.pi / 2, math.pi,
                                     math.radians(8.0), -math.radians(18.0), 0.5, 1.0])])
    for t in th:
        c, s = oracle.trig_f32(t)
        assert isinstance(c, np.float32) and isinstance(s, np.float32)
        t64 = float(np.float32(t))
        for got, exact in ((c, math.cos(t64)), (s, math.sin(t64))):
            # |got - exact| within half an ulp of got: the correctly rounded value
            # (the FP64 reference has 1e-16 relative error, far below a float32 ulp)
            half_ulp = 0.5 * float(np.spacing(np.abs(got)))
            assert abs(float(got) - exact) <= half_ulp * (1 + 1e-6) + 1e-300
            assert _nearest_f32(exact) or abs(abs(float(got) - exact) - half_ulp) < 1e-7 * half_ulp
    # the angle is rounded to float32 first: a float64 angle gives its float32 angle's values
    for t in (0.1, 1.2345678912345, -2.9999999):
        assert oracle.trig_f32(t) == oracle.trig_f32(np.float32(t))


def test_trig_f32_rounds_fp64_once(oracle):
    check_trig(oracle)


@pytest.mark.parametrize("plant", ["truncate", "unrounded_angle"])
def test_planted_trig_mistakes_are_caught(oracle, monkeypatch, plant):
    if plant == "truncate":
        def trunc(theta):
            t = np.float64(np.float32(theta))
            out = []
            for v in (math.cos(t), math.sin(t)):
                c = np.float32(v)
                if abs(float(c)) > abs(v):   # round toward zero
                    c = np.nextafter(c, np.float32(0))
                out.append(np.float32(c))
            return tuple(out)
        monkeypatch.setattr(oracle, "trig_f32", trunc)
    else:
        monkeypatch.setattr(oracle, "trig_f32",
                            lambda theta: (np.float32(math.cos(theta)), np.float32(math.sin(theta))))
    with pytest.raises(AssertionError):
        check_trig(oracle)


# ---------------------------------------------------------------------------
# run_coupled: the loop order of main_moving.py:116-122
# ---------------------------------------------------------------------------
def _small_case(pkg):
    return pkg.config.naca_case(28, 20, 3, 8.0, le_x=7.0, N_steps=2, name="pins")


def _by_hand(oracle, cfg, f0, read="k7"):
    """Two coupled steps composed from the oracle's pieces.  read="k7" (the
    reference): step n's compute_force reads K7 of the state before step n
    (for step 1, of K1's state).  Two plausible mistakes for comparison:
    read="stale" keeps reading the velocity of the initial state, read="k2"
    reads the half-force velocity K2 of the previous step."""
    kw = dict(s_v=cfg.s_v, s_e=cfg.s_e, ux_in=cfg.ux_in, uy_in=cfg.uy_in)
    f = f0.copy()
    _, ux, uy, _ = oracle.macro_post(f)
    hist = []
    for n in (1, 2):
        out = oracle.compute_force(ux, uy, cfg.X0, cfg.Y0, cfg.Z0, *cfg.motion_state(n),
                                   weight_dtype=np.float64)
        fx, fy = out["fx"].astype(np.float64), out["fy"].astype(np.float64)
        fz = np.zeros_like(fx)
        if read == "k2":
            _, ux, uy, _ = oracle.macro_with_force(f, fx, fy, fz)
        f = oracle.lbm_step(f, fx, fy, fz, **kw)[0]
        if read == "k7":
            _, ux, uy, _ = oracle.macro_post(f)
        hist.append(oracle.force_sum_f64(out["FX"], out["FY"], cfg.dV))
    return f, np.array(hist)


def test_run_coupled_is_the_reference_loop_order(oracle, pkg):
    cfg = _small_case(pkg)
    f0, rho0, ux0, uy0, _ = oracle.init_feq(cfg.Nx, cfg.Ny, cfg.Nz, cfg.ux_in, cfg.uy_in)
    # step 1 reads K1's velocity: K7 of the equilibrium state is the inflow itself
    _, u7, v7, w7 = oracle.macro_post(f0)
    np.testing.assert_allclose(u7, ux0, atol=1e-17)
    np.testing.assert_allclose(v7, uy0, atol=1e-17)
    assert np.abs(w7).max() < 1e-17
    f_rc, h_rc = oracle.run_coupled(f0.copy(), cfg.X0, cfg.Y0, cfg.Z0, 2, cfg.motion_state,
                                    cfg.s_v, cfg.s_e, cfg.ux_in, cfg.uy_in, dV=cfg.dV,
                                    dtype=np.float64)
    f_h, h_h = _by_hand(oracle, cfg, f0)
    np.testing.assert_array_equal(f_rc, f_h)
    np.testing.assert_array_equal(h_rc, h_h)
    # reading a stale or the half-force velocity at step 2 gives other forces and
    # populations, so the comparison above would catch either
    for wrong in ("stale", "k2"):
        f_w, h_w = _by_hand(oracle, cfg, f0, read=wrong)
        np.testing.assert_array_equal(h_w[0], h_h[0])
        assert np.abs(h_w[1] - h_h[1]).max() > 1e-4
        assert np.abs(f_w - f_h).max() > 1e-7
    # so does starting the motion one step late (first_step = 2)
    f_late, _ = oracle.run_coupled(f0.copy(), cfg.X0, cfg.Y0, cfg.Z0, 2, cfg.motion_state,
                                   cfg.s_v, cfg.s_e, cfg.ux_in, cfg.uy_in, dV=cfg.dV,
                                   dtype=np.float64, first_step=2)
    assert np.abs(f_late - f_h).max() > 1e-9
