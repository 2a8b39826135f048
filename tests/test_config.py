"""Host-side case logic: marker generators, the config-module convention and kinematics."""

import math
import sys

import numpy as np
import pytest

from conftest import load_sub


@pytest.fixture(scope="module")
def cfgmod():
    return load_sub("config")


@pytest.mark.parametrize("name,count,per_section", [
    ("reference_airplane", 74201, 360),   # config_moving.py:196-307
    ("reference_highre", 166765, 540),    # config_highRe.py
    ("reference_fish", 7700, 308),        # fish/config_fish.py:94-108
    ("case_c0", 1968, 82),
    ("case_c1", 8200, 164),
])
def test_marker_counts(cfgmod, name, count, per_section):
    c = getattr(cfgmod, name)()
    assert c.N_MARKERS == count
    assert c.extra["N_MARKERS_2D"] == per_section
    assert c.X0.dtype == np.float32 and len(c.Y0) == len(c.Z0) == count


def test_reference_rates(cfgmod):
    a = cfgmod.reference_airplane()
    assert abs(a.tau - 0.56) < 1e-12 and abs(a.s_v - 1 / 0.56) < 1e-12
    assert abs(cfgmod.reference_highre().tau - 0.518) < 1e-12
    assert abs(cfgmod.case_c0().tau - 0.53) < 1e-12
    # tilted inflow of clbm_moving.py:33-35
    assert abs(a.ux_in - 0.05 * math.cos(math.radians(8))) < 1e-15
    assert abs(a.uy_in + 0.05 * math.sin(math.radians(8))) < 1e-15


def test_arclength_section(cfgmod):
    X, Y, ds = cfgmod.naca0012_arclength(80.0, 1.0, 0.25)
    assert len(X) == 164 and 0.99 < ds < 1.0
    pts = np.stack([X, Y], 1).astype(float)
    seg = np.hypot(*(np.roll(pts, -1, 0) - pts).T)
    # uniform arc spacing: chords are a little shorter than the arc at the
    # leading edge (curvature) and across the 0.2-lu trailing-edge gap corner
    assert seg[:-1].max() < 1.0 and seg[:-1].min() > 0.97 and seg[-1] > 0.8
    assert np.median(seg) > 0.995
    # pivot at c/4: the trailing edge is marker 0, the leading edge lies between two markers
    assert abs(X.min() + 20.0) < 1e-2 and abs(X.max() - 60.0) < 1e-5
    assert abs(np.abs(Y).max() - 0.06 * 80 * 1.0) < 0.05   # 12 % thickness


def test_cosine_section_matches_reference_rule(cfgmod):
    X, Y = cfgmod.naca0012_cosine(80, 0.7)
    n = len(X)
    assert n == 360                                          # ceil(pi*80/0.7) -> even
    assert X[0] == np.float32(-40.0) and abs(X[n // 2] - 40.0) < 1e-5
    assert (Y[1:n // 2] > 0).all() and (Y[n // 2 + 1:] < 0).all()


def test_module_round_trip_and_alias(cfgmod):
    c = cfgmod.case_c0()
    mod = cfgmod.as_module(c)
    back = cfgmod.from_module(mod)
    assert (back.Nx, back.Ny, back.Nz) == (c.Nx, c.Ny, c.Nz)
    assert np.array_equal(back.X0, c.X0) and abs(back.s_v - c.s_v) < 1e-15
    old = sys.modules.get("config_moving")
    try:
        sys.modules["config_moving"] = mod          # main_highRe.py:26-27 convention
        r = cfgmod.resolve()
        assert r.N_MARKERS == c.N_MARKERS and r.Nx == c.Nx
    finally:
        if old is None:
            sys.modules.pop("config_moving", None)
        else:
            sys.modules["config_moving"] = old


def test_motion_state_matches_closed_form(cfgmod):
    c = cfgmod.case_c1()
    w = 2 * 0.5 * 0.05 / 80
    assert abs(c.w_motion - w) < 1e-15
    xc, yc, th, vx, vy, om = c.motion_state(1000)
    assert xc == 170.0 and vx == 0.0
    assert abs(yc - (100 + 20 * math.sin(w * 1000))) < 1e-12
    alpha = math.radians(8) + math.radians(10) * math.sin(w * 1000 + math.pi / 2)
    assert abs(th + alpha) < 1e-12                      # theta = -alpha (nose up for +x inflow)
    assert abs(om + math.radians(10) * w * math.cos(w * 1000 + math.pi / 2)) < 1e-15


def test_naca_case_box_fits(cfgmod):
    """Markers of the benchmark case stay inside the lattice through a period."""
    c = cfgmod.case_c1()
    r = np.hypot(c.X0.astype(float), c.Y0.astype(float)).max()
    for step in range(0, 10000, 500):
        xc, yc, *_ = c.motion_state(step)
        assert 2 < xc - r and xc + r < c.Nx - 3 and 2 < yc - r and yc + r < c.Ny - 3
