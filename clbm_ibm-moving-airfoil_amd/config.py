"""Case configuration: lattice, relaxation rates, marker geometry and motion.

The reference binds its parameters at import time from a module named
``config_moving`` (clbm_moving.py:15-16, airfoil_moving.py:17-19) and swaps
cases by aliasing that module name (main_highRe.py:26-27).  Here a case is a
``CaseConfig`` value.  ``from_module`` reads any module that uses the
reference's names, so the reference's own config files (or an alias placed
in ``sys.modules['config_moving']``) keep working, and ``as_module`` goes the
other way.  The geometry generators restate the reference's marker builders
host-side (config_moving.py:105-307, fish/config_fish.py:74-108) plus the
arc-length NACA 0012 section with a quarter-chord pivot used by the
benchmark cases.
"""

from __future__ import annotations

import math
import sys
import types
from dataclasses import dataclass, field, replace
from typing import Optional

import numpy as np

CS2 = 1.0 / 3.0


# ---------------------------------------------------------------------------
# geometry generators (all return float32 arrays in the body frame)
# ---------------------------------------------------------------------------
def naca0012_thickness(xu):
    """Half-thickness of NACA 0012 per unit chord (config_moving.py:116-121)."""
    t = 0.12
    return 5.0 * t * (0.2969 * np.sqrt(xu) - 0.1260 * xu - 0.3516 * xu ** 2
                      + 0.2843 * xu ** 3 - 0.1015 * xu ** 4)


def naca0012_cosine(chord: float, ds: float):
    """Surface section as config_moving.py:105-125 builds it: cosine-spaced,
    N = ceil(pi c / ds) rounded up to even (>= 64), chord-centred frame."""
    n = max(int(np.ceil(np.pi * chord / ds)), 64)
    if n % 2:
        n += 1
    th = np.linspace(0.0, 2.0 * np.pi, n + 1)[:-1]
    xu = 0.5 * (1.0 - np.cos(th))
    yt = naca0012_thickness(xu)
    sign = np.where(th <= np.pi, 1.0, -1.0)
    return ((xu - 0.5) * chord).astype(np.float32), (sign * yt * chord).astype(np.float32)


def naca0012_interior(chord: float, ds: float):
    """Interior fill of the section on a ds grid (config_moving.py:128-152)."""
    xs = np.arange(-chord / 2.0 + ds * 0.5, chord / 2.0, ds)
    ax, ay = [], []
    for xi in xs:
        xu = max(float(xi) / chord + 0.5, 1e-6)
        yt = chord * float(naca0012_thickness(xu))
        if yt < ds * 0.5:
            continue
        for yi in np.arange(-yt + ds * 0.5, yt, ds):
            ax.append(xi)
            ay.append(yi)
    return np.array(ax, dtype=np.float32), np.array(ay, dtype=np.float32)


def ellipsoid_surface(ax: float, ay: float, az: float, ds: float):
    """Ring markers on an axis-aligned ellipsoid (config_moving.py:155-190):
    rings by polar angle, Ramanujan perimeter per ring."""
    nb = max(int(np.ceil(np.pi * az / ds)), 8)
    X, Y, Z = [], [], []
    for i in range(nb + 1):
        ph = np.pi * i / nb
        zl = az * np.cos(ph)
        fr = np.sin(ph)
        if fr < 1e-8:
            X.append(0.0); Y.append(0.0); Z.append(float(zl))
            continue
        a, b = ax * fr, ay * fr
        h = ((a - b) / (a + b)) ** 2
        circ = np.pi * (a + b) * (1.0 + 3.0 * h / (10.0 + np.sqrt(4.0 - 3.0 * h)))
        nth = max(int(np.ceil(circ / ds)), 4)
        for t in np.linspace(0.0, 2.0 * np.pi, nth, endpoint=False):
            X.append(ax * fr * np.cos(t)); Y.append(ay * fr * np.sin(t)); Z.append(float(zl))
    return np.array(X, dtype=np.float32), np.array(Y, dtype=np.float32), np.array(Z, dtype=np.float32)


def ellipse_section(chord: float, thickness: float, ds: float):
    """Fish cross-section (fish/config_fish.py:74-91): ellipse, N from the
    Ramanujan perimeter, even, >= 64."""
    a, b = chord / 2.0, thickness / 2.0
    per = np.pi * (3 * (a + b) - np.sqrt((3 * a + b) * (a + 3 * b)))
    n = max(int(np.ceil(per / ds)), 64)
    if n % 2:
        n += 1
    th = np.linspace(0.0, 2.0 * np.pi, n + 1)[:-1]
    return (a * np.cos(th)).astype(np.float32), (b * np.sin(th)).astype(np.float32)


def naca0012_arclength(chord: float, spacing: float = 1.0, pivot_frac: float = 0.25,
                       n: Optional[int] = None):
    """NACA 0012 section resampled at uniform arc length (closed trailing
    edge), with the body origin at ``pivot_frac`` of the chord from the
    leading edge.  N = ceil(perimeter / spacing) unless given.  Returns
    (X, Y, ds) with ds = perimeter / N, the arc length each marker carries."""
    th = np.linspace(0.0, 2.0 * np.pi, 200001)
    xu = 0.5 * (1.0 - np.cos(th))
    sign = np.where(th <= np.pi, 1.0, -1.0)
    yv = sign * naca0012_thickness(xu)
    # trace LE -> upper surface -> TE -> lower surface -> LE
    xs = np.concatenate([xu[th <= np.pi][::-1], xu[th > np.pi][::-1]])
    ys = np.concatenate([yv[th <= np.pi][::-1], yv[th > np.pi][::-1]])
    xs = np.append(xs, xs[0])
    ys = np.append(ys, ys[0])
    seg = np.hypot(np.diff(xs), np.diff(ys)) * chord
    s = np.concatenate([[0.0], np.cumsum(seg)])
    per = s[-1]
    if n is None:
        n = int(math.ceil(per / spacing - 1e-9))
    st = np.arange(n) * (per / n)
    X = (np.interp(st, s, xs) - pivot_frac) * chord
    Y = np.interp(st, s, ys) * chord
    return X.astype(np.float32), Y.astype(np.float32), per / n


def spatial_order(X0, Y0, Z0, cell: float = 1.0) -> np.ndarray:
    """Marker visiting order along a Morton (Z-order) curve of the body-frame
    positions quantised to ``cell``: consecutive markers are neighbours, so
    the 4x4x4 stencils of a group of them overlap.  Returns int32 indices."""
    pts = np.stack([np.asarray(a, dtype=np.float64) for a in (X0, Y0, Z0)], 1)
    if len(pts) == 0:
        return np.zeros(0, np.int32)
    q = np.floor((pts - pts.min(0)) / cell).astype(np.uint64)
    key = np.zeros(len(q), dtype=np.uint64)
    for bit in range(21):
        for a in range(3):
            key |= ((q[:, a] >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit + a)
    return np.argsort(key, kind="stable").astype(np.int32)


def convex_hull_2d(X, Y) -> np.ndarray:
    """Indices of the vertices of the convex hull of the points (X, Y)
    (Andrew's monotone chain, counter-clockwise, collinear points dropped).
    Under any rotation and translation, the extreme coordinates of the point
    set are attained at these vertices."""
    x = np.asarray(X, dtype=np.float64)
    y = np.asarray(Y, dtype=np.float64)
    n = len(x)
    if n <= 2:
        return np.arange(n, dtype=np.int64)
    idx = np.lexsort((y, x))

    def half(order):
        out = []
        for i in order:
            while len(out) >= 2:
                a, b = out[-2], out[-1]
                cross = (x[b] - x[a]) * (y[i] - y[a]) - (y[b] - y[a]) * (x[i] - x[a])
                if cross > 0:
                    break
                out.pop()
            out.append(i)
        return out

    lower = half(idx)
    upper = half(idx[::-1])
    return np.array(lower[:-1] + upper[:-1], dtype=np.int64)


# ---------------------------------------------------------------------------
# case configuration
# ---------------------------------------------------------------------------
@dataclass
class CaseConfig:
    name: str
    Nx: int
    Ny: int
    Nz: int
    U_inf: float
    Re: float
    chord: float
    s_v: float
    s_e: float = 1.0
    s_q: float = 1.0
    s_pi: float = 1.0
    cs2: float = CS2
    AoA0_deg: float = 0.0
    ux_in: float = 0.0        # inlet / initial velocity (clbm_moving.py:33-35,77-78)
    uy_in: float = 0.0
    # markers in the body frame (X0, Y0) and absolute z (Z0)
    X0: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32), repr=False)
    Y0: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32), repr=False)
    Z0: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32), repr=False)
    dV: Optional[np.ndarray] = field(default=None, repr=False)  # marker volume weights (None = 1)
    # motion: yc = y0 + A_heave sin(wt); theta = theta0 + pitch_sign*A_pitch sin(wt + phi)
    x0: float = 0.0
    y0: float = 0.0
    A_heave: float = 0.0
    A_pitch: float = 0.0
    phi_pitch: float = 0.0
    theta0: float = 0.0       # mean rotation angle of the body (rad)
    f_motion: float = 0.0
    N_steps: int = 1000
    vis_interval: int = 200
    DATA_DIR: str = "frames"
    f_max: float = 0.05
    wrap_weight_quirk: bool = True
    extra: dict = field(default_factory=dict, repr=False)

    @property
    def nu(self) -> float:
        return (1.0 / self.s_v - 0.5) * self.cs2

    @property
    def tau(self) -> float:
        return 1.0 / self.s_v

    @property
    def N_MARKERS(self) -> int:
        return int(len(self.X0))

    @property
    def w_motion(self) -> float:
        return 2.0 * math.pi * self.f_motion

    def motion_state(self, step: int):
        """Prescribed heave + pitch (main_moving.py:65-74): t = float(step)."""
        t = float(step)
        w = self.w_motion
        xc = self.x0
        yc = self.y0 + self.A_heave * math.sin(w * t)
        theta = self.theta0 + self.A_pitch * math.sin(w * t + self.phi_pitch)
        vx = 0.0
        vy = self.A_heave * w * math.cos(w * t)
        omega = self.A_pitch * w * math.cos(w * t + self.phi_pitch)
        return xc, yc, theta, vx, vy, omega

    def with_(self, **kw) -> "CaseConfig":
        return replace(self, **kw)


def _rates(U: float, chord: float, Re: float):
    nu = U * chord / Re
    tau = 0.5 + nu / CS2
    return 1.0 / tau


def airplane_case(name: str, Nx: int, Ny: int, Nz: int, c: float, b_half: float, x0: float,
                  fuse: tuple, arm: float, c_h: float, bh_h: float, c_v: float, h_v: float,
                  A_heave: float, Re: float, N_steps: int, DATA_DIR: str, U: float = 0.05,
                  aoa: float = 8.0, ds: float = 0.7) -> CaseConfig:
    """The reference's airplane (config_moving.py:196-307): wing sections on
    1-lu z layers, filled tip caps, ellipsoid fuselage, H-tail with tip caps,
    V-fin stacked in local y with its tip cap.  Inflow tilted by AoA0
    (clbm_moving.py:33-35), pitch 5 deg, k = 0.5, pivot (x0, Ny/2, Nz/2)."""
    y0, z0 = Ny / 2.0, Nz / 2.0
    Xt, Yt = naca0012_cosine(c, ds)
    zl = np.arange(z0 - b_half, z0 + b_half + 0.5, 1.0, dtype=np.float32)
    Xw, Yw, Zw = np.tile(Xt, len(zl)), np.tile(Yt, len(zl)), np.repeat(zl, len(Xt))
    Xi, Yi = naca0012_interior(c, ds)
    Xtip, Ytip = np.tile(Xi, 2), np.tile(Yi, 2)
    Ztip = np.concatenate([np.full(len(Xi), z0 - b_half, np.float32),
                           np.full(len(Xi), z0 + b_half, np.float32)])
    Xf, Yf, Zf = ellipsoid_surface(*fuse, ds)
    Zf = (z0 + Zf).astype(np.float32)
    Xh, Yh = naca0012_cosine(c_h, ds)
    zh = np.arange(z0 - bh_h, z0 + bh_h + 0.5, 1.0, dtype=np.float32)
    Xht, Yht, Zht = np.tile(Xh + arm, len(zh)), np.tile(Yh, len(zh)), np.repeat(zh, len(Xh))
    Xhi, Yhi = naca0012_interior(c_h, ds)
    Xhtip, Yhtip = np.tile(Xhi + arm, 2), np.tile(Yhi, 2)
    Zhtip = np.concatenate([np.full(len(Xhi), z0 - bh_h, np.float32),
                            np.full(len(Xhi), z0 + bh_h, np.float32)])
    Xv, Yv = naca0012_cosine(c_v, ds)
    yv = np.arange(fuse[1], fuse[1] + h_v + 0.5, 1.0, dtype=np.float32)
    Xvt, Yvt, Zvt = np.tile(Xv + arm, len(yv)), np.repeat(yv, len(Xv)), np.tile(Yv + z0, len(yv))
    Xvi, Yvi = naca0012_interior(c_v, ds)
    Xvtip = Xvi + arm
    Yvtip = np.full(len(Xvi), fuse[1] + h_v, np.float32)
    Zvtip = Yvi + z0
    X0 = np.concatenate([Xw, Xtip, Xf, Xht, Xhtip, Xvt, Xvtip]).astype(np.float32)
    Y0 = np.concatenate([Yw, Ytip, Yf, Yht, Yhtip, Yvt, Yvtip]).astype(np.float32)
    Z0 = np.concatenate([Zw, Ztip, Zf, Zht, Zhtip, Zvt, Zvtip]).astype(np.float32)
    ux_in = float(U * math.cos(math.radians(aoa)))
    uy_in = float(-U * math.sin(math.radians(aoa)))
    return CaseConfig(
        name=name, Nx=Nx, Ny=Ny, Nz=Nz, U_inf=U, Re=Re, chord=c,
        s_v=_rates(U, c, Re), AoA0_deg=aoa, ux_in=ux_in, uy_in=uy_in,
        X0=X0, Y0=Y0, Z0=Z0, dV=None, x0=x0, y0=y0, A_heave=A_heave, A_pitch=math.radians(5.0),
        phi_pitch=math.pi / 2.0, theta0=math.radians(aoa), f_motion=0.5 * U / (math.pi * c),
        N_steps=N_steps, vis_interval=200, DATA_DIR=DATA_DIR,
        extra=dict(X_template=Xt, Y_template=Yt, X_htail_tmpl=Xh, Y_htail_tmpl=Yh,
                   N_MARKERS_2D=len(Xt), b_half=b_half, z0=z0, fuse=fuse, x_tail_arm=arm,
                   c_htail=c_h, b_half_htail=bh_h, c_vtail=c_v, h_vtail=h_v,
                   z_wing_start=z0 - b_half, z_wing_end=z0 + b_half,
                   z_htail_start=z0 - bh_h, z_htail_end=z0 + bh_h))


def reference_airplane() -> CaseConfig:
    """config_moving.py: chord 80 at Re 200 on 600 x 200 x 160, heave 10 lu,
    20,000 steps (config_moving.py:28-98,312-332)."""
    return airplane_case("reference-airplane", 600, 200, 160, 80, 60, 160.0, (75.0, 10.0, 10.0),
                         80.0, 36.0, 32.0, 36.0, 28.0, A_heave=10.0, Re=200, N_steps=20000,
                         DATA_DIR="drone_frames")


def reference_highre() -> CaseConfig:
    """config_highRe.py: the airplane scaled x1.5 (chord 120) at Re 1000 on
    800 x 260 x 220, heave 15 lu (config_highRe.py:34-105,288-308)."""
    return airplane_case("reference-highRe", 800, 260, 220, 120, 90, 200.0, (112.5, 15.0, 15.0),
                         120.0, 54.0, 48.0, 54.0, 42.0, A_heave=15.0, Re=1000, N_steps=20000,
                         DATA_DIR="drone_frames_highRe")


def reference_fish() -> CaseConfig:
    """fish/config_fish.py: 4:1 ellipse, chord 100, Re 300, 800 x 250 x 25,
    one layer per z = 0..24, heave 0.25c, pitch 15 deg, k = 0.8, no AoA."""
    c, U, Re, Nx, Ny, Nz = 100, 0.05, 300, 800, 250, 25
    Xt, Yt = ellipse_section(c, 25, 0.7)
    zl = np.arange(0, Nz, 1.0, dtype=np.float32)
    return CaseConfig(
        name="reference-fish", Nx=Nx, Ny=Ny, Nz=Nz, U_inf=U, Re=Re, chord=c, s_v=_rates(U, c, Re),
        AoA0_deg=0.0, ux_in=U, uy_in=0.0, X0=np.tile(Xt, Nz), Y0=np.tile(Yt, Nz),
        Z0=np.repeat(zl, len(Xt)), x0=200.0, y0=Ny / 2.0, A_heave=25.0, A_pitch=math.radians(15.0),
        phi_pitch=math.pi / 2.0, theta0=0.0, f_motion=0.8 * U / (math.pi * c), N_steps=25000,
        vis_interval=150, DATA_DIR="fish_frames",
        extra=dict(X_template=Xt, Y_template=Yt, N_MARKERS_2D=len(Xt)))


def naca_case(Nx: int, Ny: int, Nz: int, chord: float, Re: float = 200.0, U: float = 0.05,
              le_x: Optional[float] = None, k: float = 0.5, heave_frac: float = 0.25,
              aoa_mean_deg: float = 8.0, pitch_amp_deg: float = 10.0, phase_deg: float = 90.0,
              spacing: float = 1.0, N_steps: int = 2000, name: str = "naca") -> CaseConfig:
    """Extruded NACA 0012 heaving and pitching about its quarter chord.

    Inflow and initial flow are pure +x.  The section is on every z layer
    0..Nz-1 with volume weight ds*dz (dz = 1).  The aerodynamic incidence is
    alpha(t) = aoa_mean + pitch_amp sin(w t + phase) with w = 2 k U / c; the
    reference rotates markers counter-clockwise by theta (airfoil_moving.py:
    119-120), which tips the leading edge down, so theta = -alpha."""
    X, Y, ds = naca0012_arclength(chord, spacing, 0.25)
    if le_x is None:
        le_x = Nx / 4.0
    zl = np.arange(0, Nz, 1.0, dtype=np.float32)
    nsec = len(X)
    s_v = _rates(U, chord, Re)
    w = 2.0 * k * U / chord
    return CaseConfig(
        name=name, Nx=Nx, Ny=Ny, Nz=Nz, U_inf=U, Re=Re, chord=chord, s_v=s_v, AoA0_deg=0.0,
        ux_in=U, uy_in=0.0, X0=np.tile(X, Nz), Y0=np.tile(Y, Nz), Z0=np.repeat(zl, nsec),
        dV=np.full(nsec * Nz, ds, dtype=np.float32), x0=le_x + 0.25 * chord, y0=Ny / 2.0,
        A_heave=heave_frac * chord, A_pitch=-math.radians(pitch_amp_deg),
        phi_pitch=math.radians(phase_deg), theta0=-math.radians(aoa_mean_deg),
        f_motion=w / (2.0 * math.pi), N_steps=N_steps, vis_interval=max(1, N_steps // 10),
        DATA_DIR=f"{name}_frames", extra=dict(X_template=X, Y_template=Y, N_MARKERS_2D=nsec, ds=ds))


def case_c0() -> CaseConfig:
    """300 x 100 x 24, chord 40, Re 200 (tau 0.53), LE at (75, 50), 2,000 steps."""
    return naca_case(300, 100, 24, 40.0, le_x=75.0, N_steps=2000, name="naca-300x100x24")


def case_c1() -> CaseConfig:
    """600 x 200 x 50, chord 80, Re 200 (tau 0.56), LE at (150, 100), 30,000 steps."""
    return naca_case(600, 200, 50, 80.0, le_x=150.0, N_steps=30000, name="naca-600x200x50")


def case_c2_large() -> CaseConfig:
    """1200 x 400 x 100, chord 160, Re 200 (tau 0.62), LE at (300, 200)."""
    return naca_case(1200, 400, 100, 160.0, le_x=300.0, N_steps=20000, name="naca-1200x400x100")


# ---------------------------------------------------------------------------
# module interop (the reference's config-module convention)
# ---------------------------------------------------------------------------
def from_module(mod) -> CaseConfig:
    """Build a CaseConfig from a module using the reference's names
    (config_moving.py, config_highRe.py, fish/config_fish.py)."""
    g = lambda n, d=None: getattr(mod, n, d)
    aoa = float(g("AoA0_deg", 0.0))
    U = float(g("U_inf"))
    ux_in = float(U * math.cos(math.radians(aoa)))
    uy_in = float(-U * math.sin(math.radians(aoa)))
    X0 = np.asarray(g("X_marker_np"), dtype=np.float32)
    Y0 = np.asarray(g("Y_marker_np"), dtype=np.float32)
    Z0 = np.asarray(g("Z_marker_np"), dtype=np.float32)
    return CaseConfig(
        name=getattr(mod, "__name__", "module"), Nx=int(g("Nx")), Ny=int(g("Ny")), Nz=int(g("Nz")),
        U_inf=U, Re=float(g("Re", 0.0)), chord=float(g("c", 0.0)), s_v=float(g("s_v")),
        s_e=float(g("s_e", 1.0)), s_q=float(g("s_q", 1.0)), s_pi=float(g("s_pi", 1.0)),
        cs2=float(g("cs2", CS2)), AoA0_deg=aoa, ux_in=ux_in, uy_in=uy_in, X0=X0, Y0=Y0, Z0=Z0,
        dV=None, x0=float(g("x0", 0.0)), y0=float(g("y0", 0.0)), A_heave=float(g("A_heave", 0.0)),
        A_pitch=float(g("A_pitch", 0.0)), phi_pitch=float(g("phi_pitch", 0.0)),
        theta0=float(g("AoA0_rad", math.radians(aoa))), f_motion=float(g("f_motion", 0.0)),
        N_steps=int(g("N_steps", 1000)), vis_interval=int(g("vis_interval", 200)),
        DATA_DIR=str(g("DATA_DIR", "frames")))


def as_module(cfg: CaseConfig, name: str = "config_moving") -> types.ModuleType:
    """A module with the reference's names, e.g. for sys.modules aliasing."""
    m = types.ModuleType(name)
    vals = dict(Nx=cfg.Nx, Ny=cfg.Ny, Nz=cfg.Nz, U_inf=cfg.U_inf, Re=cfg.Re, c=cfg.chord,
                cs2=cfg.cs2, nu=cfg.nu, tau=cfg.tau, s_v=cfg.s_v, s_e=cfg.s_e, s_q=cfg.s_q,
                s_pi=cfg.s_pi, AoA0_deg=cfg.AoA0_deg, AoA0_rad=cfg.theta0,
                X_marker_np=cfg.X0, Y_marker_np=cfg.Y0, Z_marker_np=cfg.Z0,
                N_MARKERS=cfg.N_MARKERS, N_MARKERS_2D=cfg.extra.get("N_MARKERS_2D", cfg.N_MARKERS),
                x0=cfg.x0, y0=cfg.y0, A_heave=cfg.A_heave, A_pitch=cfg.A_pitch,
                phi_pitch=cfg.phi_pitch, f_motion=cfg.f_motion, N_steps=cfg.N_steps,
                vis_interval=cfg.vis_interval, DATA_DIR=cfg.DATA_DIR)
    for k, v in vals.items():
        setattr(m, k, v)
    return m


def resolve(config=None) -> CaseConfig:
    """The case a no-argument constructor uses: an explicit CaseConfig or
    module, else ``sys.modules['config_moving']`` when something aliased it
    there (the main_highRe.py:26-27 convention), else the reference airplane."""
    if isinstance(config, CaseConfig):
        return config
    if config is not None:
        return from_module(config)
    mod = sys.modules.get("config_moving")
    if mod is not None:
        return from_module(mod)
    return reference_airplane()
