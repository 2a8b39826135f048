"""CPU oracle: a NumPy restatement of the reference CLBM-IBM time step.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
module; only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg
of ``bench.py`` use it, and only as the checker.

Every function follows one reference kernel, formula for formula, in the
reference's array layout: populations are ``(Nx, Ny, Nz, 19)`` (q fastest)
and scalar fields ``(Nx, Ny, Nz)``.  The arithmetic runs in the dtype of the
inputs (``float64`` for the exact oracle, ``float32`` to measure the FP32
rounding spread).  Citations are to the reference repository
(kioxli/CLBM_IBM-moving-airfoil).

Pinning (what checks each function, in ``tests/test_oracle_*.py``):

* lattice constants, ``M_raw``/``M_inv``        -- exact identities (M M^-1 = I, dyadic entries)
* ``relaxation_rates`` / ``inlet_velocity``     -- closed forms at every configured (U, c, Re)
                                                   (tau 0.53, 0.56, 0.62, 0.518, 0.5024) and the
                                                   tabulated cos/sin of the 8 degree tilt
* ``init_feq``                                  -- closed-form moments
* ``macro_with_force``                          -- closed form
* ``macro_post``                                -- rho and u recovered from hand-built second-order
                                                   equilibria; fsum of the definition per cell;
                                                   planted x/y swap and dropped direction caught
* ``collide``                                   -- invariants computed by brute-force moment sums:
                                                   mass, momentum gain (1.5 - s_v/2) F, relaxed
                                                   central moments, rest fixed point, Guo moments
* ``stream``                                    -- permutation / brute force on tiny grids
* ``bc_inlet`` / ``bc_outlet``                  -- closed-form moment identities of the inlet rule
* ``bc_inlet_hecht_harting``                    -- imposed density/momentum exactly, normal
                                                   non-equilibrium bounce-back, planted drop of
                                                   the transverse f15..f18 terms caught
* ``phi1d``                                     -- Peskin identities (partition of unity, sum phi^2 = 3/8)
* ``update_markers``                            -- closed form (rotation, rigid velocity)
* ``interpolate`` / ``spread``                  -- adjointness, partition of unity, seam weight 0.75,
                                                   brute-force stencil enumeration
* ``marker_forces``                             -- closed form
* ``clip_forces``                               -- a table of known inputs (both bounds, exact
                                                   f_max, in-range values), fz zero; planted
                                                   one-sided clamp caught
* ``trig_f32``                                  -- nearest float32 of FP64 cos/sin of the float32
                                                   angle over 20,000 angles; planted truncation
                                                   and unrounded angle caught
* ``force_sum_f64``                             -- exact FP64 sum (math.fsum)
* ``f16c_encode`` / ``f16c_decode``             -- known codes, full round trip, monotonicity,
                                                   brute-force nearest rounding of |v| + 2^-15
* ``motion_state``                              -- closed form derivatives
* ``run_coupled``                               -- two steps composed by hand (step 1 reads K7 of
                                                   the initial state, step n of step n-1's state);
                                                   a stale or half-force velocity differs

Taichi itself is not installed here, so bit-level agreement with a Taichi
run is unpinned; the only deliberate choice below is the FP32 marker
rotation (see ``update_markers``), whose cos/sin are rounded from FP64.
"""

from __future__ import annotations

import math

import numpy as np

# --------------------------------------------------------------------------
# Lattice constants -- clbm_moving.py:21-30, config_moving.py:37-43,48-57
# --------------------------------------------------------------------------
EX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0], dtype=np.int64)
EY = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1], dtype=np.int64)
EZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1], dtype=np.int64)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float64)
CS2 = 1.0 / 3.0
Q = 19

# raw-moment ordering, config_moving.py:62-68
PQR = [
    (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1),
    (1, 1, 0), (1, 0, 1), (0, 1, 1),
    (2, 0, 0), (0, 2, 0), (0, 0, 2),
    (2, 1, 0), (1, 2, 0), (2, 0, 1), (0, 2, 1), (1, 0, 2), (0, 1, 2),
    (2, 2, 0), (2, 0, 2), (0, 2, 2),
]


def m_raw() -> np.ndarray:
    """19x19 raw-moment matrix, row = (p,q,r), column = direction (config_moving.py:69-72)."""
    return np.array([[int(EX[c]) ** p * int(EY[c]) ** q * int(EZ[c]) ** r for c in range(Q)]
                     for (p, q, r) in PQR], dtype=np.float64)


def m_raw_inv_f32() -> np.ndarray:
    """Its inverse cast to float32, as stored in CLBMSolver.M_inv (config_moving.py:73,
    clbm_moving.py:66-67)."""
    return np.linalg.inv(m_raw()).astype(np.float32)


def relaxation_rates(U_inf: float, chord: float, Re: float) -> dict:
    """nu, tau, s_v, s_e, s_q, s_pi as config_moving.py:37-43 define them."""
    nu = U_inf * chord / Re
    tau = 0.5 + nu / CS2
    return dict(nu=nu, tau=tau, s_v=1.0 / tau, s_e=1.0, s_q=1.0, s_pi=1.0)


def inlet_velocity(U_inf: float, aoa_deg: float) -> tuple:
    """(_UX_IN, _UY_IN) -- clbm_moving.py:33-35 (inflow tilted by the mean AoA)."""
    a = math.radians(aoa_deg)
    return float(U_inf * math.cos(a)), float(-U_inf * math.sin(a))


# --------------------------------------------------------------------------
# CLBMSolver kernels -- clbm_moving.py
# --------------------------------------------------------------------------
def init_feq(nx: int, ny: int, nz: int, ux0: float, uy0: float, dtype=np.float64):
    """K1 ``_init_feq`` (clbm_moving.py:75-91): second-order equilibrium at (ux0, uy0, 0),
    rho = 1.  Returns (f, rho, ux, uy, uz)."""
    ux0 = dtype(ux0)
    uy0 = dtype(uy0)
    cs2 = dtype(CS2)
    u2 = ux0 * ux0 + uy0 * uy0
    fq = np.empty(Q, dtype=dtype)
    for q in range(Q):
        eu = dtype(EX[q]) * ux0 + dtype(EY[q]) * uy0
        fq[q] = dtype(W[q]) * (dtype(1.0) + eu / cs2 + dtype(0.5) * eu * eu / (cs2 * cs2)
                               - dtype(0.5) * u2 / cs2)
    f = np.broadcast_to(fq, (nx, ny, nz, Q)).copy()
    rho = np.ones((nx, ny, nz), dtype=dtype)
    ux = np.full((nx, ny, nz), ux0, dtype=dtype)
    uy = np.full((nx, ny, nz), uy0, dtype=dtype)
    uz = np.zeros((nx, ny, nz), dtype=dtype)
    return f, rho, ux, uy, uz


def _moments_0_1(f):
    dt = f.dtype.type
    r = np.zeros(f.shape[:3], dtype=f.dtype)
    mx = np.zeros_like(r)
    my = np.zeros_like(r)
    mz = np.zeros_like(r)
    for q in range(Q):          # sequential q order, as the unrolled loop does
        fq = f[..., q]
        r += fq
        mx += dt(EX[q]) * fq
        my += dt(EY[q]) * fq
        mz += dt(EZ[q]) * fq
    return r, mx, my, mz


def macro_with_force(f, fx, fy, fz):
    """K2 ``_macro_with_force`` (clbm_moving.py:96-115): rho, and u = m/rho + F/(2 rho)."""
    dt = f.dtype.type
    r, mx, my, mz = _moments_0_1(f)
    half = dt(0.5)
    return r, mx / r + half * fx / r, my / r + half * fy / r, mz / r + half * fz / r


def macro_post(f):
    """K7 ``_macro_post`` (clbm_moving.py:305-321): force-free rho and u = m/rho."""
    r, mx, my, mz = _moments_0_1(f)
    return r, mx / r, my / r, mz / r


def collide(f, rho, ux, uy, uz, fx, fy, fz, s_v: float, s_e: float = 1.0, rates=None):
    """K3 ``_collide`` (clbm_moving.py:120-249).

    Step 1 (:138-168): 15 central moments about u.  Step 2 (:170-196): shear
    moments x(1-s_v); trace relaxed by s_e toward rho, deviators x(1-s_v),
    recombined with the fixed-sign formula; third order -> 0 and fourth order
    -> rho/9 (s_q = s_pi = 1 are hard-coded there).  Step 3 (:198-231): central
    -> raw moments T.  Step 4 (:233-238): f* = M_inv T with the float32 M_inv.
    Step 5 (:240-249): Guo forcing with coefficient (1 - s_v/2).

    ``rates`` (the learned-closure extension; None = the reference): a tuple
    (s3xy, s3z, s4xy, s4z) of scalars or per-cell arrays.  Third-order central
    moments relax as k3* = (1 - s3) k3 (s3xy for k210, k120; s3z for k201,
    k021, k102, k012) and fourth-order ones as k4* = rho/9 + (1 - s4)(k4 -
    rho/9) (s4xy for k220; s4z for k202, k022).  All ones is the reference.
    """
    dt = f.dtype.type
    u, v, w, r = ux, uy, uz, rho
    one, two, three, four, nine = dt(1), dt(2), dt(3), dt(4), dt(9)
    sv = dt(s_v)
    se = dt(s_e)
    z = np.zeros_like(r)
    k110 = z.copy(); k101 = z.copy(); k011 = z.copy()
    k200 = z.copy(); k020 = z.copy(); k002 = z.copy()
    k210 = z.copy(); k120 = z.copy(); k201 = z.copy(); k021 = z.copy(); k102 = z.copy(); k012 = z.copy()
    k220 = z.copy(); k202 = z.copy(); k022 = z.copy()
    for q in range(Q):
        fq = f[..., q]
        xix = dt(EX[q]) - u
        xiy = dt(EY[q]) - v
        xiz = dt(EZ[q]) - w
        xix2 = xix * xix
        xiy2 = xiy * xiy
        xiz2 = xiz * xiz
        k110 += fq * xix * xiy
        k101 += fq * xix * xiz
        k011 += fq * xiy * xiz
        k200 += fq * xix2
        k020 += fq * xiy2
        k002 += fq * xiz2
        k210 += fq * xix2 * xiy
        k120 += fq * xix * xiy2
        k201 += fq * xix2 * xiz
        k021 += fq * xiy2 * xiz
        k102 += fq * xix * xiz2
        k012 += fq * xiy * xiz2
        k220 += fq * xix2 * xiy2
        k202 += fq * xix2 * xiz2
        k022 += fq * xiy2 * xiz2
    # cascade relaxation (clbm_moving.py:170-196)
    k110s = k110 * (one - sv)
    k101s = k101 * (one - sv)
    k011s = k011 * (one - sv)
    kA = k200 + k020 + k002
    kA_eq = r
    kD = k200 - k020
    kE = k200 - k002
    kAs = kA - se * (kA - kA_eq)
    kDs = kD * (one - sv)
    kEs = kE * (one - sv)
    k200s = (kAs + kDs + kEs) / three
    k020s = (kAs - two * kDs + kEs) / three
    k002s = (kAs + kDs - two * kEs) / three
    if rates is None:
        k210s = k120s = k201s = k021s = k102s = k012s = z
        k220s = r / nine
        k202s = r / nine
        k022s = r / nine
    else:
        s3xy, s3z, s4xy, s4z = (np.asarray(a, dtype=dt) for a in rates)
        k210s, k120s = (one - s3xy) * k210, (one - s3xy) * k120
        k201s, k021s = (one - s3z) * k201, (one - s3z) * k021
        k102s, k012s = (one - s3z) * k102, (one - s3z) * k012
        k220s = r / nine + (one - s4xy) * (k220 - r / nine)
        k202s = r / nine + (one - s4z) * (k202 - r / nine)
        k022s = r / nine + (one - s4z) * (k022 - r / nine)
    # central -> raw (clbm_moving.py:198-223)
    T = [
        r, u * r, v * r, w * r,
        k110s + u * v * r, k101s + u * w * r, k011s + v * w * r,
        k200s + u * u * r, k020s + v * v * r, k002s + w * w * r,
        k210s + two * u * k110s + v * k200s + u * u * v * r,
        k120s + two * v * k110s + u * k020s + u * v * v * r,
        k201s + two * u * k101s + w * k200s + u * u * w * r,
        k021s + two * v * k011s + w * k020s + v * v * w * r,
        k102s + two * w * k101s + u * k002s + u * w * w * r,
        k012s + two * w * k011s + v * k002s + v * w * w * r,
        (k220s + two * u * k120s + two * v * k210s + four * u * v * k110s
         + v * v * k200s + u * u * k020s + u * u * v * v * r),
        (k202s + two * u * k102s + two * w * k201s + four * u * w * k101s
         + w * w * k200s + u * u * k002s + u * u * w * w * r),
        (k022s + two * v * k012s + two * w * k021s + four * v * w * k011s
         + w * w * k020s + v * v * k002s + v * v * w * w * r),
    ]
    minv = m_raw_inv_f32().astype(f.dtype)
    f_post = np.empty_like(f)
    for q in range(Q):          # dense 19x19 matvec, r_idx order (clbm_moving.py:234-238)
        val = np.zeros_like(r)
        for ridx in range(Q):
            val += minv[q, ridx] * T[ridx]
        f_post[..., q] = val
    # Guo forcing (clbm_moving.py:240-249)
    cs2 = dt(CS2)
    coeff = one - sv * dt(0.5)
    for q in range(Q):
        ex, ey, ez = dt(EX[q]), dt(EY[q]), dt(EZ[q])
        eu = ex * u + ey * v + ez * w
        cx = (ex - u) / cs2 + eu * ex / (cs2 * cs2)
        cy = (ey - v) / cs2 + eu * ey / (cs2 * cs2)
        cz = (ez - w) / cs2 + eu * ez / (cs2 * cs2)
        f_post[..., q] += dt(W[q]) * coeff * (cx * fx + cy * fy + cz * fz)
    return f_post


def stream(f_post):
    """K4 ``_stream`` (clbm_moving.py:254-261): pull with modular wrap on all three axes,
    f[i,j,k,q] = f_post[(i-ex)%Nx, (j-ey)%Ny, (k-ez)%Nz, q]."""
    f = np.empty_like(f_post)
    for q in range(Q):
        f[..., q] = np.roll(f_post[..., q], shift=(int(EX[q]), int(EY[q]), int(EZ[q])), axis=(0, 1, 2))
    return f


def bc_inlet(f, ux_in: float, uy_in: float):
    """K5 ``_bc_inlet`` (clbm_moving.py:266-294): the reference's inlet rule at x = 0,
    with its coefficients (2/3 on f1, 1/6 on the four edges), in place."""
    dt = f.dtype.type
    ux_in = dt(ux_in)
    uy_in = dt(uy_in)
    g = f[0]
    s_known = g[..., 0] + g[..., 3] + g[..., 4] + g[..., 5] + g[..., 6] + g[..., 15] + g[..., 16] + g[..., 17] + g[..., 18]
    s_out = g[..., 2] + g[..., 8] + g[..., 10] + g[..., 12] + g[..., 14]
    rho_in = (s_known + dt(2.0) * s_out) / (dt(1.0) - ux_in)
    diff_y = g[..., 3] - g[..., 4]
    diff_z = g[..., 5] - g[..., 6]
    sixth = dt(1.0) / dt(6.0)
    half = dt(0.5)
    g[..., 1] = g[..., 2] + (dt(2.0) / dt(3.0)) * rho_in * ux_in
    g[..., 7] = g[..., 10] + sixth * rho_in * ux_in + half * rho_in * uy_in - half * diff_y
    g[..., 9] = g[..., 8] + sixth * rho_in * ux_in - half * rho_in * uy_in + half * diff_y
    g[..., 11] = g[..., 14] + sixth * rho_in * ux_in - half * diff_z
    g[..., 13] = g[..., 12] + sixth * rho_in * ux_in + half * diff_z
    return f


def bc_inlet_hecht_harting(f, ux_in: float, uy_in: float):
    """The textbook alternative to the reference's inlet (no reference
    counterpart; opt-in): the on-site velocity boundary of Hecht and Harting
    (J. Stat. Mech. 2010, P01018) for D3Q19 at x = 0 with u = (ux_in, uy_in, 0),
    in place.  rho from mass and x-momentum; the normal population by
    bounce-back of its non-equilibrium part, f1 = f2 + rho ux / 3; each
    diagonal from its opposite plus the equilibrium difference
    rho (ux +- u_t) / 6 and a transverse correction N_t fixed by the
    prescribed transverse momentum:
    N_y = (f3 - f4 + f15 - f16 + f17 - f18) / 2 - rho uy / 3,
    N_z = (f5 - f6 + f15 + f16 - f17 - f18) / 2 (uz = 0)."""
    dt = f.dtype.type
    ux_in = dt(ux_in)
    uy_in = dt(uy_in)
    g = f[0]
    s_known = g[..., 0] + g[..., 3] + g[..., 4] + g[..., 5] + g[..., 6] + g[..., 15] + g[..., 16] + g[..., 17] + g[..., 18]
    s_out = g[..., 2] + g[..., 8] + g[..., 10] + g[..., 12] + g[..., 14]
    rho_in = (s_known + dt(2.0) * s_out) / (dt(1.0) - ux_in)
    half, third, sixth = dt(0.5), dt(1.0) / dt(3.0), dt(1.0) / dt(6.0)
    n_y = half * ((g[..., 3] - g[..., 4]) + (g[..., 15] - g[..., 16]) + (g[..., 17] - g[..., 18])) \
        - third * rho_in * uy_in
    n_z = half * ((g[..., 5] - g[..., 6]) + (g[..., 15] + g[..., 16]) - (g[..., 17] + g[..., 18]))
    g[..., 1] = g[..., 2] + third * rho_in * ux_in
    g[..., 7] = g[..., 10] + sixth * rho_in * (ux_in + uy_in) - n_y
    g[..., 9] = g[..., 8] + sixth * rho_in * (ux_in - uy_in) + n_y
    g[..., 11] = g[..., 14] + sixth * rho_in * ux_in - n_z
    g[..., 13] = g[..., 12] + sixth * rho_in * ux_in + n_z
    return f


def bc_outlet(f):
    """K6 ``_bc_outlet`` (clbm_moving.py:296-300): f[Nx-1] = f[Nx-2], all q, in place."""
    f[-1] = f[-2]
    return f


def lbm_step(f, fx, fy, fz, s_v: float, s_e: float, ux_in: float, uy_in: float, rates=None,
             inlet=True):
    """``CLBMSolver.step`` (clbm_moving.py:326-343): K2 -> K3 -> K4 -> K5 -> K6 -> K7.
    Returns (f_new, rho, ux, uy, uz) with the K7 (force-free) macroscopic fields.
    ``rates``: per-cell higher-order rates (see ``collide``); ``inlet`` False
    leaves x periodic, True (or "reference") applies the reference's inlet
    rule, "hecht-harting" the textbook one (``bc_inlet_hecht_harting``)."""
    rho, ux, uy, uz = macro_with_force(f, fx, fy, fz)
    f_post = collide(f, rho, ux, uy, uz, fx, fy, fz, s_v, s_e, rates=rates)
    f_new = stream(f_post)
    if inlet:
        if inlet == "hecht-harting":
            bc_inlet_hecht_harting(f_new, ux_in, uy_in)
        else:
            bc_inlet(f_new, ux_in, uy_in)
        bc_outlet(f_new)
    rho, ux, uy, uz = macro_post(f_new)
    return f_new, rho, ux, uy, uz


# --------------------------------------------------------------------------
# 16-bit deviation format 1-4-11 (the build's own storage option; the
# reference stores float32).  Restated with numpy integer operations.
# --------------------------------------------------------------------------
def f16c_encode(v):
    """float32 -> uint16: sign bit and 15-bit code u of F = fl32(|v| + 2^-15)
    rounded to 12 significant bits (nearest, ties to even), F's exponent in 4
    bits (codes 0-14: F in [2^-15, 1)) and 11 mantissa bits; exponent code 15
    (u >= 0x7800) holds the non-finite codes, which values rounding to F >= 1
    become (0x7fff for F >= 2, inf and NaN)."""
    v = np.asarray(v, dtype=np.float32)
    b = v.view(np.uint32).astype(np.uint64)
    sign = (b >> 16) & 0x8000
    a = np.abs(v)
    with np.errstate(invalid="ignore", over="ignore"):
        t = (a + np.float32(2.0 ** -15)).astype(np.float32).view(np.uint32).astype(np.uint64)
    r = t + 0x7FF + ((t >> 12) & 1)
    h = np.minimum((r - (112 << 23)) >> 12, 0x7FFF)
    return (h | sign).astype(np.uint16)


def f16c_decode(h):
    """uint16 -> float32: +-F(u) minus 2^-15 with the same sign, for exponent
    codes 0-14 (so 0x8000 gives +0); 0x7800 is infinity and the codes above
    it NaN (with the sign)."""
    h = np.asarray(h, dtype=np.uint16).astype(np.uint64)
    u = h & 0x7FFF
    F = (((u << 12) + (112 << 23)) & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    F = np.where(u >= 0x7800, np.float32(np.inf), F).astype(np.float32)
    F = np.where(u > 0x7800, np.float32(np.nan), F).astype(np.float32)
    Fs = np.where(h & 0x8000, -F, F).astype(np.float32)
    with np.errstate(invalid="ignore"):
        return (Fs - np.copysign(np.float32(2.0 ** -15), Fs)).astype(np.float32)


# --------------------------------------------------------------------------
# IBM kernels -- airfoil_moving.py
# --------------------------------------------------------------------------
def phi1d(r):
    """Peskin 4-point kernel ``_phi1d`` (airfoil_moving.py:27-38), elementwise in r's dtype."""
    r = np.asarray(r)
    dt = r.dtype.type if r.dtype.kind == "f" else np.float64
    r = r.astype(dt)
    s = np.abs(r)
    a = (dt(3.0) - dt(2.0) * s + np.sqrt(np.maximum(dt(1.0) + dt(4.0) * s - dt(4.0) * s * s, dt(0.0)))) / dt(8.0)
    b = (dt(5.0) - dt(2.0) * s - np.sqrt(np.maximum(dt(-7.0) + dt(12.0) * s - dt(4.0) * s * s, dt(0.0)))) / dt(8.0)
    return np.where(s < dt(1.0), a, np.where(s < dt(2.0), b, dt(0.0))).astype(dt)


def trig_f32(theta):
    """cos/sin of a float32 angle, evaluated in FP64 and rounded once to float32.

    The reference evaluates ti.cos/ti.sin in float32 (airfoil_moving.py:114-115),
    whose last-bit behaviour is Taichi's and unpinned; this build fixes the
    rounding as 'FP64 then one rounding', which the GPU path reproduces."""
    t = np.float64(np.float32(theta))
    return np.float32(math.cos(t)), np.float32(math.sin(t))


def update_markers(X0, Y0, Z0, xc, yc, theta, vx, vy, omega):
    """K8 ``_update_markers`` (airfoil_moving.py:100-127) in float32.

    x = xc + X0 cos - Y0 sin ; y = yc + X0 sin + Y0 cos ; z = Z0 ;
    vx_b = vx - omega (y - yc) ; vy_b = vy + omega (x - xc).
    Host scalars are truncated to float32 as the ti.f32 kernel arguments are
    (airfoil_moving.py:102-103).  Each product and sum is rounded separately
    (no fused multiply-add), left to right as written at :119-127."""
    f32 = np.float32
    X0 = np.asarray(X0, dtype=f32)
    Y0 = np.asarray(Y0, dtype=f32)
    Z0 = np.asarray(Z0, dtype=f32)
    xc, yc, vx, vy, omega = (f32(a) for a in (xc, yc, vx, vy, omega))
    c, s = trig_f32(theta)
    X = (xc + X0 * c) - Y0 * s
    Y = (yc + X0 * s) + Y0 * c
    Z = Z0.copy()
    VX = vx - omega * (Y - yc)
    VY = vy + omega * (X - xc)
    return X.astype(f32), Y.astype(f32), Z, VX.astype(f32), VY.astype(f32)


def stencil(X, Y, Z, nx: int, ny: int, nz: int, wrap_weight_quirk: bool = True):
    """Shared stencil of ``_interpolate``/``_spread`` (airfoil_moving.py:148-166,194-212).

    base = floor(pos) - 1 per axis; x index clamped to [0, Nx-1]; y and z
    indices wrapped (Python-style non-negative modulo).  The weight of each
    axis is phi(float(index used) - pos): with the quirk on (the reference),
    the *wrapped/clamped* index enters phi, so a stencil that crosses the
    periodic seam loses weight.  Returns (base (N,3) int32, ii/jj/kk (N,4)
    int64, phx/phy/phz (N,4) in the dtype of X)."""
    X = np.asarray(X)
    Y = np.asarray(Y)
    Z = np.asarray(Z)
    dt = X.dtype.type
    ib = np.floor(X).astype(np.int64) - 1
    jb = np.floor(Y).astype(np.int64) - 1
    kb = np.floor(Z).astype(np.int64) - 1
    d = np.arange(4)
    iu = ib[:, None] + d
    ju = jb[:, None] + d
    ku = kb[:, None] + d
    ii = np.clip(iu, 0, nx - 1)
    jj = np.mod(ju + ny, ny)
    kk = np.mod(ku + nz, nz)
    if wrap_weight_quirk:
        ax, ay, az = ii, jj, kk
    else:
        ax, ay, az = iu, ju, ku
    phx = phi1d(ax.astype(dt) - X[:, None])
    phy = phi1d(ay.astype(dt) - Y[:, None])
    phz = phi1d(az.astype(dt) - Z[:, None])
    base = np.stack([ib, jb, kb], axis=1).astype(np.int32)
    return base, ii, jj, kk, phx, phy, phz


def interpolate(ux, uy, X, Y, Z, wrap_weight_quirk: bool = True):
    """K10 ``_interpolate`` (airfoil_moving.py:142-170): U_k, V_k by the 4x4x4
    Peskin stencil, summed in (di, dj, dk) loop order with w = phx*phy*phz."""
    nx, ny, nz = ux.shape
    dt = np.asarray(X).dtype.type
    _, ii, jj, kk, phx, phy, phz = stencil(X, Y, Z, nx, ny, nz, wrap_weight_quirk)
    n = len(X)
    U = np.zeros(n, dtype=dt)
    V = np.zeros(n, dtype=dt)
    for di in range(4):
        for dj in range(4):
            for dk in range(4):
                w = phx[:, di] * phy[:, dj] * phz[:, dk]
                a, b, c = ii[:, di], jj[:, dj], kk[:, dk]
                U = U + ux[a, b, c].astype(dt) * w
                V = V + uy[a, b, c].astype(dt) * w
    return U, V


def marker_forces(U, V, VX, VY, f_max: float = 0.05):
    """K11 ``_marker_forces`` (airfoil_moving.py:175-183): F = clamp(-(U - V_body), +-f_max)."""
    dt = np.asarray(U).dtype.type
    fm = dt(f_max)
    FX = np.maximum(-fm, np.minimum(fm, -(U - VX)))
    FY = np.maximum(-fm, np.minimum(fm, -(V - VY)))
    return FX.astype(dt), FY.astype(dt)


def spread(X, Y, Z, FX, FY, shape, wrap_weight_quirk: bool = True, acc_dtype=np.float64):
    """K9 + K12 ``_zero_forces``/``_spread`` (airfoil_moving.py:132-137,188-214):
    fx[ii,jj,kk] += F_k * w over the same stencil as ``interpolate``.  The
    product F_k*w is formed in the dtype of the markers; the sum is accumulated
    in ``acc_dtype`` (the reference's atomic order is undefined, so only a
    tolerance can compare sums)."""
    nx, ny, nz = shape
    dt = np.asarray(X).dtype.type
    _, ii, jj, kk, phx, phy, phz = stencil(X, Y, Z, nx, ny, nz, wrap_weight_quirk)
    fx = np.zeros(shape, dtype=acc_dtype)
    fy = np.zeros(shape, dtype=acc_dtype)
    for di in range(4):
        for dj in range(4):
            for dk in range(4):
                w = phx[:, di] * phy[:, dj] * phz[:, dk]
                a, b, c = ii[:, di], jj[:, dj], kk[:, dk]
                np.add.at(fx, (a, b, c), (FX * w).astype(dt).astype(acc_dtype))
                np.add.at(fy, (a, b, c), (FY * w).astype(dt).astype(acc_dtype))
    return fx, fy


def clip_forces(fx, fy, f_max: float = 0.05, dtype=np.float32):
    """K13 ``_clip_forces`` (airfoil_moving.py:219-225): clamp fx, fy to +-f_max; fz = 0."""
    fm = f_max
    fxc = np.clip(fx, -fm, fm).astype(dtype)
    fyc = np.clip(fy, -fm, fm).astype(dtype)
    return fxc, fyc, np.zeros_like(fxc)


def compute_force(ux, uy, X0, Y0, Z0, xc, yc, theta, vx, vy, omega,
                  f_max: float = 0.05, wrap_weight_quirk: bool = True, acc_dtype=np.float64,
                  weight_dtype=np.float32):
    """``DroneMovingIBM.compute_force`` (airfoil_moving.py:230-255): K8 -> K9 -> K10 ->
    K11 -> K12 -> K13.  Marker positions are always the float32 values of K8;
    ``weight_dtype`` sets the arithmetic of the stencil weights, interpolation
    and marker forces (float32 as in the reference, float64 for the exact
    oracle).  Returns a dict with every intermediate."""
    X, Y, Z, VX, VY = update_markers(X0, Y0, Z0, xc, yc, theta, vx, vy, omega)
    wd = weight_dtype
    X, Y, Z, VX, VY = (a.astype(wd) for a in (X, Y, Z, VX, VY))
    U, V = interpolate(ux, uy, X, Y, Z, wrap_weight_quirk)
    FX, FY = marker_forces(U, V, VX, VY, f_max)
    fx, fy = spread(X, Y, Z, FX, FY, ux.shape, wrap_weight_quirk, acc_dtype)
    fxc, fyc, fzc = clip_forces(fx, fy, f_max)
    return dict(X=X, Y=Y, Z=Z, VX=VX, VY=VY, U=U, V=V, FX=FX, FY=FY,
                fx_raw=fx, fy_raw=fy, fx=fxc, fy=fyc, fz=fzc)


def force_sum_f64(FX, FY, dV=None):
    """FP64 marker force sums (sum_k F_k dV_k) with exact (fsum) accumulation.
    The reference never reduces forces; this is the build's own readout."""
    FX = np.asarray(FX, dtype=np.float64)
    FY = np.asarray(FY, dtype=np.float64)
    dv = np.ones_like(FX) if dV is None else np.asarray(dV, dtype=np.float64)
    return math.fsum(FX * dv), math.fsum(FY * dv)


def motion_state(step: int, x0: float, y0: float, A_heave: float, A_pitch: float,
                 phi_pitch: float, AoA0_rad: float, w_motion: float):
    """``_motion_state`` (main_moving.py:65-74): heave + pitch with analytic rates."""
    t = float(step)
    w = w_motion
    xc = x0
    yc = y0 + A_heave * math.sin(w * t)
    theta = AoA0_rad + A_pitch * math.sin(w * t + phi_pitch)
    vx = 0.0
    vy = A_heave * w * math.cos(w * t)
    omega = A_pitch * w * math.cos(w * t + phi_pitch)
    return xc, yc, theta, vx, vy, omega


def run_coupled(f, X0, Y0, Z0, steps, motion, s_v, s_e, ux_in, uy_in,
                f_max=0.05, wrap_weight_quirk=True, first_step=1, dV=None, dtype=None,
                inlet=True):
    """The driver loop of main_moving.py:116-122 for ``steps`` steps: per step
    motion -> compute_force(ux, uy, ...) -> lbm step.  ``motion(step)`` returns
    the 6-tuple of ``motion_state``.  ``f`` holds the state the loop starts
    from; the force-free u that the first compute_force reads is K7 of it (at
    step 1 of a fresh run that equals K1's u).  Returns (f, history) with the
    per-step FP64 force sums.  ``inlet`` as in ``lbm_step``."""
    dtype = dtype or f.dtype.type
    rho, ux, uy, uz = macro_post(f)
    hist = []
    for n in range(first_step, first_step + steps):
        xc, yc, th, vx, vy, om = motion(n)
        out = compute_force(ux, uy, X0, Y0, Z0, xc, yc, th, vx, vy, om, f_max, wrap_weight_quirk,
                            weight_dtype=dtype)
        fx = out["fx"].astype(dtype)
        fy = out["fy"].astype(dtype)
        fz = out["fz"].astype(dtype)
        f, rho, ux, uy, uz = lbm_step(f, fx, fy, fz, s_v, s_e, ux_in, uy_in, inlet=inlet)
        hist.append(force_sum_f64(out["FX"], out["FY"], dV))
    return f, np.array(hist, dtype=np.float64)


# --------------------------------------------------------------------------
# Learned closure (the build's own extension; no reference counterpart).
# --------------------------------------------------------------------------
# The 19 features of a cell: its moments about the half-force velocity u,
# divided by rho, minus their equilibrium values, in the raw-moment order of
# PQR: rho - 1, three first central moments (= -F/2/rho), k110, k101, k011,
# k200 - 1/3, k020 - 1/3, k002 - 1/3, the six third-order moments, and
# k220 - 1/9, k202 - 1/9, k022 - 1/9.
FEATURE_EQ = np.array([1.0, 0, 0, 0, 0, 0, 0, 1 / 3, 1 / 3, 1 / 3, 0, 0, 0, 0, 0, 0, 1 / 9, 1 / 9,
                       1 / 9])


def central_features(f, fx, fy, fz):
    """(..., 19) float64 features of the populations f (..., 19) with force
    (fx, fy, fz): central moments about u = (sum e f + F/2) / rho."""
    f = np.asarray(f, dtype=np.float64)
    rho = f.sum(-1)
    u = ((f * EX).sum(-1) + 0.5 * np.asarray(fx, np.float64)) / rho
    v = ((f * EY).sum(-1) + 0.5 * np.asarray(fy, np.float64)) / rho
    w = ((f * EZ).sum(-1) + 0.5 * np.asarray(fz, np.float64)) / rho
    xi = [EX - u[..., None], EY - v[..., None], EZ - w[..., None]]
    out = np.empty(f.shape[:-1] + (Q,))
    for i, (p, q, r) in enumerate(PQR):
        out[..., i] = (f * xi[0] ** p * xi[1] ** q * xi[2] ** r).sum(-1) / rho
    out[..., 0] = rho
    return out - FEATURE_EQ


def gelu(x):
    """GELU, tanh form: x/2 (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))."""
    return 0.5 * x * (1.0 + np.tanh(0.7978845608028654 * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(0.7978845608028654 * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * 0.7978845608028654 * (1.0 + 3 * 0.044715 * x * x)


def mlp_forward(params, X):
    """19 -> 32 -> 32 -> 4 with GELU: params (W1 (19,32), b1, W2 (32,32), b2,
    W3 (32,4), b3); returns (Y, cache)."""
    W1, b1, W2, b2, W3, b3 = params
    z1 = X @ W1 + b1
    h1 = gelu(z1)
    z2 = h1 @ W2 + b2
    h2 = gelu(z2)
    Y = h2 @ W3 + b3
    return Y, (X, z1, h1, z2, h2)


def mlp_backward(params, cache, dY):
    """Gradients of sum(dY * Y) with respect to the parameters."""
    W1, b1, W2, b2, W3, b3 = params
    X, z1, h1, z2, h2 = cache
    gW3 = h2.T @ dY
    gb3 = dY.sum(0)
    dh2 = dY @ W3.T
    dz2 = dh2 * gelu_grad(z2)
    gW2 = h1.T @ dz2
    gb2 = dz2.sum(0)
    dh1 = dz2 @ W2.T
    dz1 = dh1 * gelu_grad(z1)
    gW1 = X.T @ dz1
    gb1 = dz1.sum(0)
    return gW1, gb1, gW2, gb2, gW3, gb3


def closure_loss(params, X, T, wgt):
    """Weighted mean squared rate error and its gradient: with one weight per
    sample (wgt (n,)) sum_n w_n |Y_n - T_n|^2 / sum_n w_n, with one per output
    (wgt (n, 4)) sum_nj w_nj (Y_nj - T_nj)^2 / sum_nj w_nj."""
    Y, cache = mlp_forward(params, X)
    d = Y - T
    w = wgt[:, None] if wgt.ndim == 1 else wgt
    W = wgt.sum()
    loss = float((w * d * d).sum() / W)
    grads = mlp_backward(params, cache, 2.0 * w * d / W)
    return loss, grads


def closure_rates(Y, fallback_lo=0.0, fallback_hi=2.0):
    """Rates s = 1 + Y; a cell whose rates leave [lo, hi] takes the plain
    cascade (all ones)."""
    s = 1.0 + np.asarray(Y, dtype=np.float64)
    bad = np.any((s < fallback_lo) | (s > fallback_hi), axis=-1)
    s[bad] = 1.0
    return s
