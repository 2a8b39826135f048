"""GPU parity of the immersed-boundary kernels against the CPU oracle.

Bit-exact: marker positions (float32 rotation without contraction, cos/sin
rounded once from float64) and the stencil base indices floor(pos) - 1.
Tolerances: interpolated velocities and marker forces 1e-6 absolute (FP32
weights and a different summation order than the oracle's loop); spread
forces 1e-7 absolute (each cell sums at most a few dozen products <= 6e-3).
"""

import ctypes as C

import numpy as np
import pytest

from conftest import load_sub

pytestmark = pytest.mark.gpu


def _cfg(pkg, nx=24, ny=20, nz=8, chord=8.0, z_all=True, spacing=0.7):
    X, Y, ds = pkg.config.naca0012_arclength(chord, spacing, 0.25)
    zl = np.arange(0, nz, 1.0, dtype=np.float32) if z_all else np.array([3.0, 4.0], np.float32)
    return pkg.config.CaseConfig(
        name="ibm", Nx=nx, Ny=ny, Nz=nz, U_inf=0.05, Re=100, chord=chord, s_v=1 / 0.6,
        ux_in=0.05, uy_in=0.0, X0=np.tile(X, len(zl)), Y0=np.tile(Y, len(zl)),
        Z0=np.repeat(zl, len(X)), dV=np.full(len(X) * len(zl), ds, np.float32),
        x0=nx / 3.0, y0=ny / 2.0)


def _fields(shape, seed):
    rng = np.random.default_rng(seed)
    ux = (0.05 + 0.01 * rng.standard_normal(shape)).astype(np.float32)
    uy = (0.01 * rng.standard_normal(shape)).astype(np.float32)
    return ux, uy


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).transpose(2, 1, 0))).cuda()


KIN = [(8.3, 10.2, 0.31, 0.0, 0.004, -0.002), (7.9, 9.6, -0.52, 0.01, -0.003, 0.001),
       (8.0, 10.0, 0.0, 0.0, 0.0, 0.0), (8.25, 10.5, 2.9, 0.0, 0.0, 0.01)]


@pytest.mark.parametrize("kin", KIN)
def test_update_markers_bit_exact(pkg, oracle, gpu, kin):
    cfg = _cfg(pkg)
    ibm = pkg.DroneMovingIBM(cfg, record_stencil=True)
    ux, uy = _fields((cfg.Nx, cfg.Ny, cfg.Nz), 0)
    ibm.compute_force(_dev(ux), _dev(uy), *kin)
    X, Y, Z, VX, VY = oracle.update_markers(cfg.X0, cfg.Y0, cfg.Z0, *kin)
    for got, ref in ((ibm.X_marker, X), (ibm.Y_marker, Y), (ibm.Z_marker, Z), (ibm.VX_body, VX),
                     (ibm.VY_body, VY)):
        np.testing.assert_array_equal(got.to_numpy(), ref)
    base, *_ = oracle.stencil(X, Y, Z, cfg.Nx, cfg.Ny, cfg.Nz)
    np.testing.assert_array_equal(ibm.stencil_base.cpu().numpy(), base)


def test_stencil_base_bit_exact_near_integers(pkg, oracle, gpu):
    """Markers placed on and 1 ulp either side of integer coordinates, and on
    the periodic seam layers z = 0 and z = Nz - 1."""
    nx, ny, nz = 20, 16, 6
    pts = []
    for v in (5.0, 7.0, 9.0):
        for d in (np.float32(v), np.nextafter(np.float32(v), np.float32(0)),
                  np.nextafter(np.float32(v), np.float32(100))):
            pts.append(d)
    pts = np.array(pts, np.float32)
    zz = np.array([0.0, np.nextafter(np.float32(0), np.float32(1)), nz - 1.0, 2.5], np.float32)
    X0 = np.repeat(pts, len(zz)) - np.float32(10.0)
    Y0 = np.tile(pts[::-1], len(zz)) - np.float32(8.0)
    Z0 = np.repeat(zz, len(pts))
    cfg = pkg.config.CaseConfig(name="pts", Nx=nx, Ny=ny, Nz=nz, U_inf=0.05, Re=100, chord=4,
                                s_v=1.5, X0=X0.astype(np.float32), Y0=Y0.astype(np.float32),
                                Z0=Z0, x0=10.0, y0=8.0)
    ibm = pkg.DroneMovingIBM(cfg, record_stencil=True)
    ux, uy = _fields((nx, ny, nz), 1)
    ibm.compute_force(_dev(ux), _dev(uy), 10.0, 8.0, 0.0, 0.0, 0.0, 0.0)
    X, Y, Z, *_ = oracle.update_markers(cfg.X0, cfg.Y0, cfg.Z0, 10.0, 8.0, 0.0, 0.0, 0.0, 0.0)
    base, *_ = oracle.stencil(X, Y, Z, nx, ny, nz)
    np.testing.assert_array_equal(ibm.stencil_base.cpu().numpy(), base)


@pytest.mark.parametrize("quirk", [True, False])
@pytest.mark.parametrize("kin", KIN[:2])
def test_compute_force_from_fields_matches_oracle(pkg, oracle, gpu, quirk, kin):
    cfg = _cfg(pkg)
    ibm = pkg.DroneMovingIBM(cfg, wrap_weight_quirk=quirk)
    ux, uy = _fields((cfg.Nx, cfg.Ny, cfg.Nz), 2)
    fx, fy, fz = ibm.compute_force(_dev(ux), _dev(uy), *kin)
    ref = oracle.compute_force(ux.astype(np.float64), uy.astype(np.float64), cfg.X0, cfg.Y0, cfg.Z0,
                               *kin, wrap_weight_quirk=quirk, weight_dtype=np.float64)
    np.testing.assert_allclose(ibm.U_k.to_numpy(), ref["U"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(ibm.V_k.to_numpy(), ref["V"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(ibm.FX_k.to_numpy(), ref["FX"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(ibm.FY_k.to_numpy(), ref["FY"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(fx.to_numpy(), ref["fx"], rtol=0, atol=1e-7)
    np.testing.assert_allclose(fy.to_numpy(), ref["fy"], rtol=0, atol=1e-7)
    assert not fz.to_numpy().any()


def test_spread_modes_agree_and_fixed_is_reproducible(pkg, oracle, gpu):
    cfg = _cfg(pkg)
    ux, uy = _fields((cfg.Nx, cfg.Ny, cfg.Nz), 3)
    outs = {}
    for mode in ("fixed", "atomic"):
        ibm = pkg.DroneMovingIBM(cfg, spread_mode=mode)
        a = ibm.compute_force(_dev(ux), _dev(uy), *KIN[0])[0].to_numpy()
        b = ibm.compute_force(_dev(ux), _dev(uy), *KIN[0])[0].to_numpy()
        outs[mode] = (a, b)
    np.testing.assert_array_equal(outs["fixed"][0], outs["fixed"][1])
    # float32 atomic sums of <= ~30 products of size <= 6e-3: a few ulps of 0.05
    np.testing.assert_allclose(outs["atomic"][0], outs["fixed"][0], rtol=0, atol=5e-8)


def test_force_box_clears_previous_cells(pkg, oracle, gpu):
    """After the body moves, cells touched only by the previous spread read zero."""
    cfg = _cfg(pkg, nx=40)
    ibm = pkg.DroneMovingIBM(cfg)
    ux, uy = _fields((cfg.Nx, cfg.Ny, cfg.Nz), 4)
    ibm.compute_force(_dev(ux), _dev(uy), 10.0, 10.0, 0.2, 0.0, 0.0, 0.0)
    fx, fy, _ = ibm.compute_force(_dev(ux), _dev(uy), 26.0, 9.0, -0.3, 0.0, 0.0, 0.0)
    ref = oracle.compute_force(ux.astype(np.float64), uy.astype(np.float64), cfg.X0, cfg.Y0,
                               cfg.Z0, 26.0, 9.0, -0.3, 0.0, 0.0, 0.0, weight_dtype=np.float64)
    np.testing.assert_allclose(fx.to_numpy(), ref["fx"], rtol=0, atol=1e-7)
    np.testing.assert_allclose(fy.to_numpy(), ref["fy"], rtol=0, atol=1e-7)
    box = ibm.force_box()
    nz = np.argwhere(ref["fx_raw"] != 0)
    assert (box[0:3] <= nz.min(0)).all() and (box[3:6] >= nz.max(0)).all()


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_compute_force_from_solver_state(pkg, oracle, gpu, storage):
    """ux/uy passed as the solver's own fields: the kernel computes the
    force-free velocity from the populations (K7) at each stencil cell, at
    both AA parities."""
    cfg = _cfg(pkg)
    shape = (cfg.Nx, cfg.Ny, cfg.Nz)
    lbm = pkg.CLBMSolver(cfg, storage=storage)
    rng = np.random.default_rng(5)
    f, *_ = oracle.init_feq(*shape, 0.05, 0.0)
    f = f * (1 + 0.02 * rng.standard_normal(f.shape))
    lbm.f.from_numpy(f)
    ibm = pkg.DroneMovingIBM(cfg)
    for n in range(2):
        fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *KIN[n])
        _, ux, uy, _ = oracle.macro_post(f)
        ref = oracle.compute_force(ux, uy, cfg.X0, cfg.Y0, cfg.Z0, *KIN[n], weight_dtype=np.float64)
        np.testing.assert_allclose(ibm.U_k.to_numpy(), ref["U"], rtol=0, atol=1e-6)
        np.testing.assert_allclose(fx.to_numpy(), ref["fx"], rtol=0, atol=1e-7)
        lbm.step(fx, fy, fz)
        f, *_ = oracle.lbm_step(f, ref["fx"].astype(np.float64), ref["fy"].astype(np.float64),
                                0 * ref["fx"].astype(np.float64), cfg.s_v, cfg.s_e, cfg.ux_in,
                                cfg.uy_in)
        np.testing.assert_allclose(lbm.f.to_numpy(), f, rtol=0, atol=2e-6)


@pytest.mark.parametrize("nsteps", [1, 2])
def test_interp_near_inlet_and_outlet_planes(pkg, oracle, gpu, nsteps):
    """Markers whose stencils reach x = 0 and x = Nx-1 read the post-boundary
    velocity (inlet rule, outlet copy) at both AA parities, whether the
    velocity comes from the band pass or straight from the populations."""
    lib = load_sub("_lib")
    nx, ny, nz = 20, 12, 4
    X0 = np.array([0.4, 1.2, 18.3, 19.0, 10.0], np.float32) - np.float32(10.0)
    Y0 = np.array([5.5, 6.1, 4.2, 7.7, 6.0], np.float32) - np.float32(6.0)
    Z0 = np.array([1.0, 2.5, 0.0, 3.0, 2.0], np.float32)
    cfg = pkg.config.CaseConfig(name="planes", Nx=nx, Ny=ny, Nz=nz, U_inf=0.05, Re=100, chord=4,
                                s_v=1 / 0.6, ux_in=0.05, uy_in=-0.004, X0=X0, Y0=Y0, Z0=Z0,
                                x0=10.0, y0=6.0)
    lbm = pkg.CLBMSolver(cfg)
    rng = np.random.default_rng(7)
    f, *_ = oracle.init_feq(nx, ny, nz, 0.05, -0.004)
    lbm.f.from_numpy(f * (1 + 0.02 * rng.standard_normal(f.shape)))
    for _ in range(nsteps):
        lbm.step()
    assert lbm.state_word & lib.STATE_PLANES
    ibm = pkg.DroneMovingIBM(cfg)
    ibm.compute_force(lbm.ux, lbm.uy, 10.0, 6.0, 0.0, 0.0, 0.0, 0.0)
    f_ref = lbm.f.to_numpy().astype(np.float64)
    _, ux, uy, _ = oracle.macro_post(f_ref)
    ref = oracle.compute_force(ux, uy, X0, Y0, Z0, 10.0, 6.0, 0.0, 0.0, 0.0, 0.0,
                               weight_dtype=np.float64)
    np.testing.assert_allclose(ibm.U_k.to_numpy(), ref["U"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(ibm.V_k.to_numpy(), ref["V"], rtol=0, atol=1e-6)
    # the direct path: velocity computed per stencil point from the populations
    import torch
    U = torch.zeros(len(X0), dtype=torch.float32, device="cuda")
    V = torch.zeros_like(U)
    FX = torch.zeros_like(U)
    FY = torch.zeros_like(U)
    p = lambda t: C.c_void_p(t.data_ptr())
    lib.call("ibm_interp_force", C.byref(lbm.grid), C.byref(ibm.ip), C.byref(lbm.params),
             C.c_void_p(lbm.state.data_ptr()), lbm.state_word, None, None, len(X0),
             p(ibm.X_marker), p(ibm.Y_marker), p(ibm.Z_marker), p(ibm.VX_body), p(ibm.VY_body),
             p(U), p(V), p(FX), p(FY), None, None)
    np.testing.assert_allclose(U.cpu().numpy(), ref["U"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(FY.cpu().numpy(), ref["FY"], rtol=0, atol=1e-6)


def test_force_sum_f64(pkg, oracle, gpu):
    cfg = _cfg(pkg)
    ibm = pkg.DroneMovingIBM(cfg)
    ux, uy = _fields((cfg.Nx, cfg.Ny, cfg.Nz), 6)
    ibm.compute_force(_dev(ux), _dev(uy), *KIN[0])
    got = ibm.force_sums().cpu().numpy()
    ref = oracle.force_sum_f64(ibm.FX_k.to_numpy(), ibm.FY_k.to_numpy(), cfg.dV)
    np.testing.assert_allclose(got, ref, rtol=1e-14, atol=1e-15)


def test_ibm_rejects_bad_arguments(pkg, gpu):
    lib = load_sub("_lib")
    g = lib.ClbmGrid(8, 8, 8, 0)
    ip = lib.IbmParams(0.05, 1, 7, 0)
    with pytest.raises(lib.ClbmError, match="spread mode"):
        lib.call("ibm_spread", C.byref(g), C.byref(ip), 1, *([C.c_void_p(1)] * 7), None)
    ip = lib.IbmParams(-1.0, 1, 0, 0)
    with pytest.raises(lib.ClbmError, match="f_max"):
        lib.call("ibm_clip_forces", C.byref(g), C.byref(ip), *([C.c_void_p(1)] * 5), 1, None)
    ip = lib.IbmParams(0.05, 1, 0, 0)
    with pytest.raises(lib.ClbmError, match="generation"):
        lib.call("ibm_clip_forces", C.byref(g), C.byref(ip), *([C.c_void_p(1)] * 5), 0, None)


@pytest.mark.parametrize("quirk", [True, False])
@pytest.mark.parametrize("kin", KIN[:2])
def test_sorted_spread_matches_oracle_and_repeats_bit_for_bit(pkg, oracle, gpu, quirk, kin):
    """The sort-by-cell spread (each cell's stencil contributions grouped and
    added in marker order) against the FP64 oracle, to the tolerances of the
    fixed-point path; two calls give the same fields bit for bit, and the
    fixed-point path agrees within a few float32 roundings of the sums."""
    cfg = _cfg(pkg)
    ux, uy = _fields((cfg.Nx, cfg.Ny, cfg.Nz), 7)
    ibm = pkg.DroneMovingIBM(cfg, spread_mode="sorted", wrap_weight_quirk=quirk,
                             record_stencil=True)
    fx, fy, fz = ibm.compute_force(_dev(ux), _dev(uy), *kin)
    a = (fx.to_numpy(), fy.to_numpy())
    ref = oracle.compute_force(ux.astype(np.float64), uy.astype(np.float64), cfg.X0, cfg.Y0, cfg.Z0,
                               *kin, wrap_weight_quirk=quirk, weight_dtype=np.float64)
    np.testing.assert_allclose(ibm.FX_k.to_numpy(), ref["FX"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(ibm.FY_k.to_numpy(), ref["FY"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(a[0], ref["fx"], rtol=0, atol=1e-7)
    np.testing.assert_allclose(a[1], ref["fy"], rtol=0, atol=1e-7)
    assert not fz.to_numpy().any()
    X, Y, Z, *_ = oracle.update_markers(cfg.X0, cfg.Y0, cfg.Z0, *kin)
    base, *_ = oracle.stencil(X, Y, Z, cfg.Nx, cfg.Ny, cfg.Nz)
    np.testing.assert_array_equal(ibm.stencil_base.cpu().numpy(), base)
    fx2, fy2, _ = ibm.compute_force(_dev(ux), _dev(uy), *kin)
    np.testing.assert_array_equal(fx2.to_numpy(), a[0])
    np.testing.assert_array_equal(fy2.to_numpy(), a[1])
    fixed = pkg.DroneMovingIBM(cfg, wrap_weight_quirk=quirk)
    ffx, ffy, _ = fixed.compute_force(_dev(ux), _dev(uy), *kin)
    np.testing.assert_allclose(a[0], ffx.to_numpy(), rtol=0, atol=5e-8)
    np.testing.assert_allclose(a[1], ffy.to_numpy(), rtol=0, atol=5e-8)


def test_sorted_spread_equals_the_fixed_point_spread(pkg, gpu):
    """ibm_spread_sorted (stencil points grouped by cell, one fixed-point sum
    per cell) fills the accumulators with the same 64-bit sums as ibm_spread
    in the fixed-point mode (one atomic per stencil point), bit for bit, on
    the markers and forces of a coupled step of the 600 x 200 x 50 case and
    on a dense body (the reference's airplane builder with its filled tip
    caps, ellipsoid fuselage and tails at 0.7 lu spacing, on a 128 x 64 x 48
    lattice: 5,521 markers), twice in a row on one work buffer."""
    import torch
    lib = load_sub("_lib")
    P = C.c_void_p
    dense = pkg.config.airplane_case("dense", 128, 64, 48, 20.0, 15.0, 40.0, (20.0, 3.0, 3.0),
                                     25.0, 8.0, 8.0, 6.0, 6.0, 2.0, 200.0, 10, "")
    for cfg in (pkg.config.case_c1(), dense):
        lbm = pkg.CLBMSolver(cfg)
        ibm = pkg.DroneMovingIBM(cfg, spread_mode="sorted")
        for n in range(1, 3):
            fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(n))
            lbm.step(fx, fy, fz)
        g = lib.ClbmGrid(cfg.Nx, cfg.Ny, cfg.Nz, lib.STORE_F32)
        marks = [P(t._t.data_ptr()) for t in (ibm.X_marker, ibm.Y_marker, ibm.Z_marker,
                                              ibm.FX_k, ibm.FY_k)]
        ipf = lib.IbmParams(float(ibm.ip.f_max), int(ibm.ip.wrap_weight_quirk), lib.SPREAD_FIXED, 0)
        for _ in range(2):
            a = torch.zeros(2, lbm.ncell, dtype=torch.int64, device=gpu)
            b = torch.zeros_like(a)
            lib.call("ibm_spread", C.byref(g), C.byref(ipf), ibm.N, *marks, P(a[0].data_ptr()),
                     P(a[1].data_ptr()), None)
            w = ibm._sort_work
            lib.call("ibm_spread_sorted", C.byref(g), C.byref(ibm.ip), ibm.N, *marks,
                     P(b[0].data_ptr()), P(b[1].data_ptr()), P(w.data_ptr()), w.numel(), None)
            assert torch.equal(a, b)
            assert int((a != 0).sum()) > 0
        # the work buffer's per-cell counters are back at zero
        ncell = lbm.ncell
        cnt = w[:4 * ncell].view(torch.int32)
        assert int(cnt.abs().sum()) == 0


def test_sorted_spread_in_a_coupled_run(pkg, gpu):
    """A coupled run with the sorted spread (the step reads its clamped
    fields) follows the fixed-point run to float32 rounding, repeats bit for
    bit, and a moving body leaves no stale force behind (the clamp clears the
    previous generation's box in this mode too)."""
    cfg = pkg.config.naca_case(96, 48, 8, 12.0, name="sorted")
    runs = []
    for mode in ("sorted", "sorted", "fixed"):
        lbm = pkg.CLBMSolver(cfg)
        ibm = pkg.DroneMovingIBM(cfg, spread_mode=mode)
        for n in range(1, 9):
            fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(n))
            lbm.step(fx, fy, fz)
            ibm.record_force()
        runs.append((lbm.f.to_numpy(), ibm.force_history(), ibm.fx.to_numpy(), ibm))
    (fa, ha, xa, ia), (fb, hb, xb, _), (fc, hc, xc, ic) = runs
    np.testing.assert_array_equal(fa, fb)
    np.testing.assert_array_equal(ha, hb)
    np.testing.assert_array_equal(xa, xb)
    np.testing.assert_allclose(fa, fc, rtol=0, atol=1e-6)
    np.testing.assert_allclose(ha, hc, rtol=0, atol=1e-5 * np.abs(hc).max())
    np.testing.assert_allclose(xa, xc, rtol=0, atol=1e-7)
    box = ia.force_box()
    inside = np.zeros(xa.shape, bool)
    inside[box[0]:box[3] + 1, box[1]:box[4] + 1, box[2]:box[5] + 1] = True
    assert not xa[~inside].any() and np.abs(xa[inside]).max() > 0


def test_sorted_spread_is_rejected_by_the_fused_entry_points(pkg, gpu):
    lib = load_sub("_lib")
    g = lib.ClbmGrid(8, 8, 8, 0)
    ip = lib.IbmParams(0.05, 1, lib.SPREAD_SORTED, 0)
    dummy = [C.c_void_p(1)] * 7
    with pytest.raises(lib.ClbmError, match="work buffer"):
        lib.call("ibm_spread", C.byref(g), C.byref(ip), 1, *dummy, None)
    with pytest.raises(lib.ClbmError, match="ibm_spread_sorted"):
        lib.call("ibm_compute_force", C.byref(g), C.byref(ip), None, C.c_void_p(1), 0,
                 C.c_void_p(1), C.c_void_p(1), 1, None, *([C.c_void_p(1)] * 3), 0.0, 0.0, 0.0,
                 0.0, 0.0, 0.0, *([C.c_void_p(1)] * 9), None, C.c_void_p(1), C.c_void_p(1),
                 C.c_void_p(1), 1, None)
    small = lib.fn("ibm_spread_sorted_bytes")(C.byref(g), 1) - 1
    with pytest.raises(lib.ClbmError, match="smaller"):
        lib.call("ibm_spread_sorted", C.byref(g), C.byref(ip), 1, *dummy, C.c_void_p(1), small,
                 None)
    ip = lib.IbmParams(0.05, 1, lib.SPREAD_FIXED, 0)
    with pytest.raises(lib.ClbmError, match="IBM_SPREAD_SORTED"):
        lib.call("ibm_spread_sorted", C.byref(g), C.byref(ip), 1, *dummy, C.c_void_p(1),
                 small + 1, None)


@pytest.mark.parametrize("spread", ["fixed", "sorted"])
def test_body_without_markers_leaves_the_flow_force_free(pkg, gpu, spread):
    """An immersed boundary with no markers: compute_force returns zero
    fields, the coupled loop equals the force-free step bit for bit, and the
    FP64 force sums are zero."""
    cfg = pkg.config.CaseConfig(name="empty", Nx=24, Ny=16, Nz=4, U_inf=0.05, Re=100, chord=4,
                                s_v=1.7, ux_in=0.05, uy_in=0.0,
                                X0=np.zeros(0, np.float32), Y0=np.zeros(0, np.float32),
                                Z0=np.zeros(0, np.float32), x0=8.0, y0=8.0)
    a = pkg.CLBMSolver(cfg)
    b = pkg.CLBMSolver(cfg)
    ibm = pkg.DroneMovingIBM(cfg, spread_mode=spread)
    assert ibm.N == 0
    for n in range(1, 5):
        fx, fy, fz = ibm.compute_force(a.ux, a.uy, *cfg.motion_state(n))
        a.step(fx, fy, fz)
        ibm.record_force()
        b.step()
    assert not ibm.fx.to_numpy().any() and not ibm.fy.to_numpy().any()
    np.testing.assert_array_equal(a.f.to_numpy(), b.f.to_numpy())
    assert not ibm.force_history().any() and ibm.force_history().shape == (4, 2)
