"""The region a split step leaves to the marker kernels.

CPU: the host prediction holds every stencil cell of the oracle's float32
marker placement over many kinematics (including markers on cell faces and
near the y and x boundaries); the C-ABI normalisation (host-only) keeps the
outlet pair x = Nx-2, Nx-1 together.  GPU: the two parts of a split step
equal one full step bit for bit, a coupled run with the marker kernels on
their own stream equals the same run on one stream bit for bit, and that
stream (the library's one per device) runs beside the solver's stream rather
than behind it on one hardware queue.
"""

import ctypes as C

import numpy as np
import pytest

from conftest import load_sub


def _template(pkg, chord=80.0):
    X, Y, ds = pkg.config.naca0012_arclength(chord, 1.0, 0.25)
    return X, Y


def test_hull_extremes_under_rotation(pkg):
    X, Y = _template(pkg)
    h = pkg.config.convex_hull_2d(X, Y)
    rng = np.random.default_rng(3)
    for th in rng.uniform(-np.pi, np.pi, 50):
        R = X.astype(np.float64) * np.cos(th) - Y.astype(np.float64) * np.sin(th)
        assert R.max() == R[h].max() and R.min() == R[h].min()


@pytest.mark.parametrize("seed", range(4))
def test_predicted_region_holds_every_stencil_cell(pkg, oracle, seed):
    am = load_sub("airfoil_moving")
    X0, Y0 = _template(pkg)
    nx, ny, nz = 600, 200, 3
    Z0 = np.full(len(X0), 1.0, np.float32)
    h = pkg.config.convex_hull_2d(X0, Y0)
    hx, hy = X0[h], Y0[h]
    zr = am.stencil_z_range(Z0, nz)
    rng = np.random.default_rng(seed)
    for _ in range(200):
        xc = float(rng.uniform(-10, nx + 10))
        yc = float(rng.uniform(-10, ny + 10))
        th = float(rng.uniform(-np.pi, np.pi))
        if rng.random() < 0.3:       # integer pivot and round angles
            xc, yc = float(np.round(xc)), float(np.round(yc))
        r = am.stencil_region(hx, hy, zr, (nx, ny, nz), xc, yc, th)
        X, Y, Z, *_ = oracle.update_markers(X0, Y0, Z0, xc, yc, th, 0, 0, 0)
        base, ii, jj, kk, *_ = oracle.stencil(X, Y, Z, nx, ny, nz)
        if r is None:       # the whole body is off the grid in x: no split, one stream
            assert X.max() < 0 or X.min() > nx - 1
            continue
        x0, y0, z0, x1, y1, z1 = r
        assert ii.min() >= x0 and ii.max() <= x1
        assert jj.min() >= y0 and jj.max() <= y1
        assert kk.min() >= z0 and kk.max() <= z1


@pytest.mark.parametrize("seed", range(3))
def test_library_region_equals_the_python_statement(pkg, seed):
    """ibm_step_region (host code of the library) gives the normalised box of
    stencil_region, or no split where that gives none or the box exceeds half
    the lattice, bit for bit over random, rounded, off-grid, seam-crossing and
    non-finite kinematics."""
    am = load_sub("airfoil_moving")
    lib = load_sub("_lib")
    L = lib.load()
    X0, Y0 = _template(pkg)
    h = pkg.config.convex_hull_2d(X0, Y0)
    hx = np.ascontiguousarray(X0[h], dtype=np.float32)
    hy = np.ascontiguousarray(Y0[h], dtype=np.float32)
    rng = np.random.default_rng(seed)
    for nx, ny, nz, inlet in ((600, 200, 50, 1), (240, 120, 8, 1), (96, 40, 4, 0)):
        grid = lib.ClbmGrid(nx, ny, nz, lib.STORE_F32)
        params = lib.ClbmParams(1.8, 1.0, 0.05, 0.0, inlet, 0)
        for zr in ((0, nz - 1), (1, nz - 2)):
            for i in range(150):
                xc = float(rng.uniform(-60, nx + 60))
                yc = float(rng.uniform(-20, ny + 20))
                th = float(rng.uniform(-np.pi, np.pi))
                if i % 5 == 0:
                    xc, yc = float(np.round(xc)), float(np.round(yc))
                if i == 7:
                    xc = float("nan")
                if i == 8:
                    th = float("inf")
                if i == 9:
                    xc = 1e30
                r = am.stencil_region(hx, hy, zr, (nx, ny, nz), xc, yc, th)
                want = None
                if r is not None:
                    raw = (C.c_int32 * 6)(*r)
                    norm = (C.c_int32 * 6)()
                    assert L.clbm_step_region(C.byref(grid), C.byref(params), raw, norm) == 0
                    vol = (norm[3] - norm[0] + 1) * (norm[4] - norm[1] + 1) * (norm[5] - norm[2] + 1)
                    want = None if 2 * vol > nx * ny * nz else list(norm)
                out = (C.c_int32 * 6)()
                rc = L.ibm_step_region(C.byref(grid), C.byref(params), hx.ctypes.data,
                                       hy.ctypes.data, len(hx), zr[0], zr[1], xc, yc, th, out)
                assert rc in (0, 1), L.clbm_last_error()
                assert (list(out) if rc == 1 else None) == want, (xc, yc, th, zr)


def test_z_range(pkg):
    am = load_sub("airfoil_moving")
    assert am.stencil_z_range(np.array([3.0, 4.5], np.float32), 10) == (2, 6)
    assert am.stencil_z_range(np.array([0.0, 4.5], np.float32), 10) == (0, 9)   # wraps
    assert am.stencil_z_range(np.array([7.0], np.float32), 10) == (6, 9)
    assert am.stencil_z_range(np.array([7.5, 8.0], np.float32), 10) == (0, 9)


def _region_call(lib, grid, params, r):
    raw = (C.c_int32 * 6)(*r)
    out = (C.c_int32 * 6)()
    rc = lib.load().clbm_step_region(C.byref(grid), C.byref(params), raw, out)
    return rc, list(out)


def test_region_normalisation_keeps_outlet_pair(pkg):
    lib = load_sub("_lib")
    grid = lib.ClbmGrid(40, 20, 6, lib.STORE_F32)
    inlet = lib.ClbmParams(1.6, 1.0, 0.05, 0.0, 1, 0)
    periodic = lib.ClbmParams(1.6, 1.0, 0.05, 0.0, 0, 0)
    assert _region_call(lib, grid, inlet, (5, 2, 0, 20, 10, 5)) == (0, [5, 2, 0, 20, 10, 5])
    assert _region_call(lib, grid, inlet, (-3, -1, 0, 50, 30, 9)) == (0, [0, 0, 0, 39, 19, 5])
    # reaching x = Nx-2 pulls in x = Nx-1; x = Nx-1 alone pulls in Nx-2
    assert _region_call(lib, grid, inlet, (30, 2, 0, 38, 10, 5)) == (0, [30, 2, 0, 39, 10, 5])
    assert _region_call(lib, grid, inlet, (39, 2, 0, 39, 10, 5)) == (0, [38, 2, 0, 39, 10, 5])
    assert _region_call(lib, grid, periodic, (39, 2, 0, 39, 10, 5)) == (0, [39, 2, 0, 39, 10, 5])
    assert _region_call(lib, grid, inlet, (10, 2, 0, 37, 10, 5))[1][3] == 37
    assert _region_call(lib, grid, inlet, (20, 2, 0, 10, 10, 5))[0] == -1       # lo > hi
    assert _region_call(lib, grid, inlet, (50, 2, 0, 60, 10, 5))[0] == -1       # outside


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).transpose(2, 1, 0))).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["f32", "f16", "f64"])
@pytest.mark.parametrize("region", [(4, 3, 1, 12, 9, 4), (0, 0, 0, 5, 11, 5), (14, 2, 0, 23, 7, 5),
                                    (9, 0, 2, 17, 11, 2)])
def test_split_step_equals_full_step(pkg, oracle, gpu, storage, region):
    import torch
    lib = load_sub("_lib")
    nx, ny, nz = 24, 12, 6
    cfg = pkg.config.CaseConfig(name="split", Nx=nx, Ny=ny, Nz=nz, U_inf=0.05, Re=200, chord=10,
                                s_v=1 / 0.56, ux_in=0.05, uy_in=-0.007)
    a = pkg.CLBMSolver(cfg, storage=storage)
    b = pkg.CLBMSolver(cfg, storage=storage)
    rng = np.random.default_rng(11)
    f, *_ = oracle.init_feq(nx, ny, nz, cfg.ux_in, cfg.uy_in)
    f = f * (1 + 0.02 * rng.standard_normal(f.shape))
    a.f.from_numpy(f)
    b.f.from_numpy(f)
    out = (C.c_int32 * 6)()
    assert lib.load().clbm_step_region(C.byref(a.grid), C.byref(a.params), (C.c_int32 * 6)(*region),
                                       out) == 0
    x0, y0, z0, x1, y1, z1 = list(out)
    F = np.zeros((nx, ny, nz), np.float32)
    F[x0:x1 + 1, y0:y1 + 1, z0:z1 + 1] = 1e-3 * rng.standard_normal(
        (x1 - x0 + 1, y1 - y0 + 1, z1 - z0 + 1))
    fx, fy = _dev(F), _dev(-0.5 * F)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for _ in range(4):
        a.step(fx, fy)
        for part in (0, 1):
            lib.call("clbm_step_part", C.byref(b.grid), C.byref(b.params),
                     C.c_void_p(b.state.data_ptr()), b.state_word, C.c_void_p(fx.data_ptr()),
                     C.c_void_p(fy.data_ptr()), None, None, out, part, st)
        b.state_word = lib.load().clbm_state_after_step(b.state_word, C.byref(b.params))
        b._macro_valid = False
    np.testing.assert_array_equal(a.f.to_numpy(), b.f.to_numpy())


@pytest.mark.gpu
def test_step_part_rejects_unnormalised_region(pkg, gpu):
    lib = load_sub("_lib")
    cfg = pkg.config.CaseConfig(name="split", Nx=24, Ny=12, Nz=6, U_inf=0.05, Re=200, chord=10,
                                s_v=1.6, ux_in=0.05, uy_in=0.0)
    a = pkg.CLBMSolver(cfg)
    bad = (C.c_int32 * 6)(20, 0, 0, 22, 5, 5)   # reaches Nx-2 without Nx-1
    with pytest.raises(lib.ClbmError, match="normalised"):
        lib.call("clbm_step_part", C.byref(a.grid), C.byref(a.params),
                 C.c_void_p(a.state.data_ptr()), a.state_word, None, None, None, None, bad, 0, None)


def _coupled(pkg, cfg, steps, overlap, spread="fixed", inlet=True, fast_host=True,
             storage="f32"):
    lbm = pkg.CLBMSolver(cfg, inlet=inlet, storage=storage,
                         deviation="inflow" if storage in ("f16", "bf16", "f16c") else "weights")
    ibm = pkg.DroneMovingIBM(cfg, overlap=overlap, spread_mode=spread)
    ibm.fast_host = fast_host
    split = 0
    for n in range(1, steps + 1):
        fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(n))
        lbm.step(fx, fy, fz)
        split += lbm.last_step_kind in ("split", "fused")
        ibm.record_force()
    return lbm, ibm, split


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["f32", "f16c", "f64"])
@pytest.mark.parametrize("where", ["centre", "outlet", "y-seam"])
def test_overlapped_run_equals_single_stream_run(pkg, gpu, where, storage):
    nx, ny, nz = 64, 32, 4
    kw = dict(le_x=20.0)
    if where == "outlet":
        kw = dict(le_x=nx - 10.0)      # the trailing stencils reach x = Nx-1
    cfg = pkg.config.naca_case(nx, ny, nz, 12.0, name=where, **kw)
    if where == "y-seam":
        cfg = cfg.with_(y0=1.0)        # the body straddles the periodic y boundary
    a_lbm, a_ibm, split = _coupled(pkg, cfg, 12, True, storage=storage)
    b_lbm, b_ibm, nosplit = _coupled(pkg, cfg, 12, False, storage=storage)
    assert nosplit == 0 and split == 12
    for name in ("rho", "ux", "uy", "uz"):
        np.testing.assert_array_equal(getattr(a_lbm, name).to_numpy(),
                                      getattr(b_lbm, name).to_numpy(), err_msg=name)
    np.testing.assert_array_equal(a_ibm.force_history(), b_ibm.force_history())
    np.testing.assert_array_equal(a_ibm.FX_k.to_numpy(), b_ibm.FX_k.to_numpy())
    assert a_ibm.region_violations() == 0


@pytest.mark.gpu
@pytest.mark.parametrize("le_x", [20.0, 1.5])
def test_hecht_harting_coupled_run(pkg, oracle, gpu, le_x):
    """The opt-in Hecht-Harting inlet through the coupled loop: split and
    single-stream runs agree bit for bit, also when the stencils reach the
    inlet plane (le_x 1.5: the marker kernels read x = 0 through the rule),
    and both follow the FP64 oracle's coupled loop with the same inlet."""
    cfg = pkg.config.naca_case(64, 32, 4, 12.0, name="hh", le_x=le_x)
    a_lbm, a_ibm, split = _coupled(pkg, cfg, 6, True, inlet="hecht-harting")
    b_lbm, b_ibm, nosplit = _coupled(pkg, cfg, 6, False, inlet="hecht-harting")
    assert nosplit == 0 and split == 6
    np.testing.assert_array_equal(a_lbm.f.to_numpy(), b_lbm.f.to_numpy())
    np.testing.assert_array_equal(a_ibm.force_history(), b_ibm.force_history())
    assert a_ibm.region_violations() == 0
    f, *_ = oracle.init_feq(cfg.Nx, cfg.Ny, cfg.Nz, cfg.ux_in, cfg.uy_in)
    f, hist = oracle.run_coupled(f, cfg.X0, cfg.Y0, cfg.Z0, 6, cfg.motion_state, cfg.s_v, cfg.s_e,
                                 cfg.ux_in, cfg.uy_in, f_max=cfg.f_max, dV=cfg.dV,
                                 dtype=np.float64, inlet="hecht-harting")
    np.testing.assert_allclose(a_lbm.f.to_numpy(), f, rtol=0, atol=2e-6)
    np.testing.assert_allclose(a_ibm.force_history(), hist, rtol=0, atol=1e-4 * np.abs(hist).max())
    if le_x < 2:       # stencils (floor(X) - 1 .. floor(X) + 2) reach the inlet plane
        assert float(a_ibm.X_marker.to_numpy().min()) < 2.0


@pytest.mark.gpu
def test_overlapped_run_equals_single_stream_run_at_size(pkg, gpu):
    """A lattice large enough (1.6 M cells) that the outside part is still
    sweeping while the marker kernels and the region's part run beside it."""
    cfg = pkg.config.naca_case(320, 160, 32, 40.0, le_x=80.0, name="overlap-size")
    a_lbm, a_ibm, split = _coupled(pkg, cfg, 16, True)
    b_lbm, b_ibm, nosplit = _coupled(pkg, cfg, 16, False)
    assert nosplit == 0 and split == 16
    np.testing.assert_array_equal(a_lbm.f.to_numpy(), b_lbm.f.to_numpy())
    np.testing.assert_array_equal(a_ibm.force_history(), b_ibm.force_history())
    assert a_ibm.region_violations() == 0


def _fused_run(pkg, cfg, steps, fused, storage="f32"):
    lbm = pkg.CLBMSolver(cfg, storage=storage,
                         deviation="inflow" if storage in ("f16", "bf16", "f16c") else "weights")
    lbm.fused_split = fused
    ibm = pkg.DroneMovingIBM(cfg, overlap=True)
    one_launch = []
    for n in range(1, steps + 1):
        fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(n))
        lbm.step(fx, fy, fz)
        one_launch.append(lbm.last_step_kind == "fused")
    return lbm, ibm, one_launch


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["f32", "f16c", "f64"])
@pytest.mark.parametrize("where", ["centre", "outlet"])
def test_fused_split_step_equals_the_two_launch_split_step(pkg, gpu, storage, where):
    """The split step as one launch whose region waits on the device for the
    published forces (clbm_step_fused) equals the two launches around the
    join (clbm_step_split) bit for bit, and no wait runs out."""
    kw = dict(le_x=20.0) if where == "centre" else dict(le_x=64 - 10.0)
    cfg = pkg.config.naca_case(64, 32, 4, 12.0, name=f"fused-{where}", **kw)
    a_lbm, a_ibm, a_one = _fused_run(pkg, cfg, 12, True, storage)
    b_lbm, b_ibm, b_one = _fused_run(pkg, cfg, 12, False, storage)
    assert all(a_one) and not any(b_one)
    np.testing.assert_array_equal(a_lbm.f.to_numpy(), b_lbm.f.to_numpy())
    np.testing.assert_array_equal(a_ibm.FX_k.to_numpy(), b_ibm.FX_k.to_numpy())
    np.testing.assert_array_equal(a_ibm.fx.to_numpy(), b_ibm.fx.to_numpy())
    assert a_ibm.wait_expirations() == 0 and a_ibm.region_violations() == 0
    a_lbm.health()


@pytest.mark.gpu
def test_fused_split_step_equals_the_two_launch_split_step_at_size(pkg, gpu):
    """At 1.6 M cells, where the sweep is still running while the marker
    kernels finish beside it."""
    cfg = pkg.config.naca_case(320, 160, 32, 40.0, le_x=80.0, name="fused-size")
    a_lbm, a_ibm, a_one = _fused_run(pkg, cfg, 16, True)
    b_lbm, b_ibm, _ = _fused_run(pkg, cfg, 16, False)
    assert all(a_one)
    np.testing.assert_array_equal(a_lbm.f.to_numpy(), b_lbm.f.to_numpy())
    np.testing.assert_array_equal(a_ibm.FX_k.to_numpy(), b_ibm.FX_k.to_numpy())
    assert a_ibm.wait_expirations() == 0


@pytest.mark.gpu
def test_fused_split_step_needs_a_concurrent_marker_stream_or_a_join(pkg, gpu):
    """The one-launch split step is taken only while the marker stream is
    known to run beside the solver's (otherwise the two launches around the
    join), or, in the "joined" mode, after a join; all equal the two-launch
    run bit for bit."""
    cfg = pkg.config.naca_case(64, 32, 4, 12.0, le_x=20.0, name="fused-gate")
    lbm = pkg.CLBMSolver(cfg)
    ibm = pkg.DroneMovingIBM(cfg, overlap=True)
    kinds = []
    for n in range(1, 10):
        fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(n))
        if n == 4:
            ibm.side_concurrent = False     # as where kernels run one at a time
        if n == 7:
            lbm.fused_split = "joined"
        lbm.step(fx, fy, fz)
        kinds.append(lbm.last_step_kind)
    assert kinds == ["fused"] * 3 + ["split"] * 3 + ["fused"] * 3
    b_lbm, b_ibm, _ = _fused_run(pkg, cfg, 9, False)
    np.testing.assert_array_equal(lbm.f.to_numpy(), b_lbm.f.to_numpy())
    np.testing.assert_array_equal(ibm.FX_k.to_numpy(), b_ibm.FX_k.to_numpy())
    assert ibm.wait_expirations() == 0


@pytest.mark.gpu
def test_captured_split_steps_replay_as_the_eager_steps(pkg, gpu):
    """Two coupled split steps captured in a HIP graph (after an eager first
    step, which creates the marker stream) take the two launches around the
    join, with the fork recorded in the graph and no event bound to a launch,
    and one replay equals the same two steps run eagerly, bit for bit."""
    import torch
    cfg = pkg.config.naca_case(64, 32, 4, 12.0, le_x=20.0, name="captured")

    def first_step():
        lbm = pkg.CLBMSolver(cfg)
        ibm = pkg.DroneMovingIBM(cfg, overlap=True)
        fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(1))
        lbm.step(fx, fy, fz)
        return lbm, ibm

    a_lbm, a_ibm = first_step()
    for n in (2, 3):
        fx, fy, fz = a_ibm.compute_force(a_lbm.ux, a_lbm.uy, *cfg.motion_state(n))
        a_lbm.step(fx, fy, fz)
    b_lbm, b_ibm = first_step()
    torch.cuda.synchronize()
    graph, cap = torch.cuda.CUDAGraph(), torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    kinds = []
    with torch.cuda.stream(cap):
        with torch.cuda.graph(graph, stream=cap):
            for n in (2, 3):
                fx, fy, fz = b_ibm.compute_force(b_lbm.ux, b_lbm.uy, *cfg.motion_state(n))
                b_lbm.step(fx, fy, fz)
                kinds.append(b_lbm.last_step_kind)
    torch.cuda.current_stream().wait_stream(cap)
    graph.replay()
    torch.cuda.synchronize()
    assert kinds == ["split", "split"]
    np.testing.assert_array_equal(a_lbm.f.to_numpy(), b_lbm.f.to_numpy())
    np.testing.assert_array_equal(a_ibm.FX_k.to_numpy(), b_ibm.FX_k.to_numpy())


@pytest.mark.gpu
def test_fused_split_step_marks_a_wait_that_runs_out(pkg, gpu):
    """A one-launch split step asked for a generation that is never published
    (one ahead of the last, with 0.05 ms to wait) gives up its wait, marks the
    generation, and the health check raises; every kernel ends."""
    lib = load_sub("_lib")
    cfg = pkg.config.naca_case(64, 32, 4, 12.0, le_x=20.0, name="fused-late")
    lbm, ibm, one = _fused_run(pkg, cfg, 3, True)
    assert all(one)
    fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(4))
    ibm._join()
    fxp, fyp, _ = ibm._force_ptrs
    late = ibm.gen + 1
    rc = lib.load().clbm_step_fused(
        C.byref(lbm.grid), C.byref(lbm.params), C.c_void_p(lbm.state.data_ptr()),
        lbm.state_word, ibm._region_buf, fxp, fyp, ibm._box_ptr, late, 0.05,
        C.c_void_p(lib.stream_of(lbm.device)), None, None, None)
    assert rc == 0, lib.load().clbm_last_error()
    assert ibm.wait_expirations() == late
    with pytest.raises(lib.ClbmError, match="wait ran out"):
        ibm.check_region()
    assert lib.load().clbm_step_fused(
        C.byref(lbm.grid), C.byref(lbm.params), C.c_void_p(lbm.state.data_ptr()),
        lbm.state_word, ibm._region_buf, fxp, fyp, ibm._box_ptr, late, 0.0, None, None, None,
        None) < 0


@pytest.mark.gpu
def test_overlap_splits_the_benchmark_shape(pkg, gpu):
    """On a body small against the lattice every step takes the split path."""
    cfg = pkg.config.naca_case(96, 48, 4, 12.0, name="small")
    lbm, ibm, split = _coupled(pkg, cfg, 6, True)
    assert split == 6
    assert ibm.region_violations() == 0


@pytest.mark.gpu
def test_bodies_share_the_library_marker_stream(pkg, gpu):
    """Every body overlaps on the library's one marker stream per device,
    which passed the check that its kernels run beside (not behind) those of
    the solver's stream; the entry point hands out the same stream again."""
    lib = load_sub("_lib")
    cfg = pkg.config.naca_case(96, 48, 4, 12.0, name="small")
    _, a, _ = _coupled(pkg, cfg, 2, True)
    _, b, _ = _coupled(pkg, cfg, 2, True, fast_host=False)
    assert a._side_h.value and a._side_h.value == b._side_h.value
    assert a.side_concurrent is True and b.side_concurrent is True
    h, ok = C.c_void_p(), C.c_int(-1)
    so = lib.load()
    assert so.clbm_side_stream(None, C.byref(h), C.byref(ok)) == 0
    assert h.value == a._side_h.value and ok.value == 1
    assert so.clbm_side_stream(None, None, None) < 0
    # the check itself: the marker stream runs beside the null stream, a
    # stream never beside itself (the waiting wave runs out its 50 ms)
    ok.value = -1
    assert so.clbm_streams_concurrent(None, h, C.byref(ok)) == 0 and ok.value == 1
    assert so.clbm_streams_concurrent(h, h, C.byref(ok)) == 0 and ok.value == 0
    assert so.clbm_streams_concurrent(None, None, C.byref(ok)) == 0 and ok.value == 0
    assert so.clbm_streams_concurrent(None, h, None) < 0


@pytest.mark.gpu
def test_region_violation_stops_the_run_at_the_health_check(pkg, gpu, monkeypatch):
    """A predicted region that misses part of the stencils is reported by the
    run loop's health check (ClbmError naming the generation), with no
    synchronisation per step.  Here the marker kernel checks a region that is
    too small while the solver is kept off the split path, so that no outside
    part runs beside the marker kernels (the check itself is what is tested)."""
    am = load_sub("airfoil_moving")
    cm = load_sub("clbm_moving")
    cfg = pkg.config.naca_case(96, 48, 4, 12.0, name="violation")
    lbm = pkg.CLBMSolver(cfg)
    ibm = pkg.DroneMovingIBM(cfg, overlap=True)
    ibm.fast_host = False        # the calls one by one, through _region
    true_region = am.DroneMovingIBM._region

    def shrunk(self, solver, xc, yc, theta):
        r = true_region(self, solver, xc, yc, theta)
        if r is None:
            return None
        r = list(r)
        r[3] = r[0] + 2                 # three x planes: the stencils reach further
        return (C.c_int32 * 6)(*r)

    def one_ibm_split(self, fx, fy, fz):  # the IBM is remembered, no split runs
        if len(self._readers) == 1:
            self._split_ibms.add(next(iter(self._readers.values()))[0])
        return None

    monkeypatch.setattr(am.DroneMovingIBM, "_region", shrunk)
    monkeypatch.setattr(cm.CLBMSolver, "_split_source", one_ibm_split)
    for n in range(1, 4):
        fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(n))
        lbm.step(fx, fy, fz)
        assert lbm.last_step_kind == "full"          # one full step, one stream
    with pytest.raises(load_sub("_lib").ClbmError, match="generation 3"):
        lbm.health()
    assert ibm.region_violations() == 3


@pytest.mark.gpu
def test_health_passes_when_the_split_regions_hold(pkg, gpu):
    cfg = pkg.config.naca_case(96, 48, 4, 12.0, name="holds")
    lbm, ibm, split = _coupled(pkg, cfg, 5, True)
    assert split == 5 and ibm in lbm._split_ibms
    rho_min, finite = lbm.health()
    assert finite and rho_min > 0.9


@pytest.mark.gpu
@pytest.mark.parametrize("overlap", [True, False])
def test_one_call_host_path_equals_the_calls_one_by_one(pkg, gpu, overlap):
    """compute_force through ibm_compute_force_split and the split step
    through clbm_step_split equal the per-call host path (ibm_step_region,
    the fork, ibm_compute_force, ibm_clip_forces, clbm_step_part x 2), bit for
    bit, with and without the marker stream."""
    cfg = pkg.config.naca_case(96, 48, 4, 12.0, le_x=60.0, name="one-call")
    a_lbm, a_ibm, a_split = _coupled(pkg, cfg, 9, overlap, fast_host=True)
    b_lbm, b_ibm, b_split = _coupled(pkg, cfg, 9, overlap, fast_host=False)
    assert a_split == b_split == (9 if overlap else 0)
    np.testing.assert_array_equal(a_lbm.f.to_numpy(), b_lbm.f.to_numpy())
    np.testing.assert_array_equal(a_ibm.force_history(), b_ibm.force_history())
    np.testing.assert_array_equal(a_ibm.fx.to_numpy(), b_ibm.fx.to_numpy())
    assert a_ibm.region_violations() == 0 and a_ibm.gen == b_ibm.gen == 9


def _loop_seeing_forks(pkg, cfg, steps, overlap, storage, between=None):
    """A coupled loop as the benchmark runs it (no force record between
    steps); ``between(n, lbm, ibm)`` runs after step n.  Returns the solver,
    the body and, per compute_force after the first, the fork_bound argument
    it passed to ibm_compute_force_split."""
    lbm = pkg.CLBMSolver(cfg, storage=storage,
                         deviation="inflow" if storage in ("f16", "bf16", "f16c") else "weights")
    ibm = pkg.DroneMovingIBM(cfg, overlap=overlap)
    bound = []
    for n in range(1, steps + 1):
        fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(n))
        if n == 1:
            orig = ibm._split_fn

            def spy(*a, orig=orig):
                bound.append(a[10])
                return orig(*a)
            ibm._split_fn = spy
        lbm.step(fx, fy, fz)
        if between is not None:
            between(n, lbm, ibm)
    return lbm, ibm, bound


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["f32", "f16c", "f64"])
def test_bound_fork_run_equals_single_stream_run(pkg, gpu, storage):
    """With nothing queued between a step and the next compute_force, the
    body's fork event completes with the step's inside part and the marker
    stream waits for it with no record on the solver's stream: every call
    after the first takes that path, and the run equals the one-stream run
    bit for bit."""
    cfg = pkg.config.naca_case(64, 32, 4, 12.0, le_x=20.0, name="bound-fork")
    a_lbm, a_ibm, bound = _loop_seeing_forks(pkg, cfg, 12, True, storage)
    b_lbm, b_ibm, _ = _loop_seeing_forks(pkg, cfg, 12, False, storage)
    assert bound == [1] * 11
    np.testing.assert_array_equal(a_lbm.f.to_numpy(), b_lbm.f.to_numpy())
    np.testing.assert_array_equal(a_ibm.FX_k.to_numpy(), b_ibm.FX_k.to_numpy())
    np.testing.assert_array_equal(a_ibm.fx.to_numpy(), b_ibm.fx.to_numpy())
    assert a_ibm.region_violations() == 0


@pytest.mark.gpu
def test_recorded_force_sums_keep_the_fork_bound(pkg, gpu):
    """The driver's loop records the force sums after every step: they are
    queued on the marker stream, so the caller's stream gets no work between
    two steps and every compute_force after the first keeps the fork bound.
    Histories and states equal the one-stream run bit for bit, also when a
    step in between does not split (the body off the lattice), whose marker
    kernels on the caller's stream must wait for the sums still queued."""
    cfg = pkg.config.naca_case(64, 32, 4, 12.0, le_x=20.0, name="sums-side")

    def kin(n):
        xc, yc, th, vx, vy, om = cfg.motion_state(n)
        return (xc - 500.0 if n == 6 else xc), yc, th, vx, vy, om

    def run(overlap):
        lbm = pkg.CLBMSolver(cfg)
        ibm = pkg.DroneMovingIBM(cfg, overlap=overlap)
        bound = []
        for n in range(1, 11):
            fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *kin(n))
            if n == 1:
                orig = ibm._split_fn

                def spy(*a, orig=orig):
                    bound.append(a[10])
                    return orig(*a)
                ibm._split_fn = spy
            lbm.step(fx, fy, fz)
            ibm.record_force()
        return lbm, ibm, bound

    a_lbm, a_ibm, bound = run(True)
    b_lbm, b_ibm, _ = run(False)
    # the first record allocates the history (a join, on the caller's stream); step 6
    # does not split (its markers run on the caller's stream, after a join); the call
    # after each records the fork again
    assert bound == [0, 1, 1, 1, 0, 0, 1, 1, 1]
    np.testing.assert_array_equal(a_ibm.force_history(), b_ibm.force_history())
    np.testing.assert_array_equal(a_lbm.f.to_numpy(), b_lbm.f.to_numpy())
    assert a_ibm.region_violations() == 0 and a_ibm.wait_expirations() == 0


@pytest.mark.gpu
def test_work_between_steps_records_the_fork_again(pkg, gpu):
    """Reading or writing the body's fields, writing the populations, or a
    step without the body between a step and the next compute_force puts
    work on the solver's stream that the marker kernels must follow: that
    call records the fork event again (fork_bound 0), and the run equals the
    one-stream run bit for bit."""
    cfg = pkg.config.naca_case(64, 32, 4, 12.0, le_x=20.0, name="fork-again")

    def between(n, lbm, ibm):
        if n == 2:
            ibm.fx.to_numpy()                      # a read of the body's field
        elif n == 4:
            lbm.f.from_numpy(lbm.f.to_numpy())     # a write of the populations
        elif n == 6:
            ibm.fy.fill(0.0)                       # a write of the body's field
        elif n == 8:
            ibm.FX_k.to_numpy()                    # a read of a marker array
        elif n == 9:
            lbm.step()                             # a force-free step

    a_lbm, a_ibm, bound = _loop_seeing_forks(pkg, cfg, 12, True, "f32", between)
    b_lbm, b_ibm, _ = _loop_seeing_forks(pkg, cfg, 12, False, "f32", between)
    # bound[i] belongs to compute_force i + 2, right after step i + 1
    assert bound == [1, 0, 1, 0, 1, 0, 1, 0, 0, 1, 1]
    np.testing.assert_array_equal(a_lbm.f.to_numpy(), b_lbm.f.to_numpy())
    np.testing.assert_array_equal(a_ibm.FX_k.to_numpy(), b_ibm.FX_k.to_numpy())
    assert a_ibm.region_violations() == 0


@pytest.mark.gpu
def test_timed_outside_part_changes_nothing_and_times_the_kernel(pkg, gpu):
    """Timing events handed to a split step (the outside part is then
    launched with start/stop events of its own) leave the results bit for bit
    those of an untimed run, and give positive durations no longer than the
    step's stream time."""
    import torch
    cfg = pkg.config.naca_case(96, 48, 4, 12.0, le_x=60.0, name="timed")
    runs = []
    for timed in (False, True):
        lbm = pkg.CLBMSolver(cfg)
        ibm = pkg.DroneMovingIBM(cfg, overlap=True)
        evs = []
        for n in range(1, 8):
            fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(n))
            if timed and n % 2:
                e = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                for x in e:        # created now, so the step can hand them to the library
                    x.record()
                lbm.kernel_events = e
                evs.append(e)
            lbm.step(fx, fy, fz)
            ibm.record_force()
            assert lbm.last_step_kind in ("split", "fused")      # the split step ran
        torch.cuda.synchronize()
        runs.append((lbm, ibm, evs))
    (a, ai, _), (b, bi, evs) = runs
    np.testing.assert_array_equal(a.f.to_numpy(), b.f.to_numpy())
    np.testing.assert_array_equal(ai.force_history(), bi.force_history())
    assert len(evs) == 4 and all(0 < e0.elapsed_time(e1) < 50.0 for e0, e1 in evs)
