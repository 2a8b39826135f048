"""The driver counterpart and its on-disk format (main_moving.py:116-200):
frames ``frame_XXXXXX.npz`` {ux, uy, uz, rho} float32 (Nx, Ny, Nz), ``meta.npz``
with the reference's keys and the 24-entry ``grid_meta``, plus ``forces.npz``."""

import os

import numpy as np
import pytest

from conftest import load_sub

META_KEYS = {"xc", "yc", "theta", "step", "X_template", "Y_template", "X_htail_tmpl",
             "Y_htail_tmpl", "grid_meta"}


def test_grid_meta_layout(pkg):
    drv = load_sub("driver")
    cfg = pkg.config.naca_case(40, 24, 4, 8.0, name="meta")
    g = drv.grid_meta(cfg)
    assert g.dtype == np.float64 and g.shape == (24,)
    # main_moving.py:185-199: Nx, Ny, Nz, U_inf, Re, x0, y0, z0 first; vis_interval at 15
    assert tuple(g[:3]) == (40, 24, 4)
    assert g[3] == pytest.approx(0.05) and g[4] == pytest.approx(200.0)
    assert g[5] == pytest.approx(cfg.x0) and g[6] == pytest.approx(cfg.y0)
    assert g[15] == cfg.vis_interval


@pytest.mark.gpu
def test_driver_writes_reference_format_and_matches_oracle(pkg, oracle, gpu, tmp_path):
    drv = load_sub("driver")
    cfg = pkg.config.naca_case(40, 24, 4, 8.0, name="drv", N_steps=4).with_(vis_interval=2)
    res = drv.run(cfg, 4, str(tmp_path), verbose=False)
    assert not res["diverged"] and res["steps"] == 4
    files = sorted(os.listdir(tmp_path))
    assert files == ["forces.npz", "frame_000001.npz", "frame_000002.npz", "meta.npz"]
    fr = np.load(tmp_path / "frame_000002.npz")
    assert set(fr.files) == {"ux", "uy", "uz", "rho"}
    for k in fr.files:
        assert fr[k].dtype == np.float32 and fr[k].shape == (40, 24, 4)
    meta = np.load(tmp_path / "meta.npz")
    assert set(meta.files) == META_KEYS
    assert list(meta["step"]) == [2, 4] and meta["grid_meta"].shape == (24,)
    forces = np.load(tmp_path / "forces.npz")
    assert forces["sums"].shape == (4, 2) and forces["cl"].shape == (4,)
    # the frame after 4 steps against the FP64 oracle's coupled loop
    f, *_ = oracle.init_feq(cfg.Nx, cfg.Ny, cfg.Nz, cfg.ux_in, cfg.uy_in)
    f, hist = oracle.run_coupled(f, cfg.X0, cfg.Y0, cfg.Z0, 4, cfg.motion_state, cfg.s_v, cfg.s_e,
                                 cfg.ux_in, cfg.uy_in, f_max=cfg.f_max, dV=cfg.dV,
                                 dtype=np.float64)
    rho, ux, uy, uz = oracle.macro_post(f)
    np.testing.assert_allclose(fr["ux"], ux, rtol=0, atol=2e-6)
    np.testing.assert_allclose(fr["rho"], rho, rtol=0, atol=2e-6)
    np.testing.assert_allclose(forces["sums"], hist, rtol=0, atol=1e-4 * np.abs(hist).max())


@pytest.mark.parametrize("level", [None, 1, 6])
def test_save_frame_is_an_npz_np_load_reads(pkg, tmp_path, level):
    """save_frame writes what np.load returns unchanged, at every level, and
    level 6 gives the member layout of np.savez_compressed."""
    import zipfile
    drv = load_sub("driver")
    rng = np.random.default_rng(0)
    arrs = {k: rng.random((5, 4, 3), dtype=np.float32) for k in ("ux", "uy", "uz", "rho")}
    drv.save_frame(str(tmp_path / "f.npz"), level, **arrs)
    got = np.load(tmp_path / "f.npz")
    assert sorted(got.files) == sorted(arrs)
    for k, v in arrs.items():
        assert got[k].dtype == np.float32
        np.testing.assert_array_equal(got[k], v)
    np.savez_compressed(tmp_path / "g.npz", **arrs)
    names = lambda p: sorted(i.filename for i in zipfile.ZipFile(p).infolist())
    assert names(tmp_path / "f.npz") == names(tmp_path / "g.npz")
    kinds = {i.compress_type for i in zipfile.ZipFile(tmp_path / "f.npz").infolist()}
    assert kinds == {zipfile.ZIP_STORED if level is None else zipfile.ZIP_DEFLATED}


@pytest.mark.gpu
@pytest.mark.parametrize("compress", [1, False])
def test_async_frames_equal_the_synchronous_ones(pkg, gpu, tmp_path, compress):
    """Frames written through the pinned copy stream and worker threads hold
    the same arrays as frames read before the next step; six frames through
    three snapshot slots, so every slot is reused while steps go on."""
    drv = load_sub("driver")
    cfg = pkg.config.naca_case(40, 24, 4, 8.0, name="drv", N_steps=12).with_(vis_interval=2)
    a = drv.run(cfg, 12, str(tmp_path / "a"), verbose=False, async_frames=True, compress=compress)
    b = drv.run(cfg, 12, str(tmp_path / "b"), verbose=False, async_frames=False, compress=compress)
    assert a["steps"] == b["steps"] == 12
    names = sorted(os.listdir(tmp_path / "a"))
    assert names == sorted(os.listdir(tmp_path / "b"))
    assert sum(n.startswith("frame_") for n in names) == 6
    for n in names:
        fa, fb = np.load(tmp_path / "a" / n), np.load(tmp_path / "b" / n)
        assert fa.files == fb.files
        for k in fa.files:
            np.testing.assert_array_equal(fa[k], fb[k], err_msg=f"{n}:{k}")
    # consecutive frames differ: each holds its own step's fields
    f1, f2 = np.load(tmp_path / "a" / "frame_000001.npz"), np.load(tmp_path / "a" / "frame_000002.npz")
    assert not np.array_equal(f1["ux"], f2["ux"])


def test_main_maps_the_inlet_flag(pkg, monkeypatch):
    """--inlet reaches run(): "reference" as the default rule (True), and
    "hecht-harting" as itself."""
    drv = load_sub("driver")
    seen = []
    monkeypatch.setattr(drv, "run", lambda *a, **k: seen.append(k["inlet"]) or {"diverged": False})
    assert drv.main(["--case", "c0", "--steps", "1", "--no-frames"]) == 0
    assert drv.main(["--case", "c0", "--steps", "1", "--no-frames", "--inlet", "hecht-harting"]) == 0
    assert seen == [True, "hecht-harting"]


@pytest.mark.gpu
def test_driver_hecht_harting_inlet_is_the_solver_option(pkg, gpu, tmp_path):
    """run(inlet="hecht-harting") steps a CLBMSolver(inlet="hecht-harting"):
    its force history equals a hand loop with that solver, bit for bit, and
    differs from the reference rule's."""
    drv = load_sub("driver")
    cfg = pkg.config.naca_case(40, 24, 4, 8.0, le_x=3.0, name="drv-hh", N_steps=5)
    res = drv.run(cfg, 5, str(tmp_path), write_frames=False, verbose=False, inlet="hecht-harting")
    lbm = pkg.CLBMSolver(cfg, inlet="hecht-harting")
    ibm = pkg.DroneMovingIBM(cfg)
    for n in range(1, 6):
        fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(n))
        lbm.step(fx, fy, fz)
        ibm.record_force()
    np.testing.assert_array_equal(res["history"], ibm.force_history())
    ref = drv.run(cfg, 5, str(tmp_path / "ref"), write_frames=False, verbose=False)
    assert not np.array_equal(ref["history"], res["history"])


def test_frame_writer_depth_respects_the_pinned_budget(pkg):
    """The default number of snapshot slots: one per two cores (2 to 8),
    capped by the pinned host memory a frame needs (16 bytes a cell), at
    least one."""
    drv = load_sub("driver")
    d = drv.FrameWriter.default_depth
    assert d(600 * 200 * 50, cores=16, budget=1 << 30) == 8       # 96 MB frames
    assert d(600 * 200 * 50, cores=6, budget=1 << 30) == 3
    assert d(600 * 200 * 50, cores=1, budget=1 << 30) == 2
    assert d(800 * 260 * 220, cores=16, budget=1 << 30) == 1      # 732 MB frames
    assert d(800 * 260 * 220, cores=16, budget=4 << 30) == 5
    assert d(1200 * 400 * 100, cores=16, budget=1 << 30) == 1     # over budget: still one


@pytest.mark.gpu
@pytest.mark.parametrize("storage,deviation", [("f32", "weights"), ("f16c", "inflow"),
                                               ("f64", "weights")])
def test_checkpoint_resume_continues_bit_for_bit(pkg, gpu, tmp_path, storage, deviation):
    """save_state after 3 coupled steps, load_state into new objects, 4 more
    steps: populations, marker forces and force sums equal the uninterrupted
    run's bit for bit (the raw buffer keeps 16-bit codes and the AA state
    word); a checkpoint of another storage, grid or rate is refused."""
    cfg = pkg.config.naca_case(40, 24, 4, 8.0, name="ckpt")
    a_l = pkg.CLBMSolver(cfg, storage=storage, deviation=deviation)
    a_i = pkg.DroneMovingIBM(cfg)

    def steps(l, i, n0, n1):
        for n in range(n0, n1):
            fx, fy, fz = i.compute_force(l.ux, l.uy, *cfg.motion_state(n))
            l.step(fx, fy, fz)
            i.record_force()

    steps(a_l, a_i, 1, 4)
    path = str(tmp_path / "state.npz")
    a_l.save_state(path, force_history=a_i.force_history())
    steps(a_l, a_i, 4, 8)
    b_l = pkg.CLBMSolver(cfg, storage=storage, deviation=deviation)
    b_i = pkg.DroneMovingIBM(cfg)
    extra = b_l.load_state(path)
    assert b_l.steps == 3 and b_l.state_word == 1 | 2
    b_i.restore_history(extra["force_history"])
    steps(b_l, b_i, 4, 8)
    assert b_l.steps == a_l.steps == 7 and b_l.state_word == a_l.state_word
    np.testing.assert_array_equal(b_l.f.to_numpy(), a_l.f.to_numpy())
    np.testing.assert_array_equal(b_i.FX_k.to_numpy(), a_i.FX_k.to_numpy())
    np.testing.assert_array_equal(b_i.force_history(), a_i.force_history())
    other = "f32" if storage != "f32" else "f16"
    for wrong in (pkg.CLBMSolver(cfg, storage=other),
                  pkg.CLBMSolver(cfg.with_(Nz=6), storage=storage, deviation=deviation),
                  pkg.CLBMSolver(cfg.with_(s_v=1.7), storage=storage, deviation=deviation)):
        with pytest.raises(ValueError, match="checkpoint"):
            wrong.load_state(path)


@pytest.mark.gpu
def test_driver_resumes_from_its_checkpoint(pkg, gpu, tmp_path):
    """run(8 steps) against run(4 steps, checkpoint) + run(8 steps, resume):
    the same force history, the same frames and meta records."""
    drv = load_sub("driver")
    cfg = pkg.config.naca_case(40, 24, 4, 8.0, name="resume").with_(vis_interval=2)
    one, two = tmp_path / "one", tmp_path / "two"
    ref = drv.run(cfg, 8, str(one), verbose=False)
    ck = str(tmp_path / "run.npz")
    drv.run(cfg, 4, str(two), verbose=False, checkpoint=ck, checkpoint_every=3)
    res = drv.run(cfg, 8, str(two), verbose=False, checkpoint=ck, resume=ck)
    assert res["steps"] == 8 and not res["diverged"]
    np.testing.assert_array_equal(res["history"], ref["history"])
    assert sorted(os.listdir(two)) == sorted(os.listdir(one))
    for name in ("frame_000004.npz", "meta.npz", "forces.npz"):
        a, b = np.load(one / name), np.load(two / name)
        assert set(a.files) == set(b.files)
        for k in a.files:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"{name}:{k}")
