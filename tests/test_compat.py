"""The import-compatible modules in compat/: the reference drivers'
``from clbm_moving import CLBMSolver`` / ``from airfoil_moving import
DroneMovingIBM`` (main_moving.py:55-56), the fish pair (fish/main_fish.py:51-52)
and ``ti.init`` (main_moving.py:28-37) resolve to the HIP-backed objects,
built from whatever module the driver imported as its config."""

import os
import sys

import numpy as np
import pytest

from conftest import ROOT, PKG_NAME

COMPAT = os.path.join(ROOT, PKG_NAME, "compat")


@pytest.fixture()
def compat(monkeypatch, pkg):
    monkeypatch.syspath_prepend(COMPAT)
    for m in ("_pkg", "clbm_moving", "airfoil_moving", "clbm_fish", "ibm_fish", "taichi",
              "config_moving", "config_fish"):
        monkeypatch.delitem(sys.modules, m, raising=False)
    cfg = pkg.config.naca_case(40, 24, 4, 8.0, name="compat")
    monkeypatch.setitem(sys.modules, "config_moving", pkg.config.as_module(cfg, "config_moving"))
    monkeypatch.setitem(sys.modules, "config_fish", pkg.config.as_module(cfg, "config_fish"))
    return cfg


def test_compat_modules_resolve_the_aliased_config(compat, pkg):
    import _pkg
    import airfoil_moving
    import clbm_fish
    import clbm_moving
    import ibm_fish
    got = _pkg.config_from("config_moving")
    assert (got.Nx, got.Ny, got.Nz) == (40, 24, 4)
    np.testing.assert_array_equal(got.X0, compat.X0)
    assert got.s_v == pytest.approx(compat.s_v)
    assert issubclass(clbm_moving.CLBMSolver, pkg.CLBMSolver)
    assert issubclass(airfoil_moving.DroneMovingIBM, pkg.DroneMovingIBM)
    assert issubclass(clbm_fish.CLBMSolver, pkg.CLBMSolver)
    assert issubclass(ibm_fish.FishIBM, pkg.FishIBM)


def test_taichi_stand_in_accepts_the_drivers_init(compat):
    import taichi as ti
    ti.init(arch=ti.gpu, default_fp=ti.f32, default_ip=ti.i32, fast_math=True)
    ti.init(arch=ti.cpu, default_fp=ti.f32, default_ip=ti.i32, cpu_max_num_threads=8)
    assert ti._settings["arch"] == ti.cpu


@pytest.mark.gpu
def test_compat_objects_step_like_the_package(compat, pkg, gpu):
    """Three reference-driver steps through the compat classes equal the same
    steps through the package classes, bit for bit."""
    import airfoil_moving
    import clbm_moving
    a_lbm, a_ibm = clbm_moving.CLBMSolver(), airfoil_moving.DroneMovingIBM()
    b_lbm, b_ibm = pkg.CLBMSolver(compat), pkg.DroneMovingIBM(compat)
    for n in range(1, 4):
        kin = compat.motion_state(n)
        fa = a_ibm.compute_force(a_lbm.ux, a_lbm.uy, *kin)
        a_lbm.step(*fa)
        fb = b_ibm.compute_force(b_lbm.ux, b_lbm.uy, *kin)
        b_lbm.step(*fb)
    for name in ("rho", "ux", "uy", "uz"):
        np.testing.assert_array_equal(getattr(a_lbm, name).to_numpy(),
                                      getattr(b_lbm, name).to_numpy())
    np.testing.assert_array_equal(a_ibm.fx.to_numpy(), b_ibm.fx.to_numpy())
