"""The committed golden vectors (tests/golden/*.npz, written by
tests/golden/make_golden.py from the FP64 oracle).

CPU: the oracle still reproduces them, so an edit to the oracle that changes
its results shows up here.  GPU: the HIP path is compared with them.

GPU tolerances: populations 2e-6 absolute for FP32 storage and 1e-12 for
FP64 storage (the FP64 path matches the FP64 oracle to rounding); marker
positions and stencil base indices exact; interpolated velocities and marker
forces 1e-6 absolute (FP32 weights, wave-order sums).  The coupled case: the
GPU IBM works in FP32 while the fixture's IBM ran in FP64, so the FP64-storage
run agrees to 1e-9 on populations and 1e-6 relative on the force sums, and
the FP32-storage run to 2e-6 and 1e-4 relative.
"""

import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name))


# ---------------------------------------------------------------------------
# CPU: the oracle against the fixtures
# ---------------------------------------------------------------------------
def test_oracle_reproduces_lbm_steps(oracle):
    g = _load("lbm_step_small.npz")
    s_v, s_e, ux_in, uy_in = g["params"]
    f = g["f0"].copy()
    for n in range(3):
        f, *_ = oracle.lbm_step(f, g["fx"], g["fy"], g["fz"], s_v, s_e, ux_in, uy_in)
        np.testing.assert_allclose(f, g["f_steps"][n], rtol=0, atol=1e-15)


def test_oracle_reproduces_stencil_edges(oracle):
    g = _load("stencil_edges.npz")
    nx, ny, nz = (int(v) for v in g["grid"])
    for i, kin in enumerate(g["kin"]):
        X, Y, Z, VX, VY = oracle.update_markers(g["X0"], g["Y0"], g["Z0"], *kin)
        np.testing.assert_array_equal(X, g[f"X_{i}"])
        np.testing.assert_array_equal(Y, g[f"Y_{i}"])
        base, *_, phx, phy, phz = oracle.stencil(X, Y, Z, nx, ny, nz, True)
        np.testing.assert_array_equal(base, g[f"base_{i}"])
        np.testing.assert_array_equal(phz, g[f"phz_{i}"])
        U, V = oracle.interpolate(g["ux"], g["uy"], X, Y, Z, True)
        np.testing.assert_array_equal(U, g[f"U_{i}"])


def test_stencil_fixture_covers_the_edges():
    """The fixture holds what it claims: points one ulp either side of a
    cell face, markers on both z seam layers, stencils clamped at both x ends."""
    g = _load("stencil_edges.npz")
    X, Z, base = g["X_0"], g["Z_0"], g["base_0"]
    assert np.any(X == np.float32(5.0))
    assert np.any(X == np.nextafter(np.float32(5.0), np.float32(9)))
    assert np.any(X == np.nextafter(np.float32(5.0), np.float32(0)))
    assert np.any(Z == 0.0) and np.any(Z == 5.0)
    assert base[:, 0].min() < 0 and base[:, 0].max() + 3 > 15
    # the floor rule separates x = 5 from the float32 just below it
    just_below = X == np.nextafter(np.float32(5.0), np.float32(0))
    at = X == np.float32(5.0)
    assert set(base[just_below, 0]) == {3} and set(base[at, 0]) == {4}


# ---------------------------------------------------------------------------
# GPU: the HIP path against the fixtures
# ---------------------------------------------------------------------------
def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).transpose(2, 1, 0))).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("storage,atol", [("f32", 2e-6), ("f64", 1e-12)])
def test_gpu_lbm_steps_match_fixture(pkg, gpu, storage, atol):
    g = _load("lbm_step_small.npz")
    s_v, s_e, ux_in, uy_in = (float(v) for v in g["params"])
    nx, ny, nz = g["f0"].shape[:3]
    cfg = pkg.config.CaseConfig(name="gold", Nx=nx, Ny=ny, Nz=nz, U_inf=0.05, Re=200, chord=10,
                                s_v=s_v, s_e=s_e, ux_in=ux_in, uy_in=uy_in)
    s = pkg.CLBMSolver(cfg, storage=storage)
    s.f.from_numpy(g["f0"])
    F = [_dev(g[k]) for k in ("fx", "fy", "fz")]
    for n in range(3):
        s.step(*F)
        np.testing.assert_allclose(s.f.to_numpy(), g["f_steps"][n], rtol=0, atol=atol,
                                   err_msg=f"step {n + 1}")
    mtol = max(10 * atol, 2e-7)
    np.testing.assert_allclose(s.ux.to_numpy(), g["ux"], rtol=0, atol=mtol)
    np.testing.assert_allclose(s.rho.to_numpy(), g["rho"], rtol=0, atol=mtol)


@pytest.mark.gpu
def test_gpu_stencil_edges_match_fixture(pkg, gpu):
    g = _load("stencil_edges.npz")
    nx, ny, nz = (int(v) for v in g["grid"])
    cfg = pkg.config.CaseConfig(name="edges", Nx=nx, Ny=ny, Nz=nz, U_inf=0.05, Re=100, chord=4,
                                s_v=1.5, X0=g["X0"], Y0=g["Y0"], Z0=g["Z0"], x0=8.0, y0=6.0)
    ibm = pkg.DroneMovingIBM(cfg, record_stencil=True)
    ux, uy = _dev(g["ux"]), _dev(g["uy"])
    for i, kin in enumerate(g["kin"]):
        ibm.compute_force(ux, uy, *(float(v) for v in kin))
        np.testing.assert_array_equal(ibm.X_marker.to_numpy(), g[f"X_{i}"])
        np.testing.assert_array_equal(ibm.Y_marker.to_numpy(), g[f"Y_{i}"])
        np.testing.assert_array_equal(ibm.Z_marker.to_numpy(), g[f"Z_{i}"])
        np.testing.assert_array_equal(ibm.stencil_base.cpu().numpy(), g[f"base_{i}"])
        for name, ref in (("U_k", "U"), ("V_k", "V"), ("FX_k", "FX"), ("FY_k", "FY")):
            np.testing.assert_allclose(getattr(ibm, name).to_numpy(), g[f"{ref}_{i}"], rtol=0,
                                       atol=1e-6, err_msg=f"{name} kinematics {i}")


@pytest.mark.gpu
@pytest.mark.parametrize("storage,ftol,htol", [("f64", 1e-9, 1e-6), ("f32", 2e-6, 1e-4)])
def test_gpu_coupled_c0_matches_fixture(pkg, gpu, storage, ftol, htol):
    g = _load("coupled_c0.npz")
    cfg = pkg.config.case_c0()
    lbm = pkg.CLBMSolver(cfg, storage=storage)
    ibm = pkg.DroneMovingIBM(cfg)
    hist = []
    for n in range(1, int(g["steps"]) + 1):
        fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(n))
        hist.append(ibm.force_sums().cpu().numpy().copy())
        lbm.step(fx, fy, fz)
    hist = np.array(hist)
    scale = np.abs(g["hist"]).max()
    np.testing.assert_allclose(hist, g["hist"], rtol=0, atol=htol * scale)
    f = lbm.f.to_numpy().reshape(-1, 19)
    np.testing.assert_allclose(f[g["idx"]], g["f_idx"], rtol=0, atol=ftol)
    np.testing.assert_allclose(f.astype(np.float64).sum(axis=0), g["qsum"], rtol=1e-9 if storage == "f64" else 1e-6)
