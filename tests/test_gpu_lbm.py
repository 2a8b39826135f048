"""GPU parity of the lattice-Boltzmann kernels against the CPU oracle.

Tolerances (stated per test): FP32 storage is compared with the FP64 oracle
at an absolute 2e-6 on populations (~1e-5 relative to the smallest lattice
weight 1/36) for up to ten steps; the FP64 storage path must match the FP64
oracle to 1e-12, which fixes the algebra of the closed-form collision.
"""

import numpy as np
import pytest

from conftest import load_sub

pytestmark = pytest.mark.gpu


def _case(pkg, nx, ny, nz, s_v=1 / 0.56, ux=0.05, uy=-0.007):
    cfg = pkg.config.CaseConfig(name="t", Nx=nx, Ny=ny, Nz=nz, U_inf=0.05, Re=200, chord=10,
                                s_v=s_v, ux_in=ux, uy_in=uy)
    return cfg


def _rand_state(oracle, shape, seed, amp=0.02, ux=0.05, uy=-0.007):
    rng = np.random.default_rng(seed)
    f, *_ = oracle.init_feq(*shape, ux, uy)
    return f * (1.0 + amp * rng.standard_normal(f.shape))


def _rand_force(shape, seed, amp=1e-3):
    """Force fields as the solver receives them (float32), widened for the oracle."""
    rng = np.random.default_rng(seed + 1000)
    return [(amp * rng.standard_normal(shape)).astype(np.float32).astype(np.float64)
            for _ in range(3)]


def _set_force(solver, arrs):
    import torch
    out = []
    for a in arrs:
        t = torch.from_numpy(np.ascontiguousarray(a.astype(np.float32).transpose(2, 1, 0))).cuda()
        out.append(t)
    return out


@pytest.mark.parametrize("storage,atol", [("f32", 2e-7), ("f64", 1e-15)])
def test_init_feq(pkg, oracle, gpu, storage, atol):
    cfg = _case(pkg, 7, 5, 4)
    s = pkg.CLBMSolver(cfg, storage=storage)
    f_ref, *_ = oracle.init_feq(7, 5, 4, cfg.ux_in, cfg.uy_in)
    np.testing.assert_allclose(s.f.to_numpy(), f_ref, rtol=0, atol=atol)


@pytest.mark.parametrize("storage,atol", [("f32", 2e-6), ("f64", 1e-12)])
@pytest.mark.parametrize("shape", [(9, 7, 5), (16, 12, 8), (4, 3, 3), (5, 1, 4), (6, 4, 1),
                                   (3, 1, 1), (67, 5, 3)])
def test_steps_with_force_match_oracle(pkg, oracle, gpu, storage, atol, shape):
    """Random post-stream state, random dense force, inlet + outlet: 1..6 steps
    (both AA parities) against clbm_oracle.lbm_step in FP64.  Shapes include
    periodic axes one cell thick (every y or z neighbour is the cell itself),
    the narrowest lattice the inlet/outlet planes allow (Nx = 3: inlet,
    one interior column, outlet) and rows that straddle 256-cell workgroups."""
    cfg = _case(pkg, *shape)
    s = pkg.CLBMSolver(cfg, storage=storage)
    f = _rand_state(oracle, shape, 1)
    F = _rand_force(shape, 1)
    s.f.from_numpy(f)
    tF = _set_force(s, F)
    for n in range(6):
        s.step(*tF)
        f, rho, ux, uy, uz = oracle.lbm_step(f, *F, cfg.s_v, cfg.s_e, cfg.ux_in, cfg.uy_in)
        assert s.parity == (n + 1) % 2
        np.testing.assert_allclose(s.f.to_numpy(), f, rtol=0, atol=atol, err_msg=f"step {n + 1}")
    # the macroscopic readout is float32 (as the reference's fields): rho ~ 1 carries 6e-8
    mtol = max(10 * atol, 2e-7)
    np.testing.assert_allclose(s.rho.to_numpy(), rho, rtol=0, atol=mtol)
    np.testing.assert_allclose(s.ux.to_numpy(), ux, rtol=0, atol=mtol)
    np.testing.assert_allclose(s.uy.to_numpy(), uy, rtol=0, atol=mtol)
    np.testing.assert_allclose(s.uz.to_numpy(), uz, rtol=0, atol=mtol)


@pytest.mark.parametrize("storage,atol", [("f32", 2e-6), ("f64", 1e-12)])
@pytest.mark.parametrize("shape", [(9, 7, 5), (16, 12, 8)])
def test_hecht_harting_inlet_matches_oracle(pkg, oracle, gpu, storage, atol, shape):
    """The opt-in Hecht-Harting inlet (no reference counterpart): 1..6 steps
    against clbm_oracle.lbm_step(inlet="hecht-harting"), and the x = 0 cells
    then carry exactly the momentum rho_in (ux_in, uy_in, 0) of their density."""
    cfg = _case(pkg, *shape)
    s = pkg.CLBMSolver(cfg, storage=storage, inlet="hecht-harting")
    f = _rand_state(oracle, shape, 2)
    F = _rand_force(shape, 2)
    s.f.from_numpy(f)
    tF = _set_force(s, F)
    for n in range(6):
        s.step(*tF)
        f, *_ = oracle.lbm_step(f, *F, cfg.s_v, cfg.s_e, cfg.ux_in, cfg.uy_in,
                                inlet="hecht-harting")
        np.testing.assert_allclose(s.f.to_numpy(), f, rtol=0, atol=atol, err_msg=f"step {n + 1}")
    g = s.f.to_numpy()[0].astype(np.float64)
    e = np.stack([oracle.EX, oracle.EY, oracle.EZ], 1).astype(np.float64)
    rho, m = g.sum(-1), g @ e
    mtol = 1e-13 if storage == "f64" else 2e-6
    np.testing.assert_allclose(m[..., 0], rho * cfg.ux_in, rtol=0, atol=mtol)
    np.testing.assert_allclose(m[..., 1], rho * cfg.uy_in, rtol=0, atol=mtol)
    np.testing.assert_allclose(m[..., 2], 0.0, rtol=0, atol=mtol)


@pytest.mark.parametrize("storage,atol", [("f32", 2e-6), ("f64", 1e-12)])
@pytest.mark.parametrize("dense", [True, False])
def test_outlet_rows_with_equal_and_unequal_force(pkg, oracle, gpu, storage, atol, dense):
    """The x = Nx-1 column collides the outlet copy of x = Nx-2.  Rows whose
    force equals that of their x = Nx-2 neighbour take one code path, the others
    another; a force where half the rows match (and, without ``dense``, only a
    force box well inside the grid) covers both against the oracle."""
    shape = (11, 6, 5)
    cfg = _case(pkg, *shape)
    s = pkg.CLBMSolver(cfg, storage=storage)
    f = _rand_state(oracle, shape, 3)
    F = _rand_force(shape, 3)
    if not dense:
        for a in F:
            a[:3] = 0.0
            a[8:] = 0.0
    else:
        rows = np.random.default_rng(5).random(shape[1:]) < 0.5
        for a in F:
            a[-1][rows] = a[-2][rows]
    s.f.from_numpy(f)
    tF = _set_force(s, F)
    for n in range(4):
        s.step(*tF)
        f, *_ = oracle.lbm_step(f, *F, cfg.s_v, cfg.s_e, cfg.ux_in, cfg.uy_in)
        np.testing.assert_allclose(s.f.to_numpy(), f, rtol=0, atol=atol, err_msg=f"step {n + 1}")


@pytest.mark.parametrize("storage,atol", [("f32", 2e-6), ("f64", 1e-12)])
def test_periodic_box_matches_oracle_and_conserves_mass(pkg, oracle, gpu, storage, atol):
    shape = (10, 6, 5)
    cfg = _case(pkg, *shape)
    s = pkg.CLBMSolver(cfg, storage=storage, inlet=False)
    f = _rand_state(oracle, shape, 2)
    s.f.from_numpy(f)
    m0 = f.sum()
    for _ in range(5):
        s.step()
        rho, ux, uy, uz = oracle.macro_with_force(f, 0 * f[..., 0], 0 * f[..., 0], 0 * f[..., 0])
        z = np.zeros(shape)
        f = oracle.stream(oracle.collide(f, rho, ux, uy, uz, z, z, z, cfg.s_v, cfg.s_e))
    got = s.f.to_numpy()
    np.testing.assert_allclose(got, f, rtol=0, atol=atol)
    assert abs(got.astype(np.float64).sum() - m0) < (1e-3 if storage == "f32" else 1e-10)


def test_init_state_is_fixed_point_of_rest(pkg, oracle, gpu):
    """rho = 1, u = 0 with no force and a periodic box: f = w is a fixed point."""
    cfg = _case(pkg, 8, 6, 4, ux=0.0, uy=0.0)
    s = pkg.CLBMSolver(cfg, storage="f64", inlet=False)
    f0 = s.f.to_numpy()
    for _ in range(4):
        s.step()
    np.testing.assert_allclose(s.f.to_numpy(), f0, rtol=0, atol=1e-15)
    np.testing.assert_allclose(f0[0, 0, 0], oracle.W, rtol=0, atol=1e-16)


MANTISSA = {"f16": 10, "bf16": 7, "f16c": 11}


@pytest.mark.parametrize("deviation", ["weights", "inflow"])
@pytest.mark.parametrize("storage", ["f16", "bf16", "f16c"])
def test_16bit_storage_tracks_fp64(pkg, oracle, gpu, storage, deviation):
    """Deviation storage: one rounding of f - base per step; after 6 steps the
    populations stay within the format's resolution (4 roundings of the
    largest deviation) of the FP64 result."""
    shape = (12, 8, 6)
    cfg = _case(pkg, *shape)
    s = pkg.CLBMSolver(cfg, storage=storage, deviation=deviation)
    f = _rand_state(oracle, shape, 3, amp=0.005)
    F = _rand_force(shape, 3, amp=2e-4)
    s.f.from_numpy(f)
    tF = _set_force(s, F)
    for _ in range(6):
        s.step(*tF)
        f, *_ = oracle.lbm_step(f, *F, cfg.s_v, cfg.s_e, cfg.ux_in, cfg.uy_in)
    base = oracle.W if deviation == "weights" else oracle.init_feq(1, 1, 1, cfg.ux_in, cfg.uy_in)[0][0, 0, 0]
    dev = np.abs(f - base).max()
    tol = dev * 2.0 ** -MANTISSA[storage] * 4
    np.testing.assert_allclose(s.f.to_numpy(), f, rtol=0, atol=tol)


@pytest.mark.parametrize("deviation", ["weights", "inflow"])
def test_f16c_import_export_round_trip_is_the_codec(pkg, oracle, gpu, deviation):
    """clbm_import_state + clbm_export_state of the 1-4-11 format equal the
    oracle's codec on f - base, bit for bit (base: float32 equilibrium at
    the deviation velocity, rounded once from float64)."""
    shape = (10, 6, 4)
    cfg = _case(pkg, *shape)
    s = pkg.CLBMSolver(cfg, storage="f16c", deviation=deviation)
    f = _rand_state(oracle, shape, 9, amp=0.05).astype(np.float32)
    f[0, 0, 0, 3] = np.float32(np.nan)
    s.f.from_numpy(f)
    # the base as the library forms it: float32 grid velocity, second-order
    # equilibrium at rho = 1 in float64, rounded once
    u, v = ((float(np.float32(cfg.ux_in)), float(np.float32(cfg.uy_in))) if deviation == "inflow"
            else (0.0, 0.0))
    eu = oracle.EX * u + oracle.EY * v
    base = (oracle.W * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * (u * u + v * v))).astype(np.float32)
    want = base + oracle.f16c_decode(oracle.f16c_encode(f - base))
    got = s.f.to_numpy()
    np.testing.assert_array_equal(got, want)
    assert np.isnan(got[0, 0, 0, 3])


def test_health_reports_min_and_nonfinite(pkg, oracle, gpu):
    shape = (8, 6, 4)
    cfg = _case(pkg, *shape)
    s = pkg.CLBMSolver(cfg)
    f = _rand_state(oracle, shape, 4)
    s.f.from_numpy(f)
    rmin, fin = s.health()
    assert fin
    assert abs(rmin - f.sum(-1).min()) < 1e-6
    f[3, 2, 1, 7] = np.nan
    s.f.from_numpy(f)
    _, fin = s.health()
    assert not fin
    f[3, 2, 1, 7] = 0.0
    f[1, 1, 1, :] *= 0.1
    s.f.from_numpy(f)
    rmin, fin = s.health()
    assert fin and abs(rmin - f.sum(-1).min()) < 1e-6


def test_step_rejects_bad_arguments(pkg, gpu):
    lib = load_sub("_lib")
    import ctypes as C
    g = lib.ClbmGrid(2, 4, 4, 0)
    p = lib.ClbmParams(1.5, 1.0, 0.05, 0.0, 1, 0)
    with pytest.raises(lib.ClbmError, match="nx >= 3"):
        lib.call("clbm_step", C.byref(g), C.byref(p), C.c_void_p(1), 0, None, None, None, None, None)
    g = lib.ClbmGrid(8, 4, 4, 0)
    with pytest.raises(lib.ClbmError, match="unknown bits"):
        lib.call("clbm_step", C.byref(g), C.byref(p), C.c_void_p(1), 4, None, None, None, None, None)
    with pytest.raises(lib.ClbmError, match="needs the params"):
        lib.call("clbm_macro", C.byref(g), None, C.c_void_p(1), lib.STATE_PLANES, None, None, None,
                 None, None)


@pytest.mark.parametrize("parity", [0, 1])
def test_stream_probe_moves_the_step_slots(pkg, oracle, gpu, parity):
    """clbm_stream_probe (the measurement probe) loads each cell's slots and
    stores them where a step from that parity stores: parity 0 new[opp q][c] =
    old[q][c]; parity 1 new[q][c + e_q] = old[opp q][c - e_q] (periodic), i.e.
    new[q] = old[opp q] shifted by 2 e_q.  Exact."""
    import ctypes as C
    import torch
    lib = load_sub("_lib")
    nx, ny, nz = 9, 7, 5
    s = pkg.CLBMSolver(_case(pkg, nx, ny, nz))
    old = np.random.default_rng(3).random((19, nz, ny, nx), dtype=np.float32)
    s.state.copy_(torch.from_numpy(old.reshape(-1).view(np.uint8)).cuda())
    lib.call("clbm_stream_probe", C.byref(s.grid), C.c_void_p(s.state.data_ptr()), parity, None)
    torch.cuda.synchronize()
    new = s.state.cpu().numpy().view(np.float32).reshape(19, nz, ny, nx)
    e = np.stack([oracle.EX, oracle.EY, oracle.EZ], 1).astype(int)
    opp = [int(np.flatnonzero((e == -e[q]).all(1))[0]) for q in range(19)]
    exp = np.empty_like(old)
    for q in range(19):
        if parity == 0:
            exp[opp[q]] = old[q]
        else:
            exp[q] = np.roll(old[opp[q]], shift=(2 * e[q, 2], 2 * e[q, 1], 2 * e[q, 0]),
                             axis=(0, 1, 2))
    np.testing.assert_array_equal(new, exp)
