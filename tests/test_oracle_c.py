"""The C restatement (oracle/clbm_ref.c, also the CPU baseline) against the
NumPy oracle: the lattice step is bit-for-bit the same float32 arithmetic;
the coupled step differs only in the order of the spreading sums."""

import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_sub


@pytest.fixture(scope="module")
def ref_c():
    so = os.path.join(ROOT, "oracle", "libclbm_ref.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       stdout=subprocess.DEVNULL)
    import clbm_ref_c
    return clbm_ref_c


def _case():
    return load_sub("config").naca_case(40, 24, 6, 10.0, le_x=10.0, N_steps=5, name="t")


def test_lbm_step_bit_exact(oracle, ref_c):
    cfg = _case()
    sim = ref_c.RefSim(cfg, threads=2)
    rng = np.random.default_rng(0)
    f, *_ = oracle.init_feq(cfg.Nx, cfg.Ny, cfg.Nz, cfg.ux_in, cfg.uy_in, dtype=np.float32)
    f = (f * (1 + 0.02 * rng.standard_normal(f.shape))).astype(np.float32)
    sim.set_f(f)
    # the C step reads the force fields of its last compute_force: take them
    sim.step_coupled(1)
    fx, fy = sim.forces()
    sim.set_f(f)
    import clbm_ref_c  # noqa: F401
    lib = ref_c._load()
    lib.ref_step(sim.h)
    g, *_ = oracle.lbm_step(f, fx, fy, np.zeros_like(fx), np.float32(cfg.s_v), np.float32(cfg.s_e),
                            cfg.ux_in, cfg.uy_in)
    assert np.array_equal(sim.f(), g)


def test_hecht_harting_lbm_step_bit_exact(oracle, ref_c):
    """The C restatement's opt-in Hecht-Harting inlet is the NumPy oracle's
    float32 arithmetic, bit for bit (and differs from the reference's rule)."""
    cfg = _case()
    sim = ref_c.RefSim(cfg, threads=2, inlet="hecht-harting")
    rng = np.random.default_rng(4)
    f, *_ = oracle.init_feq(cfg.Nx, cfg.Ny, cfg.Nz, cfg.ux_in, cfg.uy_in, dtype=np.float32)
    f = (f * (1 + 0.02 * rng.standard_normal(f.shape))).astype(np.float32)
    sim.set_f(f)
    ref_c._load().ref_step(sim.h)
    z = np.zeros(f.shape[:3], np.float32)
    g, *_ = oracle.lbm_step(f, z, z, z, np.float32(cfg.s_v), np.float32(cfg.s_e), cfg.ux_in,
                            cfg.uy_in, inlet="hecht-harting")
    assert np.array_equal(sim.f(), g)
    h, *_ = oracle.lbm_step(f, z, z, z, np.float32(cfg.s_v), np.float32(cfg.s_e), cfg.ux_in,
                            cfg.uy_in)
    assert not np.array_equal(sim.f(), h)


def test_init_and_coupled_steps_close(oracle, ref_c):
    cfg = _case()
    sim = ref_c.RefSim(cfg, threads=1)
    f, rho, ux, uy, uz = oracle.init_feq(cfg.Nx, cfg.Ny, cfg.Nz, cfg.ux_in, cfg.uy_in, dtype=np.float32)
    assert np.array_equal(sim.f(), f)
    for n in range(1, 4):
        kin = cfg.motion_state(n)
        r = oracle.compute_force(ux, uy, cfg.X0, cfg.Y0, cfg.Z0, *kin, acc_dtype=np.float32)
        sim.step_coupled(n)
        m = sim.markers()
        assert np.array_equal(m["X"], r["X"]) and np.array_equal(m["Y"], r["Y"])
        np.testing.assert_allclose(m["U"], r["U"], rtol=0, atol=1e-7)
        f, rho, ux, uy, uz = oracle.lbm_step(f, r["fx"], r["fy"], r["fz"], np.float32(cfg.s_v),
                                             np.float32(cfg.s_e), cfg.ux_in, cfg.uy_in)
        np.testing.assert_allclose(sim.f(), f, rtol=0, atol=5e-7)


def test_force_history_and_threads_agree(oracle, ref_c):
    cfg = _case()
    a = ref_c.RefSim(cfg, threads=1).run(4, record=True)
    b = ref_c.RefSim(cfg, threads=4).run(4, record=True)
    assert a.shape == (4, 2)
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-9)
