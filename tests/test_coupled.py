"""clbm_ibm_step: compute_force + step as one launch and a tail on one stream.

GPU: its populations, marker arrays and force box equal those of
ibm_compute_force + ibm_clip_forces + clbm_step (the package's single-stream
coupled loop), bit for bit, with the region stepped in the tail kernel, at
the end of the main launch after the bounded wait, and through the give-up
path (the inside workgroups step nothing, the tail steps the region); at
96 x 48 x 8 and at 600 x 200 x 50; f32, f16c and f64.  The fields fx, fy are
not written by it: ibm_spread + ibm_clip_forces rebuild them from the
markers, within 1e-8 of the clamp's fields.  CPU: argument
errors.
"""

import ctypes as C

import numpy as np
import pytest

from conftest import load_sub


def _lib():
    return load_sub("_lib")


def coupled_step(lbm, ibm, cfg, n, interleave=4, spread=4, spins=100000):
    """One clbm_ibm_step of (lbm, ibm) at step n of cfg's motion."""
    load_sub("coupled").step_coupled(lbm, ibm, *cfg.motion_state(n), interleave=interleave,
                                     inside_spread=spread, wait_spins=spins)


def _pair(pkg, cfg, storage):
    a = (pkg.CLBMSolver(cfg, storage=storage, deviation="inflow"),
         pkg.DroneMovingIBM(cfg, overlap=False))
    b = (pkg.CLBMSolver(cfg, storage=storage, deviation="inflow"),
         pkg.DroneMovingIBM(cfg, overlap=False))
    return a, b


MODES = {"tail": dict(spins=0), "in-launch": dict(spins=100000, spread=4),
         "give-up": dict(spins=-1, spread=2)}


@pytest.mark.gpu
@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("storage", ["f32", "f16c", "f64"])
def test_coupled_step_equals_the_separate_calls(pkg, gpu, storage, mode):
    cfg = pkg.config.naca_case(96, 48, 8, 12.0, name="coupled")
    (la, ia), (lb, ib) = _pair(pkg, cfg, storage)
    for n in range(1, 7):
        fx, fy, fz = ia.compute_force(la.ux, la.uy, *cfg.motion_state(n))
        la.step(fx, fy, fz)
        coupled_step(lb, ib, cfg, n, interleave=3, **MODES[mode])
        assert lb.state_word == la.state_word
    np.testing.assert_array_equal(la.f.to_numpy(), lb.f.to_numpy())
    for k in ("X_marker", "Y_marker", "U_k", "V_k", "FX_k", "FY_k"):
        np.testing.assert_array_equal(getattr(ia, k).to_numpy(), getattr(ib, k).to_numpy(), k)
    np.testing.assert_array_equal(ia.force_box(), ib.force_box())
    assert ib.region_violations() == 0
    # the accumulators are clear again, and the clamp's fields rebuild from the
    # markers when read; the stand-alone spread kernel evaluates the delta
    # weights in its own instruction order (a weight may differ by a float32
    # rounding): the rebuilt fields agree to 1e-8 absolute (|F| <= f_max = 0.05)
    assert not ib._acc.any().item()
    np.testing.assert_allclose(ib.fx.to_numpy(), ia.fx.to_numpy(), rtol=0, atol=1e-8)
    np.testing.assert_allclose(ib.fy.to_numpy(), ia.fy.to_numpy(), rtol=0, atol=1e-8)
    assert not np.any(ib.fz.to_numpy())
    # then a plain compute_force + step carries on from the coupled state
    fx, fy, fz = ia.compute_force(la.ux, la.uy, *cfg.motion_state(7))
    la.step(fx, fy, fz)
    fx, fy, fz = ib.compute_force(lb.ux, lb.uy, *cfg.motion_state(7))
    lb.step(fx, fy, fz)
    np.testing.assert_array_equal(la.f.to_numpy(), lb.f.to_numpy())
    np.testing.assert_array_equal(ia.fx.to_numpy(), ib.fx.to_numpy())


@pytest.mark.gpu
def test_plain_step_after_unread_coupled_steps_clears_stale_forces(pkg, gpu):
    """Coupled steps whose fields nobody read, then a plain compute_force:
    fx, fy hold exactly this generation's clamped forces (no stale cells)."""
    cfg = pkg.config.naca_case(96, 48, 8, 12.0, name="stale")
    (la, ia), (lb, ib) = _pair(pkg, cfg, "f32")
    for n in range(1, 4):
        fx, fy, fz = ia.compute_force(la.ux, la.uy, *cfg.motion_state(n))
        la.step(fx, fy, fz)
    fx, fy, fz = ib.compute_force(lb.ux, lb.uy, *cfg.motion_state(1))
    lb.step(fx, fy, fz)
    ib.fx.to_numpy()                            # fields of step 1 materialised
    for n in (2, 3):
        coupled_step(lb, ib, cfg, n)            # fields of steps 2, 3 never read
    for n in range(4, 30):                      # the body has moved on by then
        for l, i in ((la, ia), (lb, ib)):
            fx, fy, fz = i.compute_force(l.ux, l.uy, *cfg.motion_state(n))
            l.step(fx, fy, fz)
    np.testing.assert_array_equal(ia.fx.to_numpy(), ib.fx.to_numpy())
    np.testing.assert_array_equal(ia.fy.to_numpy(), ib.fy.to_numpy())
    np.testing.assert_array_equal(la.f.to_numpy(), lb.f.to_numpy())


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["f32", "f16c"])
def test_coupled_step_at_600x200x50(pkg, gpu, storage):
    cfg = pkg.config.case_c1()
    (la, ia), (lb, ib) = _pair(pkg, cfg, storage)
    for n in range(1, 5):
        fx, fy, fz = ia.compute_force(la.ux, la.uy, *cfg.motion_state(n))
        la.step(fx, fy, fz)
        coupled_step(lb, ib, cfg, n, **(MODES["in-launch"] if n % 2 else MODES["tail"]))
    np.testing.assert_array_equal(la.f.to_numpy(), lb.f.to_numpy())
    np.testing.assert_array_equal(ia.FX_k.to_numpy(), ib.FX_k.to_numpy())
    assert ib.region_violations() == 0


@pytest.mark.parametrize("change,match", [
    (dict(interleave=0), "interleave"),
    (dict(spread=0), "inside_spread"),
    (dict(n=0), "n must be"),
    (dict(gen=0), "generation"),
    (dict(mode=1), "fixed-point"),
    (dict(region=(0, 0, 0, 10, 5, 3)), "normalised"),
])
def test_coupled_step_argument_errors(pkg, change, match):
    """Rejected on the host before any launch (no GPU is needed)."""
    lib = _lib()
    g = lib.ClbmGrid(12, 6, 4, lib.STORE_F32)
    p = lib.ClbmParams(1.7, 1.0, 0.05, 0.0, 1, 0)
    ip = lib.IbmParams(0.05, 1, change.get("mode", 0), 0, 8, 0)
    one = C.c_void_p(16)
    r = change.get("region", (2, 0, 0, 5, 5, 3))
    region = (C.c_int32 * 6)(*r)
    with pytest.raises(lib.ClbmError, match=match):
        lib.call("clbm_ibm_step", C.byref(g), C.byref(p), one, 0, C.byref(ip),
                 change.get("n", 4), None, one, one, one, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0,
                 *([one] * 9), None, one, one, one, change.get("gen", 1), region,
                 change.get("interleave", 1), change.get("spread", 1), 0, None)
