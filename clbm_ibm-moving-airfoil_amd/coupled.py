"""The coupled step on one stream: compute_force followed by step in one call.

No reference counterpart as one call: the reference runs
``DroneMovingIBM.compute_force`` and then ``CLBMSolver.step``
(main_moving.py:116-122).  ``step_coupled`` does the same work through
``clbm_ibm_step``: one launch holding the marker workgroups and the update of
the lattice away from the body, and a short tail (or the end of the same
launch, after a bounded wait for the markers) for the body's region, which
takes the force straight from the spread accumulators.  The results equal
those of the two calls bit for bit; only one stream is used, with no events.

The IBM object's ``fx, fy`` fields are not written by it.  They are rebuilt
from the marker forces (spread + clamp) when something reads them, within
1e-8 of what the clamp kernel writes (the stand-alone spread evaluates the
delta weights in its own instruction order).

The default coupled loop of the package keeps the two-stream split step
(``DroneMovingIBM(overlap=True)``), which measured faster on MI355X.
"""

from __future__ import annotations

import ctypes as C

from . import _lib


def step_coupled(lbm, ibm, xc, yc, theta, vx, vy, omega, interleave: int = 2,
                 inside_spread: int = 1, wait_spins: int = 100000) -> None:
    """compute_force(lbm.ux, lbm.uy, xc, ...) then lbm.step(fx, fy, fz), in
    one call on the current stream.  ``interleave``: one super-block of eight
    marker workgroups every ``interleave`` super-blocks of the launch;
    ``wait_spins`` > 0: the body's region is stepped at the end of the launch
    (one inside super-block every ``inside_spread``) after waiting at most that
    many polls for the markers, else in the tail kernel; 0: always in the
    tail; < 0: the inside workgroups give up at once (the tail steps the
    region)."""
    import torch
    if ibm.ip.spread_mode != _lib.SPREAD_FIXED:
        raise ValueError("step_coupled needs the fixed-point spread mode")
    if (lbm.Nx, lbm.Ny, lbm.Nz) != (ibm.Nx, ibm.Ny, ibm.Nz):
        raise ValueError("solver and IBM lattices differ")
    ibm._join()
    ibm._last_side = False      # this step's marker work runs in the launch below
    lbm._wait_readers()
    region = (C.c_int32 * 6)()
    rc = _lib.fn("ibm_step_region")(C.byref(lbm.grid), C.byref(lbm.params), *ibm._hull_args,
                                    float(xc), float(yc), float(theta), region)
    if rc < 0:
        _lib.check(rc, "ibm_step_region")
    if rc == 0:
        # no split region (body off the grid, or over half of it): both calls
        fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, xc, yc, theta, vx, vy, omega)
        lbm.step(fx, fy, fz)
        return
    ibm.gen += 1
    P = C.c_void_p
    m = lambda a: P(a._t.data_ptr())
    st = P(torch.cuda.current_stream(lbm.device).cuda_stream)
    base = ibm._stencil_base
    _lib.call("clbm_ibm_step", C.byref(lbm.grid), C.byref(lbm.params), P(lbm.state.data_ptr()),
              lbm.state_word, C.byref(ibm.ip), ibm.N,
              P(ibm.order.data_ptr() if ibm.order is not None else 0), m(ibm.X0), m(ibm.Y0),
              m(ibm.Z0), float(xc), float(yc), float(theta), float(vx), float(vy), float(omega),
              m(ibm.X_marker), m(ibm.Y_marker), m(ibm.Z_marker), m(ibm.VX_body), m(ibm.VY_body),
              m(ibm.U_k), m(ibm.V_k), m(ibm.FX_k), m(ibm.FY_k),
              P(base.data_ptr() if base is not None else 0), P(ibm._acc[0].data_ptr()),
              P(ibm._acc[1].data_ptr()), P(ibm.box.data_ptr()), ibm.gen, region,
              int(interleave), int(inside_spread), int(wait_spins), st)
    ibm._fields_stale = True
    lbm._split_ibms.add(ibm)     # health() checks its region flag
    lbm.state_word = _lib.fn("clbm_state_after_step")(lbm.state_word, lbm._pref)
    lbm.steps += 1
    lbm._macro_valid = False
    lbm.last_kernel_cells = lbm.ncell
    lbm.last_step_kind = "coupled"
