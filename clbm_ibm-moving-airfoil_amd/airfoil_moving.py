"""DroneMovingIBM / FishIBM: the reference's immersed-boundary object over HIP kernels.

Mirrors ``DroneMovingIBM`` of airfoil_moving.py:44-255 (and its token-identical
clone ``FishIBM``, fish/ibm_fish.py): constructed without arguments from the
``config_moving`` convention (or a CaseConfig), ``compute_force(ux, uy, xc,
yc, theta, vx, vy, omega)`` returns its own ``fx, fy, fz`` fields, ready for
``CLBMSolver.step``.  Per call it launches two kernels instead of six: one
for marker update, interpolation, marker force and spreading (K8, K10-K12;
markers visited in a spatially sorted order so that each workgroup can
pre-sum its stencils), and one that clamps the accumulated force into the
fields (K9 + K13) on the bounding box of the touched cells.

When the velocity comes from a CLBMSolver (the reference's call
``compute_force(lbm.ux, lbm.uy, ...)``), the two launches run on a stream of
this object's own (``overlap=True``): the host predicts the box of cells the
stencils can reach from the kinematics (the convex hull of the template,
rotated in float32 as the kernel does, plus one cell), the kernel checks
that no stencil leaves it, and the solver's next ``step`` updates the rest
of the lattice while these kernels run.  Reading any of this object's fields
first orders the current stream after them.

Added readout with no reference counterpart: FP64 sums of the marker forces
(``force_sums``) and a device-side history of them for C_D / C_L.
"""

from __future__ import annotations

import ctypes as C
import math
import weakref

import numpy as np

from . import _lib
from .config import CaseConfig, convex_hull_2d, resolve, spatial_order
from .fields import Field, MarkerField


def stencil_region(hx, hy, zr, shape, xc, yc, theta):
    """Inclusive (x0, y0, z0, x1, y1, z1) box holding every stencil cell of a
    rigid template placed by (xc, yc, theta), from the template's hull
    vertices (hx, hy) and the z range ``zr`` of its stencils.

    The hull vertices are placed with the kernel's float32 arithmetic (cos/sin
    of the float32 angle rounded once from float64, no contraction).  The
    extreme coordinates of the exact placement are attained at hull vertices;
    every other marker's float32 position differs from its exact one by a few
    ulp, so one cell of margin on each side covers it.  x is clamped to the
    grid; a y range that crosses the periodic boundary takes the whole axis.
    None for non-finite kinematics.  The library's host function
    ibm_step_region computes the same box (and its normalisation) per step;
    this is its statement in NumPy."""
    nx, ny, nz = shape
    f32 = np.float32
    t = np.float64(f32(theta))
    if not math.isfinite(t):
        return None
    c, s = f32(math.cos(t)), f32(math.sin(t))
    X = (f32(xc) + hx * c) - hy * s
    Y = (f32(yc) + hx * s) + hy * c
    if len(X) == 0 or not (np.all(np.isfinite(X)) and np.all(np.isfinite(Y))):
        return None
    # stencil cells floor(p) - 1 .. floor(p) + 2, plus one cell of margin
    x0, x1 = int(np.floor(X.min())) - 2, int(np.floor(X.max())) + 3
    y0, y1 = int(np.floor(Y.min())) - 2, int(np.floor(Y.max())) + 3
    x0, x1 = max(x0, 0), min(x1, nx - 1)
    if x0 > x1:
        return None
    if y0 < 0 or y1 > ny - 1:
        y0, y1 = 0, ny - 1
    return (x0, y0, int(zr[0]), x1, y1, int(zr[1]))


def _capturing() -> bool:
    import torch
    return torch.cuda.is_current_stream_capturing()


def stencil_z_range(Z0, nz):
    """The z range of the stencils of markers at heights Z0 (z never moves):
    floor(z) - 1 .. floor(z) + 2, or the whole axis when a stencil wraps."""
    z = np.asarray(Z0, dtype=np.float32)
    k0, k1 = int(np.floor(z.min())) - 1, int(np.floor(z.max())) + 2
    return (k0, k1) if (k0 >= 0 and k1 <= nz - 1) else (0, nz - 1)


class DroneMovingIBM:
    """Direct-forcing IBM with the Peskin 4-point delta for a rigid body in heave + pitch.

    Options beyond the reference: ``spread_mode`` "fixed" (order-independent
    fixed-point sums, the default), "atomic" (float atomics like
    ti.atomic_add) or "sorted" (the stencil points grouped by cell and each
    cell's contributions added in marker order, in float32: deterministic
    without atomics on the fields; separate marker, interpolation, spread and
    clamp launches on the caller's stream); ``wrap_weight_quirk`` False to weight wrapped stencil
    points by their unwrapped offset; ``record_stencil`` to keep floor(pos)-1
    per marker (for index checks); ``order`` "natural" (the template's own
    order, neighbours along each section) or "morton" (a Z-order sort of the
    template) for the marker groups of the fused kernel; ``group`` markers
    per workgroup (4, 8 or 16); ``overlap`` True to run the marker kernels on
    a stream of their own, concurrently with the solver's next step, False to
    keep them on the caller's stream, "auto" (the default) for True on
    lattices of ``OVERLAP_MIN_CELLS`` cells or more, where the step kernels
    outlast the host work of the split step.  None of these tuning options
    change results in the fixed spread mode."""

    OVERLAP_MIN_CELLS = 4_000_000

    def __init__(self, config=None, device=None, spread_mode: str = "fixed",
                 f_max=None, wrap_weight_quirk=None, record_stencil: bool = False,
                 order: str = "natural", group: int = 8, overlap="auto"):
        import torch
        cfg: CaseConfig = resolve(config)
        self.cfg = cfg
        self.device = torch.device(device if device is not None else "cuda")
        self.Nx, self.Ny, self.Nz = int(cfg.Nx), int(cfg.Ny), int(cfg.Nz)
        self.N = cfg.N_MARKERS
        _lib.load()
        dev = self.device
        f32 = torch.float32

        def marr(a=None):
            if a is None:
                return MarkerField(torch.zeros(self.N, dtype=f32, device=dev), sync=self._join)
            return MarkerField(torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev),
                               sync=self._join)

        self._pending = False

        self.X0, self.Y0, self.Z0 = marr(cfg.X0), marr(cfg.Y0), marr(cfg.Z0)
        self.X_marker, self.Y_marker, self.Z_marker = marr(), marr(), marr()
        self.VX_body, self.VY_body = marr(), marr()
        self.U_k, self.V_k = marr(), marr()
        self.FX_k, self.FY_k = marr(), marr()
        self.dV = marr(cfg.dV) if cfg.dV is not None else None
        # visiting order: neighbours in the body frame stay neighbours after any rigid motion
        self.order = (torch.from_numpy(spatial_order(cfg.X0, cfg.Y0, cfg.Z0)).to(dev)
                      if order == "morton" else None)
        shape = (self.Nx, self.Ny, self.Nz)
        ncell = self.Nx * self.Ny * self.Nz
        # force-box record (uint64 words, zero = no generation yet) and the generation counter
        self.box = torch.zeros(_lib.BOX_WORDS, dtype=torch.int64, device=dev)
        self.gen = 0
        box_ptr = self.box.data_ptr() + 8 * _lib.FORCE_BOX_OFFSET
        # True after coupled.step_coupled, which leaves fx, fy unwritten
        self._fields_stale = False
        self.fx = Field(shape, torch.zeros((self.Nz, self.Ny, self.Nx), dtype=f32, device=dev),
                        refresh=self._refresh_fields, owner=self, box=box_ptr)
        self.fy = Field(shape, torch.zeros((self.Nz, self.Ny, self.Nx), dtype=f32, device=dev),
                        refresh=self._refresh_fields, owner=self, box=box_ptr)
        self.fz = Field(shape, torch.zeros((self.Nz, self.Ny, self.Nx), dtype=f32, device=dev),
                        refresh=self._refresh_fields, owner=self, box=box_ptr)
        self._acc = torch.zeros((2, ncell), dtype=torch.int64, device=dev)
        # what the solver's split step reads: the force fields and their box
        self._force_ptrs = (C.c_void_p(self.fx._t.data_ptr()), C.c_void_p(self.fy._t.data_ptr()),
                            C.c_void_p(box_ptr))
        self.ip = _lib.IbmParams(float(cfg.f_max if f_max is None else f_max),
                                 int(cfg.wrap_weight_quirk if wrap_weight_quirk is None
                                     else wrap_weight_quirk),
                                 _lib.SPREAD_MODES[spread_mode], 0, int(group), 0)
        self._stencil_base = (torch.zeros((self.N, 3), dtype=torch.int32, device=dev)
                              if record_stencil else None)
        # work buffer of the sorted spread (zero before its first call; each call
        # leaves it so)
        self._sort_work = None
        if spread_mode == "sorted":
            nb = _lib.fn("ibm_spread_sorted_bytes")(
                C.byref(_lib.ClbmGrid(self.Nx, self.Ny, self.Nz, _lib.STORE_F32)), self.N)
            self._sort_work = torch.zeros(max(int(nb), 1), dtype=torch.uint8, device=dev)
        # region prediction: hull of the template in the body frame, z range of its stencils
        if overlap == "auto":
            overlap = self.Nx * self.Ny * self.Nz >= self.OVERLAP_MIN_CELLS
        elif not isinstance(overlap, bool):
            raise ValueError("overlap must be True, False or 'auto'")
        self.overlap = overlap
        hull = convex_hull_2d(cfg.X0, cfg.Y0) if self.N else np.zeros(0, np.int64)
        self._hx = np.ascontiguousarray(np.asarray(cfg.X0, dtype=np.float32)[hull])
        self._hy = np.ascontiguousarray(np.asarray(cfg.Y0, dtype=np.float32)[hull])
        self._zr = stencil_z_range(cfg.Z0, self.Nz) if self.N else (0, self.Nz - 1)
        self._hull_args = (C.c_void_p(self._hx.ctypes.data), C.c_void_p(self._hy.ctypes.data),
                           len(self._hx), int(self._zr[0]), int(self._zr[1]))
        # the marker stream (the library's, shared by all bodies on the device),
        # whether it runs beside the caller's stream, and the two events that
        # order it against that stream (fork before the marker kernels, join
        # after them)
        self._side_h = None
        self.side_concurrent = None
        self._ev = []
        # whether the last compute_force ran its marker kernels on the marker
        # stream, and whether force sums were queued there since the last join
        self._last_side = False
        self._side_sums = False
        # the generation whose clamped forces the last split compute_force
        # published for a one-launch split step (clbm_step_fused)
        self._published = 0
        # True once anything may have touched this object's buffers on the
        # caller's stream since the solver bound the fork event to its last
        # step (compute_force then records the event again)
        self._fork_void = True
        # ctypes arguments of the marker and clamp calls that stay fixed for a
        # velocity source: (weak reference to the solver, before-state,
        # after-state, clamp prefix)
        self._argcache = None
        self._hist = None
        self._hist_n = 0
        # with a solver as the velocity source, compute_force is one library call
        # over a launch plan (False: the calls one by one, as _region et al. do)
        self.fast_host = True
        self._plan = None
        self._plan_solver = None
        self._sums = torch.zeros(2, dtype=torch.float64, device=dev)

    # -- reference surface ------------------------------------------------
    def compute_force(self, ux, uy, xc: float, yc: float, theta: float, vx: float, vy: float,
                      omega: float):
        """Markers, interpolation, marker forces, spreading and clamping
        (airfoil_moving.py:230-255).  ``ux``/``uy`` are the solver's velocity
        fields (read straight from its populations) or any [Nz,Ny,Nx] float32
        device Field/tensor.  Returns (fx, fy, fz), this object's own fields."""
        self._last_side = False
        if self._fields_stale:
            # the last coupled step left fx, fy unwritten: clear them so that the
            # clamp of this call (which clears the last two boxes) leaves them exact
            self._fields_stale = False
            self._fork_void = True
            self.fx._t.zero_()
            self.fy._t.zero_()
        if self._sort_work is not None:
            return self._compute_force_sorted(ux, uy, xc, yc, theta, vx, vy, omega)
        pre, post, clip, solver, state = self._call_args(ux, uy)
        if solver is not None and self.fast_host:
            return self._compute_force_plan(solver, xc, yc, theta, vx, vy, omega)
        region = self._region(solver, xc, yc, theta) if solver is not None else None
        main_h = _lib.stream_of(self.device)
        if region is not None:
            # run on the marker stream once everything queued so far is done
            if self._side_h is None:
                self._open_side()
            _lib.check(_lib.fn("clbm_stream_after")(self._side_h, main_h, self._ev[0]),
                       "clbm_stream_after")
            st = self._side_h
            self.ip.check_region = 1
            self.ip.region[:] = list(region)
        else:
            self._join()
            st = main_h
            self.ip.check_region = 0
        self.gen += 1
        _lib.check(_lib.fn("ibm_compute_force")(*pre, state, *post[0], float(xc), float(yc),
                                                float(theta), float(vx), float(vy), float(omega),
                                                *post[1], self.gen, st), "ibm_compute_force")
        _lib.check(_lib.fn("ibm_clip_forces")(*clip, self.gen, st), "ibm_clip_forces")
        if region is not None:
            self._pending = True
            self._last_side = True
            solver._add_reader(self, region)
        return self.fx, self.fy, self.fz

    def _compute_force_sorted(self, ux, uy, xc, yc, theta, vx, vy, omega):
        """compute_force with the sorted spread: ibm_update_markers,
        ibm_interp_force, ibm_spread_sorted and ibm_clip_forces on the
        caller's stream (K8, K10 + K11, K12, K9 + K13)."""
        self._join()
        grid, lp, f, state, uxp, uyp = self._velocity_source(ux, uy)
        P = C.c_void_p
        m = lambda a: P(a._t.data_ptr())
        st = self._stream()
        self.gen += 1
        gref, ipref = C.byref(grid), C.byref(self.ip)
        boxp = P(self.box.data_ptr())
        acc0, acc1 = P(self._acc[0].data_ptr()), P(self._acc[1].data_ptr())
        _lib.call("ibm_update_markers", gref, self.N, m(self.X0), m(self.Y0), m(self.Z0),
                  float(xc), float(yc), float(theta), float(vx), float(vy), float(omega),
                  m(self.X_marker), m(self.Y_marker), m(self.Z_marker), m(self.VX_body),
                  m(self.VY_body), boxp, self.gen, st)
        _lib.call("ibm_interp_force", gref, ipref, lp, P(f), state, P(uxp), P(uyp), self.N,
                  m(self.X_marker), m(self.Y_marker), m(self.Z_marker), m(self.VX_body),
                  m(self.VY_body), m(self.U_k), m(self.V_k), m(self.FX_k), m(self.FY_k),
                  P(self._stencil_base.data_ptr() if self._stencil_base is not None else 0), st)
        _lib.call("ibm_spread_sorted", gref, ipref, self.N, m(self.X_marker), m(self.Y_marker),
                  m(self.Z_marker), m(self.FX_k), m(self.FY_k), acc0, acc1,
                  P(self._sort_work.data_ptr()), self._sort_work.numel(), st)
        _lib.call("ibm_clip_forces", gref, ipref, acc0, acc1, m(self.fx), m(self.fy), boxp,
                  self.gen, st)
        return self.fx, self.fy, self.fz

    def _compute_force_plan(self, solver, xc, yc, theta, vx, vy, omega):
        """compute_force with the velocity from ``solver``'s populations as one
        library call (ibm_compute_force_split) over a launch plan built once:
        the step region, the fork of the marker stream and both launches."""
        plan = self._plan
        ip = self.ip
        if (plan is None or self._plan_solver() is not solver or
                (plan.ip.f_max, plan.ip.wrap_weight_quirk, plan.ip.spread_mode,
                 plan.ip.fixed_shift, plan.ip.group) !=
                (ip.f_max, ip.wrap_weight_quirk, ip.spread_mode, ip.fixed_shift, ip.group)):
            plan = self._build_plan(solver)
        if self._pending:
            self._join()
        if self._side_sums:
            # force sums queued on the marker stream still read FX_k: marker
            # kernels queued there follow them, but ones that run on the caller's
            # stream (no split this time) must wait for them
            probe = (C.c_int32 * 6)()
            rc = _lib.fn("ibm_step_region")(C.byref(plan.grid), C.byref(plan.params),
                                            *self._hull_args, float(xc), float(yc), float(theta),
                                            probe)
            if rc != 1:
                self._join()
        main_h = _lib.stream_of(self.device)
        # the fork event already completes with the solver's last step when
        # nothing else went onto this stream since (no new record needed)
        fork = solver._fork
        bound = int(fork is not None and fork[0] is self and fork[1] == main_h
                    and not self._fork_void and not _capturing())
        solver._fork = None
        self.gen += 1
        region = self._region_buf
        rc = self._split_fn(C.byref(plan), solver.state_word, self.gen, float(xc), float(yc),
                            float(theta), float(vx), float(vy), float(omega), main_h, bound,
                            region)
        if rc < 0:
            _lib.check(rc, "ibm_compute_force_split")
        if rc == 1:
            self._pending = True
            self._last_side = True
            self._published = self.gen
            solver._add_reader(self, region)
        return self.fx, self.fy, self.fz

    def _build_plan(self, solver):
        """The ibm_split_plan_t of this body on ``solver``'s lattice (and the
        marker stream with its two events when overlapping)."""
        if (solver.Nx, solver.Ny, solver.Nz) != (self.Nx, self.Ny, self.Nz):
            raise ValueError("solver and IBM lattices differ")
        if self.overlap and self.N and self._side_h is None:
            self._open_side()
        pl = _lib.IbmSplitPlan()
        pl.grid = solver.grid
        pl.params = solver.params
        pl.ip = self.ip
        pl.f = solver.state.data_ptr()
        pl.n = self.N
        pl.order = self.order.data_ptr() if self.order is not None else None
        for name, fld in (("X0", self.X0), ("Y0", self.Y0), ("Z0", self.Z0), ("X", self.X_marker),
                          ("Y", self.Y_marker), ("Z", self.Z_marker), ("VX", self.VX_body),
                          ("VY", self.VY_body), ("U", self.U_k), ("V", self.V_k),
                          ("FX", self.FX_k), ("FY", self.FY_k)):
            setattr(pl, name, fld._t.data_ptr())
        pl.stencil_base = (self._stencil_base.data_ptr() if self._stencil_base is not None
                           else None)
        pl.acc_x, pl.acc_y = self._acc[0].data_ptr(), self._acc[1].data_ptr()
        pl.box = self.box.data_ptr()
        pl.fx, pl.fy = self.fx._t.data_ptr(), self.fy._t.data_ptr()
        pl.hx, pl.hy = self._hx.ctypes.data, self._hy.ctypes.data
        pl.nh, pl.z0, pl.z1 = len(self._hx), int(self._zr[0]), int(self._zr[1])
        on = self.overlap and self.N > 0
        pl.side = self._side_h.value if on else None
        pl.fork_ev = self._ev[0].value if on else None
        self._plan = pl
        self._plan_solver = weakref.ref(solver)
        self._region_buf = (C.c_int32 * 6)()
        self._split_fn = _lib.fn("ibm_compute_force_split")
        self._box_ptr = C.c_void_p(self.box.data_ptr())
        return pl

    def _call_args(self, ux, uy):
        """ctypes arguments of ibm_compute_force / ibm_clip_forces around the
        per-call ones, the solver whose populations give the velocity (or
        None) and its state word.  Cached for a solver source; built anew for
        velocity tensors."""
        from .clbm_moving import CLBMSolver
        owner = getattr(ux, "owner", None)
        solver = owner if (isinstance(owner, CLBMSolver) and getattr(uy, "owner", None) is owner) \
            else None
        if solver is not None and self._argcache is not None and self._argcache[0]() is solver:
            return (*self._argcache[1:], solver, solver.state_word)
        grid, lp, f, state, uxp, uyp = self._velocity_source(ux, uy)
        P = C.c_void_p
        m = lambda a: P(a._t.data_ptr())
        gref, ipref = C.byref(grid), C.byref(self.ip)
        acc0, acc1, boxp = P(self._acc[0].data_ptr()), P(self._acc[1].data_ptr()), P(self.box.data_ptr())
        pre = (gref, ipref, lp, P(f))
        post = ((P(uxp), P(uyp), self.N, P(self.order.data_ptr() if self.order is not None else 0),
                 m(self.X0), m(self.Y0), m(self.Z0)),
                (m(self.X_marker), m(self.Y_marker), m(self.Z_marker), m(self.VX_body),
                 m(self.VY_body), m(self.U_k), m(self.V_k), m(self.FX_k), m(self.FY_k),
                 P(self._stencil_base.data_ptr() if self._stencil_base is not None else 0),
                 acc0, acc1, boxp))
        clip = (gref, ipref, acc0, acc1, P(self.fx._t.data_ptr()), P(self.fy._t.data_ptr()), boxp)
        if solver is not None:
            # a weak reference: the cache must not keep a discarded solver's buffers alive
            self._argcache = (weakref.ref(solver), pre, post, clip)
        return pre, post, clip, solver, state

    # -- the concurrent stream ------------------------------------------------
    def _open_side(self) -> None:
        self._side_h, self.side_concurrent = _lib.side_stream(self.device)
        self._ev = [self._new_event(), self._new_event()]

    def _new_event(self) -> C.c_void_p:
        ev = C.c_void_p()
        _lib.call("clbm_event_create", C.byref(ev))
        return ev

    def __del__(self):
        try:
            for ev in self._ev:
                _lib.fn("clbm_event_destroy")(ev)
        except Exception:  # noqa: BLE001 (interpreter shutdown)
            pass

    def _region(self, solver, xc, yc, theta):
        """The normalised int32[6] region holding every stencil cell of this
        call (ibm_step_region), or None when overlapping would not pay (region
        over half the lattice) or is switched off."""
        if not self.overlap or self.N == 0:
            return None
        out = (C.c_int32 * 6)()
        rc = _lib.fn("ibm_step_region")(C.byref(solver.grid), C.byref(solver.params),
                                        *self._hull_args, float(xc), float(yc), float(theta), out)
        if rc < 0:
            _lib.check(rc, "ibm_step_region")
        return out if rc == 1 else None

    def _join(self) -> None:
        """Order the current stream after this object's pending kernels
        (before the caller's stream touches its buffers)."""
        self._fork_void = True
        if self._pending or self._side_sums:
            main_h = _lib.stream_of(self.device)
            _lib.check(_lib.fn("clbm_stream_after")(main_h, self._side_h, self._ev[1]),
                       "clbm_stream_after")
            self._pending = False
            self._side_sums = False

    def _refresh_fields(self) -> None:
        """Before fx, fy, fz are read: order after pending marker kernels and,
        after a coupled step (which leaves the fields unwritten), rebuild
        them from the markers: spread of F_k into the accumulators, then the
        clamp of this generation."""
        self._join()
        if not self._fields_stale:
            return
        self._fields_stale = False
        P = C.c_void_p
        st = self._stream()
        grid = _lib.ClbmGrid(self.Nx, self.Ny, self.Nz, _lib.STORE_F32)
        self.fx._t.zero_()
        self.fy._t.zero_()
        acc0, acc1 = P(self._acc[0].data_ptr()), P(self._acc[1].data_ptr())
        _lib.call("ibm_spread", C.byref(grid), C.byref(self.ip), self.N,
                  P(self.X_marker._t.data_ptr()), P(self.Y_marker._t.data_ptr()),
                  P(self.Z_marker._t.data_ptr()), P(self.FX_k._t.data_ptr()),
                  P(self.FY_k._t.data_ptr()), acc0, acc1, st)
        _lib.call("ibm_clip_forces", C.byref(grid), C.byref(self.ip), acc0, acc1,
                  P(self.fx._t.data_ptr()), P(self.fy._t.data_ptr()), P(self.box.data_ptr()),
                  self.gen, st)

    @property
    def stencil_base(self):
        """(N, 3) int32 device tensor of floor(pos) - 1 per marker (record_stencil=True)."""
        self._join()
        return self._stencil_base

    def region_violations(self) -> int:
        """The last generation whose stencils left the predicted region (0 if none)."""
        self._join()
        return int(self.box[_lib.REGION_FLAG_OFFSET].item())

    def wait_expirations(self) -> int:
        """The last generation whose one-launch split step (clbm_step_fused)
        ran out of its wait for the clamped forces (0 if none)."""
        self._join()
        return int(self.box[_lib.SPLIT_EXPIRED_OFFSET].item())

    def check_region(self) -> None:
        """Raise ClbmError when the stencils of some compute_force call left
        the region predicted for the split step (the marker kernel marks the
        generation in the box record; reading it synchronises).  A split step
        whose markers left their region ran them concurrently with the update
        of cells they read, so the results after that generation are not
        trustworthy.  Likewise when a one-launch split step stepped the region
        without waiting for its forces (the wait for them ran out)."""
        self._join()
        g, late = (int(v) for v in self.box[[_lib.REGION_FLAG_OFFSET,
                                             _lib.SPLIT_EXPIRED_OFFSET]].tolist())
        if g:
            raise _lib.ClbmError(
                f"the stencils of compute_force call {g} (generation {g}) left the region "
                f"predicted for the split step; the marker kernels raced the lattice update "
                f"away from the body (use overlap=False for this body)")
        if late:
            raise _lib.ClbmError(
                f"the split step of compute_force call {late} (generation {late}) stepped the "
                f"body's region before its forces were published (its wait ran out); the "
                f"results from that step on are not trustworthy")

    def force_box(self):
        """The (x0, y0, z0, x1, y1, z1) box of the cells the last compute_force reached."""
        self._refresh_fields()     # a step that left the fields unwritten left the box too
        return self.box[_lib.FORCE_BOX_OFFSET:_lib.FORCE_BOX_OFFSET + 3].cpu().numpy().view(
            np.int32).copy()

    def _velocity_source(self, ux, uy):
        from .clbm_moving import CLBMSolver
        owner = getattr(ux, "owner", None)
        if isinstance(owner, CLBMSolver) and getattr(uy, "owner", None) is owner:
            if (owner.Nx, owner.Ny, owner.Nz) != (self.Nx, self.Ny, self.Nz):
                raise ValueError("solver and IBM lattices differ")
            return (owner.grid, C.byref(owner.params), owner.state.data_ptr(), owner.state_word,
                    0, 0)
        from .fields import as_device_ptr
        shape = (self.Nx, self.Ny, self.Nz)
        grid = _lib.ClbmGrid(self.Nx, self.Ny, self.Nz, _lib.STORE_F32)
        return grid, None, 0, 0, as_device_ptr(ux, shape), as_device_ptr(uy, shape)

    # -- force readout (no reference counterpart) ----------------------------
    def force_sums(self, out=None):
        """Launch the FP64 sums (sum FX dV, sum FY dV) of the current marker
        forces into ``out`` (a float64 device tensor of 2, default an internal
        one) and return it; nothing is synchronised."""
        out = self._sums if out is None else out
        self._join()
        _lib.call("ibm_force_sum_f64", self.N, C.c_void_p(self.FX_k.data_ptr()),
                  C.c_void_p(self.FY_k.data_ptr()),
                  C.c_void_p(self.dV.data_ptr() if self.dV is not None else 0),
                  C.c_void_p(out.data_ptr()), self._stream())
        return out

    def record_force(self) -> None:
        """Append this step's FP64 force sums to the device history.  After a
        compute_force whose marker kernels ran on the marker stream the sums
        are queued there too, after them and before the next step's: the
        caller's stream gets no work, so the solver's next step keeps its
        fork bound (force_history and any other reader join first)."""
        import torch
        side = self._last_side and self._side_h is not None
        if self._hist is None or self._hist_n == self._hist.shape[0]:
            self._join()
            cap = 1024 if self._hist is None else 2 * self._hist.shape[0]
            h = torch.zeros((cap, 2), dtype=torch.float64, device=self.device)
            if self._hist is not None:
                h[: self._hist_n] = self._hist[: self._hist_n]
            self._hist = h
            if side:
                # the marker stream writes into the new buffer after it is filled
                _lib.check(_lib.fn("clbm_stream_after")(self._side_h,
                                                        _lib.stream_of(self.device), self._ev[0]),
                           "clbm_stream_after")
        if side:
            t = lambda a: C.c_void_p(a._t.data_ptr())
            _lib.call("ibm_force_sum_f64", self.N, t(self.FX_k), t(self.FY_k),
                      t(self.dV) if self.dV is not None else C.c_void_p(0),
                      C.c_void_p(self._hist[self._hist_n].data_ptr()), self._side_h)
            self._side_sums = True
        else:
            self.force_sums(self._hist[self._hist_n])
        self._hist_n += 1

    def force_history(self) -> np.ndarray:
        """(steps, 2) float64: per recorded step, sum_k FX_k dV_k and sum_k FY_k dV_k
        (the force the markers put on the fluid)."""
        if self._hist is None:
            return np.zeros((0, 2))
        self._join()
        return self._hist[: self._hist_n].cpu().numpy().copy()

    def restore_history(self, sums) -> None:
        """Replace the recorded history by ``sums`` ((steps, 2), as
        force_history returns it), e.g. when a run resumes from a checkpoint;
        record_force appends after it."""
        import torch
        a = np.asarray(sums, dtype=np.float64).reshape(-1, 2)
        self._join()
        cap = max(1024, 1 << max(0, int(len(a) - 1).bit_length()))
        h = torch.zeros((cap, 2), dtype=torch.float64, device=self.device)
        if len(a):
            h[: len(a)] = torch.from_numpy(a).to(self.device)
        self._hist, self._hist_n = h, len(a)

    def coefficients(self, sums, rho_inf: float = 1.0, span=None, momentum_factor: float = 1.0):
        """Drag and lift coefficients from force sums (array (..., 2)).

        The body feels -sum_k F_k dV_k.  Drag is its component along the
        inflow direction (ux_in, uy_in), lift the component 90 degrees
        counter-clockwise of it; both are divided by 1/2 rho U^2 c span
        (span = Nz by default).  ``momentum_factor`` scales the sums; the
        solver's momentum gain per unit force per step is (1.5 - s_v/2)."""
        s = np.asarray(sums, dtype=np.float64) * momentum_factor
        cfg = self.cfg
        U = math.hypot(cfg.ux_in, cfg.uy_in) or cfg.U_inf
        ex, ey = (cfg.ux_in / U, cfg.uy_in / U) if U > 0 else (1.0, 0.0)
        span = float(self.Nz if span is None else span)
        q = 0.5 * rho_inf * U * U * cfg.chord * span
        fxb, fyb = -s[..., 0], -s[..., 1]
        cd = (fxb * ex + fyb * ey) / q
        cl = (-fxb * ey + fyb * ex) / q
        return cd, cl

    def _stream(self) -> C.c_void_p:
        return C.c_void_p(_lib.stream_of(self.device))


class FishIBM(DroneMovingIBM):
    """The fish-case clone (fish/ibm_fish.py:41-173); same kernels and surface."""
