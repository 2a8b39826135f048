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

Added readout with no reference counterpart: FP64 sums of the marker forces
(``force_sums``) and a device-side history of them for C_D / C_L.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from .config import CaseConfig, resolve, spatial_order
from .fields import Field, MarkerField


class DroneMovingIBM:
    """Direct-forcing IBM with the Peskin 4-point delta for a rigid body in heave + pitch.

    Options beyond the reference: ``spread_mode`` "fixed" (order-independent
    fixed-point sums, the default) or "atomic" (float atomics like
    ti.atomic_add); ``wrap_weight_quirk`` False to weight wrapped stencil
    points by their unwrapped offset; ``record_stencil`` to keep floor(pos)-1
    per marker (for index checks); ``order`` "natural" (the template's own
    order, neighbours along each section) or "morton" (a Z-order sort of the
    template) for the marker groups of the fused kernel; ``group`` markers
    per workgroup (4, 8 or 16).  None of these tuning options change results
    in the fixed spread mode."""

    def __init__(self, config=None, device=None, spread_mode: str = "fixed",
                 f_max=None, wrap_weight_quirk=None, record_stencil: bool = False,
                 order: str = "natural", group: int = 8):
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
                return MarkerField(torch.zeros(self.N, dtype=f32, device=dev))
            return MarkerField(torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev))

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
        self.fx = Field(shape, torch.zeros((self.Nz, self.Ny, self.Nx), dtype=f32, device=dev),
                        owner=self, box=box_ptr)
        self.fy = Field(shape, torch.zeros((self.Nz, self.Ny, self.Nx), dtype=f32, device=dev),
                        owner=self, box=box_ptr)
        self.fz = Field(shape, torch.zeros((self.Nz, self.Ny, self.Nx), dtype=f32, device=dev),
                        owner=self, box=box_ptr)
        self._acc = torch.zeros((2, ncell), dtype=torch.int64, device=dev)
        self.ip = _lib.IbmParams(float(cfg.f_max if f_max is None else f_max),
                                 int(cfg.wrap_weight_quirk if wrap_weight_quirk is None
                                     else wrap_weight_quirk),
                                 _lib.SPREAD_MODES[spread_mode], 0, int(group), 0)
        self.stencil_base = (torch.zeros((self.N, 3), dtype=torch.int32, device=dev)
                             if record_stencil else None)
        self._hist = None
        self._hist_n = 0
        self._sums = torch.zeros(2, dtype=torch.float64, device=dev)

    # -- reference surface ------------------------------------------------
    def compute_force(self, ux, uy, xc: float, yc: float, theta: float, vx: float, vy: float,
                      omega: float):
        """Markers, interpolation, marker forces, spreading and clamping
        (airfoil_moving.py:230-255).  ``ux``/``uy`` are the solver's velocity
        fields (read straight from its populations) or any [Nz,Ny,Nx] float32
        device Field/tensor.  Returns (fx, fy, fz), this object's own fields."""
        st = self._stream()
        p = lambda m: C.c_void_p(m.data_ptr())
        grid, lp, f, state, uxp, uyp = self._velocity_source(ux, uy)
        self.gen += 1
        boxp = C.c_void_p(self.box.data_ptr())
        acc0, acc1 = C.c_void_p(self._acc[0].data_ptr()), C.c_void_p(self._acc[1].data_ptr())
        _lib.call("ibm_compute_force", C.byref(grid), C.byref(self.ip), lp, C.c_void_p(f), state,
                  C.c_void_p(uxp), C.c_void_p(uyp), self.N,
                  C.c_void_p(self.order.data_ptr() if self.order is not None else 0),
                  p(self.X0), p(self.Y0),
                  p(self.Z0), float(xc), float(yc), float(theta), float(vx), float(vy),
                  float(omega), p(self.X_marker), p(self.Y_marker), p(self.Z_marker),
                  p(self.VX_body), p(self.VY_body), p(self.U_k), p(self.V_k), p(self.FX_k),
                  p(self.FY_k),
                  C.c_void_p(self.stencil_base.data_ptr() if self.stencil_base is not None else 0),
                  acc0, acc1, boxp, self.gen, st)
        _lib.call("ibm_clip_forces", C.byref(grid), C.byref(self.ip), acc0, acc1,
                  C.c_void_p(self.fx._t.data_ptr()), C.c_void_p(self.fy._t.data_ptr()), boxp,
                  self.gen, st)
        return self.fx, self.fy, self.fz

    def force_box(self):
        """The (x0, y0, z0, x1, y1, z1) box of the cells the last compute_force reached."""
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
        _lib.call("ibm_force_sum_f64", self.N, C.c_void_p(self.FX_k.data_ptr()),
                  C.c_void_p(self.FY_k.data_ptr()),
                  C.c_void_p(self.dV.data_ptr() if self.dV is not None else 0),
                  C.c_void_p(out.data_ptr()), self._stream())
        return out

    def record_force(self) -> None:
        """Append this step's FP64 force sums to the device history."""
        import torch
        if self._hist is None or self._hist_n == self._hist.shape[0]:
            cap = 1024 if self._hist is None else 2 * self._hist.shape[0]
            h = torch.zeros((cap, 2), dtype=torch.float64, device=self.device)
            if self._hist is not None:
                h[: self._hist_n] = self._hist[: self._hist_n]
            self._hist = h
        self.force_sums(self._hist[self._hist_n])
        self._hist_n += 1

    def force_history(self) -> np.ndarray:
        """(steps, 2) float64: per recorded step, sum_k FX_k dV_k and sum_k FY_k dV_k
        (the force the markers put on the fluid)."""
        if self._hist is None:
            return np.zeros((0, 2))
        return self._hist[: self._hist_n].cpu().numpy().copy()

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
        import torch
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)


class FishIBM(DroneMovingIBM):
    """The fish-case clone (fish/ibm_fish.py:41-173); same kernels and surface."""
