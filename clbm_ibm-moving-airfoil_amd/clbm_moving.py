"""CLBMSolver: the reference's solver object over the HIP kernels.

Mirrors ``CLBMSolver`` of clbm_moving.py:41-343: constructed without
arguments from the ``config_moving`` convention (or given a CaseConfig),
advanced by ``step(fx, fy, fz)``, read through ``f, rho, ux, uy, uz`` and
``M_inv`` with ``.to_numpy()``.

Differences of mechanism (not of results):
* one population buffer updated in place (AA pattern) instead of f/f_post;
  ``state_word`` records which of the two in-place layouts it holds (bit 0)
  and whether the inlet/outlet planes of the last step are still to be
  applied by the next reader (bit 1), which is how the reference's K5/K6
  run without a pass of their own;
* ``rho, ux, uy, uz`` are computed on demand (K7 of clbm_moving.py:305-321)
  when read, instead of after every step; the immersed-boundary object reads
  the force-free velocity straight from the populations;
* when the forces come from an immersed-boundary object that computed them on
  its own stream (``DroneMovingIBM(overlap=...)``), the step runs in two
  parts: the cells outside the body's region first, concurrently with the
  marker kernels (they touch no population the markers read), then the
  region with the force once the marker kernels are done.
"""

from __future__ import annotations

import ctypes as C
import json
import os
import weakref

import numpy as np

from . import _lib
from .config import CaseConfig, resolve
from .fields import Field, HostConstant, as_device_ptr

Q = 19


def _m_raw_inv_f32() -> np.ndarray:
    """M_raw^-1 in float32 (config_moving.py:62-73); exact dyadic entries."""
    ex = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0]
    ey = [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1]
    ez = [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1]
    pqr = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (2, 0, 0),
           (0, 2, 0), (0, 0, 2), (2, 1, 0), (1, 2, 0), (2, 0, 1), (0, 2, 1), (1, 0, 2), (0, 1, 2),
           (2, 2, 0), (2, 0, 2), (0, 2, 2)]
    m = np.array([[ex[c] ** p * ey[c] ** q * ez[c] ** r for c in range(Q)] for (p, q, r) in pqr],
                 dtype=np.float64)
    return np.linalg.inv(m).astype(np.float32)


def _stream_handle(device) -> C.c_void_p:
    return C.c_void_p(_lib.stream_of(device))


class PopulationField:
    """``CLBMSolver.f``: (Nx, Ny, Nz, 19) post-stream populations on the host side."""

    def __init__(self, solver: "CLBMSolver"):
        self._s = solver
        self.shape = (solver.Nx, solver.Ny, solver.Nz, Q)

    def to_numpy(self) -> np.ndarray:
        import torch
        s = self._s
        wide = s.storage == "f64"
        out = torch.empty((Q, s.Nz, s.Ny, s.Nx), dtype=torch.float64 if wide else torch.float32,
                          device=s.device)
        _lib.call("clbm_export_state_f64" if wide else "clbm_export_state", C.byref(s.grid),
                  C.byref(s.params), C.c_void_p(s.state.data_ptr()), s.state_word,
                  C.c_void_p(out.data_ptr()), s._stream())
        return np.ascontiguousarray(out.cpu().numpy().transpose(3, 2, 1, 0))

    def from_numpy(self, arr) -> None:
        import torch
        s = self._s
        wide = s.storage == "f64"
        a = np.asarray(arr, dtype=np.float64 if wide else np.float32)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != {self.shape}")
        src = torch.from_numpy(np.ascontiguousarray(a.transpose(3, 2, 1, 0))).to(s.device)
        s._wait_readers()
        _lib.call("clbm_import_state_f64" if wide else "clbm_import_state", C.byref(s.grid),
                  C.c_void_p(s.state.data_ptr()), C.c_void_p(src.data_ptr()), s._stream())
        s.state_word = 0
        s._macro_valid = False
        if s.device.type == "cuda":
            torch.cuda.current_stream(s.device).synchronize()


# inlet argument -> clbm_params_t.inlet (CLBM_INLET_*)
_INLETS = {False: 0, True: 1, "hecht-harting": 2}


class CLBMSolver:
    """D3Q19 cascaded LBM with Guo forcing on an Nx x Ny x Nz lattice (HIP kernels).

    Parameters mirror the reference's import-time constants; all are optional:
    ``config`` (CaseConfig or config module; default: ``sys.modules
    ['config_moving']`` or the reference airplane case), ``storage`` one of
    f32 (drop-in), f16, bf16, f16c (16-bit deviation storage) or f64 (error
    reference),
    ``inlet`` True for the reference's inlet rule at x = 0 and outlet copy at
    x = Nx-1, "hecht-harting" for the same outlet with the on-site velocity
    inlet of Hecht and Harting (J. Stat. Mech. 2010, P01018; not in the
    reference, opt-in), False for an x-periodic box, ``outside_lds`` the LDS bytes per
    workgroup the outside part of a split step reserves (a tuning choice that
    leaves each CU room for a concurrent marker workgroup; by default 33 KiB,
    4 workgroups per CU, for 4- and 8-byte populations, which are bound by
    HBM, and 27 KiB, 5 per CU, for the 16-bit formats, whose step kernel also
    leans on instruction issue), ``deviation`` what 16-bit
    storage subtracts before rounding: "weights" f_q - w_q, or "inflow"
    f_q - feq_q(1, inlet velocity), which is near zero wherever the flow is
    near the free stream.

    ``fused_split`` (attribute, default True): a split step runs as one
    launch whose last workgroups wait on the device for the body's published
    forces (clbm_step_fused, waiting at most FUSED_WAIT_MS) instead of two
    launches around a join of the marker stream (clbm_step_split); same
    results.  Used only when the marker stream passed the check that it runs
    beside the solver's stream (DroneMovingIBM.side_concurrent): where kernels
    run one at a time (a profiler collecting counters) no kernel can wait for
    another, and the two launches are used.  "joined": the one launch after a
    join of the marker stream, whatever the check said (for measuring that
    kernel where kernels run one at a time)."""

    # the longest a one-launch split step's region waits for its forces
    FUSED_WAIT_MS = 2000.0

    def __init__(self, config=None, storage: str = "f32", device=None, inlet=True,
                 outside_lds=None, deviation: str = "weights"):
        import torch
        cfg: CaseConfig = resolve(config)
        self.cfg = cfg
        self.Nx, self.Ny, self.Nz = int(cfg.Nx), int(cfg.Ny), int(cfg.Nz)
        self.device = torch.device(device if device is not None else "cuda")
        if storage not in _lib.STORAGE:
            raise ValueError(f"storage must be one of {sorted(_lib.STORAGE)}")
        self.storage = storage
        _lib.load()
        if deviation not in ("weights", "inflow"):
            raise ValueError("deviation must be 'weights' or 'inflow'")
        self.deviation = deviation
        if inlet not in _INLETS:
            raise ValueError("inlet must be True, False or 'hecht-harting'")
        self.inlet = inlet
        du, dv = (float(cfg.ux_in), float(cfg.uy_in)) if deviation == "inflow" else (0.0, 0.0)
        self.grid = _lib.ClbmGrid(self.Nx, self.Ny, self.Nz, _lib.STORAGE[storage], du, dv)
        if outside_lds is None:
            outside_lds = 27648 if _lib.ELEM_BYTES[self.grid.storage] == 2 else 33792
        self.params = _lib.ClbmParams(float(cfg.s_v), float(cfg.s_e), float(cfg.ux_in),
                                      float(cfg.uy_in), _INLETS[inlet], int(outside_lds))
        nbytes = Q * self.Nx * self.Ny * self.Nz * _lib.ELEM_BYTES[self.grid.storage]
        self.state = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        self.state_word = 0
        # ctypes arguments that stay fixed for this solver
        self._gref, self._pref = C.byref(self.grid), C.byref(self.params)
        self._state_ptr = C.c_void_p(self.state.data_ptr())
        self.steps = 0
        shape = (self.Nx, self.Ny, self.Nz)
        self._macro = None
        self._macro_valid = False
        self.rho = Field(shape, None, refresh=self._refresh_macro, owner=self)
        self.ux = Field(shape, None, refresh=self._refresh_macro, owner=self)
        self.uy = Field(shape, None, refresh=self._refresh_macro, owner=self)
        self.uz = Field(shape, None, refresh=self._refresh_macro, owner=self)
        self.f = PopulationField(self)
        self.M_inv = HostConstant(_m_raw_inv_f32())
        # immersed-boundary objects reading the populations on their own
        # stream: {id: (ibm, region int32[6], event after their last read)}
        self._readers = {}
        # (ibm, stream): the immersed-boundary object whose fork event completes
        # with this solver's last kernel, the inside part of a split step queued
        # on that stream; its next compute_force then needs no record of the
        # event (None after anything else that writes the populations)
        self._fork = None
        # the immersed-boundary objects whose forces a split step has used:
        # health() checks that their stencils stayed in the predicted regions
        self._split_ibms = weakref.WeakSet()
        # optional (start, end) events recorded around the first kernel of the
        # next step (the whole step, or its outside part); cells that kernel updated
        self.kernel_events = None
        self.last_kernel_cells = 0
        # how the last step ran: "full" (one launch over the lattice), "split"
        # (outside part, join, inside part), "fused" (both parts in one launch)
        # or "coupled" (coupled.step_coupled)
        self.last_step_kind = None
        # a learned closure (closure.ClosureNet) replaces the plain cascade when set
        self.closure = None
        self.fused_split = True
        self._init_feq()

    # -- reference surface ------------------------------------------------
    def _init_feq(self) -> None:
        """K1 (clbm_moving.py:75-91): equilibrium at the inlet velocity, rho = 1."""
        self._wait_readers()
        _lib.call("clbm_init_feq", C.byref(self.grid), C.c_void_p(self.state.data_ptr()), 1.0,
                  float(self.cfg.ux_in), float(self.cfg.uy_in), 0.0, self._stream())
        self.state_word = 0
        self._macro_valid = False

    def step(self, fx=None, fy=None, fz=None) -> None:
        """Advance one step with body force (fx, fy, fz) (clbm_moving.py:326-343)."""
        import torch
        split = self._split_source(fx, fy, fz) if self.closure is None else None
        ev = self.kernel_events
        self.kernel_events = None
        self._fork = None
        st_h = _lib.stream_of(self.device)
        st = C.c_void_p(st_h)
        state = self._state_ptr
        if split is not None and (ev is None or (ev[0].cuda_event and ev[1].cuda_event)):
            # part 0, the join of the marker stream and part 1 in one call
            ibm, region = split
            self._split_ibms.add(ibm)
            fxp, fyp, boxp = ibm._force_ptrs
            e0 = ev[0].cuda_event if ev is not None else None
            e1 = ev[1].cuda_event if ev is not None else None
            # the body's fork event completes with the step's last kernel (the
            # next marker kernels read the populations it writes); not while a graph
            # is being captured, where launches carry no events and the marker stream
            # must rejoin this one before the capture ends (the two launches' join)
            capturing = torch.cuda.is_current_stream_capturing()
            done = None if capturing else ibm._ev[0]
            joined = self.fused_split == "joined"
            if (ibm._published == ibm.gen and (joined or (self.fused_split and ibm.side_concurrent))
                    and not capturing):
                # one launch; its region waits on the device for the published forces
                # (only where the marker stream was seen running beside this one: a
                # profiler that serialises kernels fails that check at its creation),
                # or finds them published at once after a join
                if joined:
                    ibm._join()
                _lib.check(_lib.fn("clbm_step_fused")(
                    self._gref, self._pref, state, self.state_word, region, fxp, fyp,
                    ibm._box_ptr, ibm.gen, self.FUSED_WAIT_MS, st, e0, e1, done),
                    "clbm_step_fused")
                fused = True
            else:
                _lib.check(_lib.fn("clbm_step_split")(
                    self._gref, self._pref, state, self.state_word, region, fxp, fyp, boxp, st,
                    ibm._side_h, ibm._ev[1], e0, e1, done), "clbm_step_split")
                fused = False
            ibm._pending = False
            if done is not None:
                self._fork = (ibm, st_h)
                ibm._fork_void = False
            self._readers.clear()
            r = list(region)
            self.last_kernel_cells = self.ncell if fused else self.ncell - (
                (r[3] - r[0] + 1) * (r[4] - r[1] + 1) * (r[5] - r[2] + 1))
            self.last_step_kind = "fused" if fused else "split"
        elif split is not None:
            ibm, region = split
            self._split_ibms.add(ibm)
            part = _lib.fn("clbm_step_part")
            if ev is not None:
                ev[0].record(torch.cuda.current_stream(self.device))
            _lib.check(part(self._gref, self._pref, state, self.state_word, None, None, None,
                            None, region, 0, st), "clbm_step_part")
            if ev is not None:
                ev[1].record(torch.cuda.current_stream(self.device))
            ibm._join()
            self._readers.clear()
            fxp, fyp, boxp = ibm._force_ptrs
            _lib.check(part(self._gref, self._pref, state, self.state_word, fxp, fyp, None, boxp,
                            region, 1, st), "clbm_step_part")
            r = list(region)
            self.last_kernel_cells = self.ncell - ((r[3] - r[0] + 1) * (r[4] - r[1] + 1)
                                                   * (r[5] - r[2] + 1))
            self.last_step_kind = "split"
        else:
            self._wait_readers()
            shape = (self.Nx, self.Ny, self.Nz)
            box = 0
            for fld in (fx, fy):
                if isinstance(fld, Field) and fld.box is not None:
                    box = fld.box
            fzp = as_device_ptr(fz, shape)
            if isinstance(fz, Field) and fz.box is not None:
                fzp = 0     # the immersed-boundary fz is identically zero (airfoil_moving.py:225)
            fxp, fyp = as_device_ptr(fx, shape), as_device_ptr(fy, shape)
            if ev is not None:
                ev[0].record(torch.cuda.current_stream(self.device))
            if self.closure is None:
                _lib.call("clbm_step", C.byref(self.grid), C.byref(self.params), state,
                          self.state_word, C.c_void_p(fxp), C.c_void_p(fyp), C.c_void_p(fzp),
                          C.c_void_p(box), st)
            else:
                net = self.closure
                _lib.call("clbm_step_closure", C.byref(self.grid), C.byref(self.params), state,
                          self.state_word, C.c_void_p(fxp), C.c_void_p(fyp), C.c_void_p(fzp),
                          C.c_void_p(box), C.c_void_p(net.packed.data_ptr()),
                          C.c_void_p(net.params.data_ptr()), net.lo, net.hi, st)
            if ev is not None:
                ev[1].record(torch.cuda.current_stream(self.device))
            self.last_kernel_cells = self.ncell
            self.last_step_kind = "full"
        self.state_word = _lib.fn("clbm_state_after_step")(self.state_word, self._pref)
        self.steps += 1
        self._macro_valid = False

    # -- checkpoint / resume (no reference counterpart: the reference saves
    #    only macroscopic frames, main_moving.py:148-155, and cannot resume) --
    CHECKPOINT_VERSION = 1

    def _checkpoint_meta(self) -> dict:
        g, p = self.grid, self.params
        return {"version": self.CHECKPOINT_VERSION, "grid": [self.Nx, self.Ny, self.Nz],
                "storage": self.storage, "deviation_velocity": [float(g.dev_ux), float(g.dev_uy)],
                "inlet": int(p.inlet), "s_v": float(p.s_v), "s_e": float(p.s_e),
                "ux_in": float(p.ux_in), "uy_in": float(p.uy_in)}

    def save_state(self, path: str, **extra) -> None:
        """Write the populations as they are held (the raw in-place buffer of
        this storage format, so that 16-bit codes are kept exactly), the state
        word and the step count to ``path`` (an .npz file, replaced
        atomically), with ``extra`` arrays beside them.  A solver of the same
        grid, storage, deviation base, inlet and rates that loads it continues
        bit for bit."""
        import torch
        self._wait_readers()
        raw = self.state.cpu().numpy()
        tmp = f"{path}.tmp{os.getpid()}"
        with open(tmp, "wb") as fh:
            np.savez(fh, state=raw, state_word=np.int32(self.state_word),
                     steps=np.int64(self.steps), meta=json.dumps(self._checkpoint_meta()),
                     **{k: np.asarray(v) for k, v in extra.items()})
        os.replace(tmp, path)
        torch.cuda.current_stream(self.device).synchronize()

    def load_state(self, path: str) -> dict:
        """Load what ``save_state`` wrote (ValueError when the file is for
        another grid, storage, deviation base, inlet or rates); returns its
        extra arrays."""
        import torch
        with np.load(path, allow_pickle=False) as d:
            meta = json.loads(str(d["meta"]))
            mine = self._checkpoint_meta()
            for k, v in mine.items():
                if meta.get(k) != v:
                    raise ValueError(f"checkpoint {path}: {k} is {meta.get(k)!r}, "
                                     f"this solver's {v!r}")
            raw = d["state"]
            word = int(d["state_word"])
            steps = int(d["steps"])
            extra = {k: d[k] for k in d.files if k not in ("state", "state_word", "steps", "meta")}
        if raw.dtype != np.uint8 or raw.size != self.state.numel():
            raise ValueError(f"checkpoint {path}: population buffer of {raw.size} bytes, "
                             f"this solver holds {self.state.numel()}")
        if word & ~(_lib.STATE_PARITY | _lib.STATE_PLANES) or steps < 0:
            raise ValueError(f"checkpoint {path}: bad state word {word} or step count {steps}")
        self._wait_readers()
        self.state.copy_(torch.from_numpy(raw).to(self.device))
        self.state_word = word
        self.steps = steps
        self._macro_valid = False
        torch.cuda.current_stream(self.device).synchronize()
        return extra

    # -- concurrent readers (immersed-boundary objects on their own stream) --
    def _add_reader(self, ibm, region) -> None:
        self._readers[id(ibm)] = (ibm, region)

    def _wait_readers(self) -> None:
        """Order the current stream after every pending concurrent reader
        (before anything that writes the populations)."""
        self._fork = None
        if not self._readers:
            return
        for ibm, _ in self._readers.values():
            ibm._join()
        self._readers.clear()

    def _split_source(self, fx, fy, fz):
        """(ibm, region) when (fx, fy, fz) are the fields of the one
        immersed-boundary object reading this solver on its own stream."""
        if len(self._readers) != 1:
            return None
        ibm, region = next(iter(self._readers.values()))
        if fx is not ibm.fx or fy is not ibm.fy or (fz is not None and fz is not ibm.fz):
            return None
        return ibm, region

    @property
    def parity(self) -> int:
        """Number of steps taken mod 2 (which AA layout the buffer holds)."""
        return self.state_word & _lib.STATE_PARITY

    @property
    def f_post(self):
        raise AttributeError("the in-place (AA) population layout keeps no separate post-collision "
                             "buffer; read CLBMSolver.f for the post-stream state")

    # -- readout ------------------------------------------------------------
    def _refresh_macro(self) -> None:
        import torch
        if self._macro_valid:
            return
        if self._macro is None:
            self._macro = torch.empty((4, self.Nz, self.Ny, self.Nx), dtype=torch.float32,
                                      device=self.device)
            for i, fld in enumerate((self.rho, self.ux, self.uy, self.uz)):
                fld._t = self._macro[i]
        m = self._macro
        _lib.call("clbm_macro", C.byref(self.grid), C.byref(self.params),
                  C.c_void_p(self.state.data_ptr()), self.state_word,
                  C.c_void_p(m[0].data_ptr()), C.c_void_p(m[1].data_ptr()),
                  C.c_void_p(m[2].data_ptr()), C.c_void_p(m[3].data_ptr()), self._stream())
        self._macro_valid = True

    def macro_into(self, out) -> None:
        """K7 (clbm_moving.py:305-321) into a caller's float32 device tensor
        [4, Nz, Ny, Nx] (rho, ux, uy, uz), queued on the current stream: a
        snapshot that later steps do not overwrite."""
        import torch
        if (not isinstance(out, torch.Tensor) or out.dtype != torch.float32 or not out.is_cuda
                or tuple(out.shape) != (4, self.Nz, self.Ny, self.Nx) or not out.is_contiguous()):
            raise TypeError("out must be a contiguous float32 device tensor [4, Nz, Ny, Nx]")
        _lib.call("clbm_macro", C.byref(self.grid), C.byref(self.params),
                  C.c_void_p(self.state.data_ptr()), self.state_word,
                  C.c_void_p(out[0].data_ptr()), C.c_void_p(out[1].data_ptr()),
                  C.c_void_p(out[2].data_ptr()), C.c_void_p(out[3].data_ptr()), self._stream())

    def health(self):
        """(rho_min, all_finite) of the current state, computed on the device
        (the divergence guard of main_moving.py:143-145).  Also checks, with
        the same synchronisation, that the stencils of every immersed-boundary
        object whose forces a split step used stayed in the region predicted
        for it (DroneMovingIBM.check_region raises ClbmError otherwise: the
        marker kernels then raced the lattice update away from the body)."""
        import torch
        out = torch.empty(2, dtype=torch.float32, device=self.device)
        _lib.call("clbm_health", C.byref(self.grid), C.byref(self.params),
                  C.c_void_p(self.state.data_ptr()), self.state_word, C.c_void_p(out.data_ptr()),
                  self._stream())
        h = out.cpu().numpy()
        for ibm in list(self._split_ibms):
            ibm.check_region()
        return float(h[0]), bool(h[1] > 0.5)

    def _stream(self) -> C.c_void_p:
        return _stream_handle(self.device)

    @property
    def ncell(self) -> int:
        return self.Nx * self.Ny * self.Nz
