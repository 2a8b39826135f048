"""Batched sweeps: many independent moving-airfoil instances advanced together.

No reference counterpart: the reference runs one case per process
(main_moving.py).  A sweep holds n instances, each a CLBMSolver and a
DroneMovingIBM of its own grid, rates and kinematics.  All of them advance
in lockstep, two launches per step for a whole group:

* ``ibm_batch_markers``: the marker kernel of every body, its stencils
  checked against the region that the body's table predicts for the step
  (``ibm_step_region``, as for the split step), and the next value of the
  device step counter;
* ``clbm_batch_step_acc``: the fused step of every lattice, with the force
  decoded and clamped straight from the body's fixed-point accumulators
  inside that region (the clamp needs no launch; the bodies' fields fx, fy
  and force box are rebuilt from the markers when read).  A run raises
  ClbmError when some body's stencils left their region.

With the float-atomic spread mode (or ``fused_clamp=False``) the pair is
``ibm_batch_compute`` (markers, then a clamp kernel into the fields, whose
last workgroup bumps the counter) and ``clbm_batch_step``.

The instances are grouped so that a group's populations fit in the GPU's
last-level cache, and a run advances one group for all its steps before the
next, so a group is read from that cache rather than HBM from step to step
(``cache_bytes`` 0 keeps one group of everything).

Kinematics come from per-instance device tables indexed by that counter.
The step sequence therefore needs no host argument and is captured once as a
graph of two steps (both AA parities), which is replayed.  Every instance
computes exactly what it computes when run on its own (fixed-point
spreading), which ``sequential`` checks and times.

``autotune`` picks the lattice workgroup size, the markers per workgroup and
the LDS each lattice workgroup reserves (which caps the workgroups per CU)
by timing the candidates on the batch itself, then restores the initial
state.  The choice is cached per batch signature in a JSON file.

Usage: python -m <package>.batch [--instances 64] [--steps 5000] [--seed 0] [--json out.json]
"""

from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import json
import math
import os
import sys
import time

import numpy as np

from . import _lib
from . import config as cfgmod
from .airfoil_moving import DroneMovingIBM
from .clbm_moving import CLBMSolver

GRIDS = [(200, 64, 16), (300, 100, 24), (450, 150, 36), (600, 200, 50)]
# (lattice block, markers per group, LDS bytes each lattice workgroup reserves: 24 KiB caps a
# CU at 6 workgroups of 256, 40 KiB at 4)
CANDIDATES = [(256, 8, 0), (256, 4, 0), (128, 8, 0), (128, 4, 0),
              (256, 8, 24576), (256, 4, 24576), (256, 8, 40960), (256, 4, 40960)]
TUNE_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "clbm_amd_batch_tune.json")


def draw_instances(n: int, seed: int = 0, grids=GRIDS):
    """Seeded instances: a grid from ``grids``, chord Ny/2.5, Re in [100, 400],
    heave in [0.1c, 0.5c], pitch amplitude in [0, 20] deg, phase in [0, 180] deg,
    reduced frequency in [0.25, 1], mean angle of attack in [0, 12] deg."""
    out = []
    for i in range(n):
        rng = np.random.default_rng(seed * 100003 + i)
        nx, ny, nz = grids[int(rng.integers(len(grids)))]
        chord = ny / 2.5
        out.append(cfgmod.naca_case(
            nx, ny, nz, chord, Re=float(rng.uniform(100, 400)), k=float(rng.uniform(0.25, 1.0)),
            heave_frac=float(rng.uniform(0.1, 0.5)), aoa_mean_deg=float(rng.uniform(0, 12)),
            pitch_amp_deg=float(rng.uniform(0, 20)), phase_deg=float(rng.uniform(0, 180)),
            le_x=nx / 4.0, name=f"inst{i}-{nx}x{ny}x{nz}"))
    return out


def kinematics_table(cfg, first_step: int, steps: int) -> np.ndarray:
    """(steps, 7) float32: xc, yc, cos, sin, vx, vy, omega for steps first_step,
    first_step+1, ...  Each entry is what ibm_compute_force receives for that
    step: float32 arguments, cos/sin of the float32 angle in float64 rounded once."""
    t = np.zeros((steps, 7), dtype=np.float32)
    for j in range(steps):
        xc, yc, th, vx, vy, om = cfg.motion_state(first_step + j)
        a = np.float64(np.float32(th))
        t[j] = (np.float32(xc), np.float32(yc), np.float32(math.cos(a)), np.float32(math.sin(a)),
                np.float32(vx), np.float32(vy), np.float32(om))
    return t


def region_table(lbm, ibm, cfg, first_step: int, steps: int) -> np.ndarray:
    """(steps, 6) int32: for the steps first_step .. first_step + steps - 1 of
    cfg's motion, the region (x0, y0, z0, x1, y1, z1) that ibm_step_region
    predicts for the body's stencils (the whole lattice when it predicts
    none: the body off the grid or over half of it)."""
    out = np.empty((steps, 6), dtype=np.int32)
    r = (C.c_int32 * 6)()
    fn = _lib.fn("ibm_step_region")
    whole = (0, 0, 0, lbm.Nx - 1, lbm.Ny - 1, lbm.Nz - 1)
    for s in range(steps):
        xc, yc, th = cfg.motion_state(first_step + s)[:3]
        rc = fn(C.byref(lbm.grid), C.byref(lbm.params), *ibm._hull_args, float(xc), float(yc),
                float(th), r)
        if rc < 0:
            _lib.check(rc, "ibm_step_region")
        out[s] = tuple(r) if rc == 1 else whole
    return out


def cache_groups(cfgs, storage: str, cache_bytes: int):
    """Partition instance indices into groups whose populations together fit
    in ``cache_bytes`` (first fit, largest first); an instance larger than the
    budget forms a group of its own, and ``cache_bytes`` 0 keeps one group."""
    idx = list(range(len(cfgs)))
    if cache_bytes <= 0:
        return [idx]
    e = _lib.ELEM_BYTES[_lib.STORAGE[storage]]
    size = [19 * e * c.Nx * c.Ny * c.Nz for c in cfgs]
    groups, load = [], []
    for k in sorted(idx, key=lambda k: -size[k]):
        for g, used in enumerate(load):
            if used + size[k] <= cache_bytes:
                groups[g].append(k)
                load[g] += size[k]
                break
        else:
            groups.append([k])
            load.append(size[k])
    return [sorted(g) for g in groups]


class _Group:
    """Device tables and captured graph of the instances of one cache group.
    A ``solo`` group (one lattice large enough that its step is bound by HBM
    on its own) has neither: its instance steps through its own objects."""

    def __init__(self, idx, solo: bool = False):
        self.idx = idx
        self.solo = solo
        self.lb = self.ib = self.lwork = self.iwork = None
        self.graph = None
        self.graph_state = None


class BatchSweep:
    """n instances on one GPU, stepped together (see the module docstring).

    The instances are split into groups that fit in ``cache_bytes`` (by
    default 192 MiB of the MI355X's 256 MB last-level cache) and ``run``
    advances one group at a time for all its steps, so a group's populations
    stay in that cache from step to step.

    A group of one lattice of ``solo_min_cells`` cells or more (by default
    the size from which DroneMovingIBM overlaps its marker kernels with the
    step) is stepped through that instance's own solver and IBM objects:
    the split step with the marker kernels on a second stream beside the
    lattice away from the body, which batching cannot improve on for a
    lattice that fills the GPU by itself.  Results are the same bit for bit."""

    def __init__(self, cfgs, storage: str = "f32", device=None, steps: int = 1000,
                 block: int = 256, group: int = 8, first_step: int = 1,
                 cache_bytes: int = 192 << 20, deviation: str = "weights",
                 solo_min_cells: int = DroneMovingIBM.OVERLAP_MIN_CELLS, inlet=True,
                 lanes: int = 1, lds: int = 0, fused_clamp: bool = True):
        import torch
        self.cfgs = list(cfgs)
        self.device = torch.device(device if device is not None else "cuda")
        self.storage = storage
        self.deviation = deviation
        self.first_step = first_step
        self.capacity = int(steps)
        self.lbms = [CLBMSolver(c, storage=storage, device=self.device, deviation=deviation,
                                inlet=inlet) for c in self.cfgs]
        self.ibms = [DroneMovingIBM(c, device=self.device, overlap=False) for c in self.cfgs]
        self._kin = [torch.from_numpy(kinematics_table(c, first_step, self.capacity)).to(self.device)
                     for c in self.cfgs]
        self.steps_done = 0
        self.state = 0
        if int(lanes) < 1:
            raise ValueError("lanes must be >= 1")
        self.lanes = int(lanes)
        self._lanes = None
        # two launches per step (markers; lattice with the clamp folded in)
        # instead of three, with fixed-point spreading; the batched bodies'
        # fields fx, fy are then rebuilt when read
        self.fused_clamp = bool(fused_clamp) and \
            self.ibms[0].ip.spread_mode == _lib.SPREAD_FIXED
        cells = [c.Nx * c.Ny * c.Nz for c in self.cfgs]
        # the groups in flight together share the cache budget
        self.groups = [_Group(g, solo=len(g) == 1 and cells[g[0]] >= solo_min_cells)
                       for g in cache_groups(self.cfgs, storage, cache_bytes // self.lanes)]
        self._solo = set()
        for gr in self.groups:
            if gr.solo:
                self._solo.add(gr.idx[0])
                self.ibms[gr.idx[0]].overlap = True
        # per batched body and step of the kinematics table, the region its
        # stencils stay in (predicted as for the split step); the fused
        # lattice step reads the force inside it
        self._regions = [None if n in self._solo or not self.fused_clamp else
                         torch.from_numpy(region_table(self.lbms[n], self.ibms[n], self.cfgs[n],
                                                       first_step, self.capacity)).to(self.device)
                         for n in range(len(self.cfgs))]
        self.build(block, group, lds)

    # -- tables ---------------------------------------------------------------
    def build(self, block: int, group: int, lds: int = 0) -> None:
        import torch
        self.block, self.group, self.lds = int(block), int(group), int(lds)
        first = self.ibms[0].ip
        for ibm in self.ibms[1:]:
            if (ibm.ip.f_max, ibm.ip.wrap_weight_quirk, ibm.ip.spread_mode) != \
                    (first.f_max, first.wrap_weight_quirk, first.spread_mode):
                raise ValueError("instances of a batch share the spread mode, clamp and quirk")
        torch.cuda.synchronize(self.device)
        for gr in self.groups:
            if not gr.solo:
                self._build_group(gr)

    def _build_group(self, gr: "_Group") -> None:
        import torch
        lbms = [self.lbms[k] for k in gr.idx]
        ibms = [self.ibms[k] for k in gr.idx]
        n = len(gr.idx)
        grids = (_lib.ClbmGrid * n)(*[l.grid for l in lbms])
        params = (_lib.ClbmParams * n)(*[l.params for l in lbms])
        f = (C.c_void_p * n)(*[l.state.data_ptr() for l in lbms])
        fx = (C.c_void_p * n)(*[i.fx._t.data_ptr() for i in ibms])
        fy = (C.c_void_p * n)(*[i.fy._t.data_ptr() for i in ibms])
        boxes = (C.c_void_p * n)(*[i.box.data_ptr() for i in ibms])
        lib = _lib.load()
        nb = lib.clbm_batch_bytes(n, grids, self.block)
        if nb == 0:
            raise _lib.ClbmError(lib.clbm_last_error().decode())
        gr.lwork = torch.empty(nb, dtype=torch.uint8, device=self.device)
        gr.lb = _lib.ClbmBatch()
        torch.cuda.synchronize(self.device)
        _lib.call("clbm_batch_build", n, grids, params, f, fx, fy, boxes, self.block,
                  C.c_void_p(gr.lwork.data_ptr()), nb, C.byref(gr.lb))
        gr.lb.lds = self.lds
        bufs = (_lib.IbmBuffers * n)()
        for k, ibm in enumerate(ibms):
            b = bufs[k]
            b.n = ibm.N
            b.order = ibm.order.data_ptr() if ibm.order is not None else None
            for name, t in (("X0", ibm.X0), ("Y0", ibm.Y0), ("Z0", ibm.Z0), ("X", ibm.X_marker),
                            ("Y", ibm.Y_marker), ("Z", ibm.Z_marker), ("VX", ibm.VX_body),
                            ("VY", ibm.VY_body), ("U", ibm.U_k), ("V", ibm.V_k),
                            ("FX", ibm.FX_k), ("FY", ibm.FY_k)):
                setattr(b, name, t._t.data_ptr())
            b.acc_x, b.acc_y = ibm._acc[0].data_ptr(), ibm._acc[1].data_ptr()
            b.box = ibm.box.data_ptr()
            b.fx, b.fy = ibm.fx._t.data_ptr(), ibm.fy._t.data_ptr()
            b.kin = self._kin[gr.idx[k]].data_ptr()
            b.gen0 = ibm.gen - self.steps_done    # the generation of counter value c is gen0 + 1 + c
            b.regions = self._regions[gr.idx[k]].data_ptr() if self.fused_clamp else None
        first = self.ibms[0].ip
        ip = _lib.IbmParams(float(first.f_max), int(first.wrap_weight_quirk),
                            int(first.spread_mode), 0, self.group, 0)
        ib = lib.ibm_batch_bytes(n, bufs, self.group)
        if ib == 0:
            raise _lib.ClbmError(lib.clbm_last_error().decode())
        gr.iwork = torch.empty(ib, dtype=torch.uint8, device=self.device)
        gr.ib = _lib.IbmBatch()
        _lib.call("ibm_batch_build", n, grids, params, f, C.byref(ip), bufs,
                  C.c_void_p(gr.iwork.data_ptr()), ib, C.byref(gr.ib))
        # the counter starts at 0: move it to the steps already taken (with
        # region tables a step of AA parity p reads counter[p] and writes
        # counter[1 - p]; both start at the steps taken, whatever the parity)
        off = (gr.ib.counter - gr.iwork.data_ptr()) // 4
        cnt = gr.iwork.view(torch.int32)
        cnt[off:off + (2 if self.fused_clamp else 1)] = self.steps_done
        gr.graph = None

    # -- stepping -------------------------------------------------------------
    def _next_state(self, state: int) -> int:
        return _lib.load().clbm_state_after_step(state, C.byref(self.lbms[0].params))

    def _launch_step(self, gr: "_Group", state: int, stream) -> int:
        if self.fused_clamp:
            _lib.call("ibm_batch_markers", C.byref(gr.ib), state, stream)  # advances the counter
            _lib.call("clbm_batch_step_acc", C.byref(gr.lb), C.byref(gr.ib), state, stream)
        else:
            _lib.call("ibm_batch_compute", C.byref(gr.ib), state, stream)  # advances the counter
            _lib.call("clbm_batch_step", C.byref(gr.lb), state, stream)
        return self._next_state(state)

    def _solo_steps(self, gr: "_Group", k: int) -> None:
        """k steps of a solo group's instance through its own objects (on the
        current stream, with its marker kernels on their own)."""
        i = gr.idx[0]
        lbm, ibm, cfg = self.lbms[i], self.ibms[i], self.cfgs[i]
        for j in range(k):
            fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy,
                                           *cfg.motion_state(self.first_step + self.steps_done + j))
            lbm.step(fx, fy, fz)

    def _advance_objects(self, k: int) -> None:
        """Keep the solver/IBM objects consistent with k batched steps (solo
        instances keep their own count)."""
        self.steps_done += k
        for n, l in enumerate(self.lbms):
            if n not in self._solo:
                l.state_word = self.state
                l._fork = None
                l.steps += k
                l._macro_valid = False
        for n, i in enumerate(self.ibms):
            if n not in self._solo:
                i.gen += k
                if self.fused_clamp and k:
                    i._fields_stale = True

    def step(self, k: int = 1) -> None:
        """k eager steps of the whole batch, group by group."""
        import torch
        if self.steps_done + k > self.capacity:
            raise ValueError("kinematics table exhausted; build the sweep with more steps")
        st = C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        for gr in self.groups:
            if gr.solo:
                self._solo_steps(gr, k)
                continue
            s = self.state
            for _ in range(k):
                s = self._launch_step(gr, s, st)
        self.state = self._state_after(k)
        self._advance_objects(k)
        self.check_regions()

    def run(self, k: int, graph: bool = True, times=None) -> None:
        """k steps, group by group; with ``graph``, a group's steps after the
        first go through its captured two-step graph (the inlet/outlet planes
        are pending from step 2 on, so the state word cycles with period two).
        With ``lanes`` > 1, that many groups advance at once, each on a stream
        of its own, their launches interleaved step by step (or pair by pair
        of graph steps), so that one group's marker kernels run beside
        another's lattice update.  ``times``: a list that receives (group
        instance indices, start event, end event) per group, events on the
        group's stream around its work (read them after a synchronisation)."""
        import torch
        if not graph:
            self.step(k)
            return
        if self.steps_done + k > self.capacity:
            raise ValueError("kinematics table exhausted; build the sweep with more steps")
        if self.steps_done == 0 and k > 0:
            self.step(1)
            k -= 1
        main = torch.cuda.current_stream(self.device)
        if self.lanes > 1 and self._lanes is None:
            self._lanes = [torch.cuda.Stream(device=self.device) for _ in range(self.lanes)]
        order = [g for g in self.groups if g.solo] + [g for g in self.groups if not g.solo]
        for c in range(0, len(order), self.lanes):
            chunk = order[c:c + self.lanes]
            streams = [main] if self.lanes == 1 else self._lanes[:len(chunk)]
            for st in streams:
                if st is not main:
                    st.wait_stream(main)
            work = [self._units(gr, k, st, times) for gr, st in zip(chunk, streams)]
            while work:
                for w in list(work):
                    try:
                        next(w)
                    except StopIteration:
                        work.remove(w)
            for st in streams:
                if st is not main:
                    main.wait_stream(st)
        self.state = self._state_after(k)
        self._advance_objects(k)
        self.check_regions()

    def check_regions(self) -> None:
        """Raise ClbmError when a body's stencils left the region that its
        step read the force from (the split step of a solo group, or the
        region table of a batched body); one synchronisation."""
        import torch
        checked = [n for n in range(len(self.ibms)) if n in self._solo or self.fused_clamp]
        if not checked:
            return
        for n in checked:
            self.ibms[n]._join()
        flags = torch.stack([self.ibms[n].box[_lib.REGION_FLAG_OFFSET] for n in checked]).cpu()
        for n, g in zip(checked, flags.tolist()):
            if not g:
                continue
            if n in self._solo:
                self.ibms[n].check_region()
            raise _lib.ClbmError(
                f"instance {n}: the stencils of generation {g} left the region of its "
                f"step's table; its lattice step missed part of the force")

    def _units(self, gr: "_Group", k: int, stream, times):
        """Issue a group's k steps on ``stream``, yielding after each unit of
        work (a step, or a replay of the captured pair of steps)."""
        import torch
        if times is not None:
            ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            ev[0].record(stream)
            times.append((gr.idx, ev[0], ev[1]))
        if gr.solo:
            i = gr.idx[0]
            lbm, ibm, cfg = self.lbms[i], self.ibms[i], self.cfgs[i]
            for j in range(k):
                with torch.cuda.stream(stream):
                    fx, fy, fz = ibm.compute_force(
                        lbm.ux, lbm.uy, *cfg.motion_state(self.first_step + self.steps_done + j))
                    lbm.step(fx, fy, fz)
                yield
        else:
            st = C.c_void_p(stream.cuda_stream)
            s, left = self.state, k
            if left >= 2 and gr.graph is not None and gr.graph_state != s:
                s = self._launch_step(gr, s, st)   # realign with the captured pair
                left -= 1
                yield
            if left >= 2 and gr.graph is None:
                with torch.cuda.stream(stream):
                    self._capture(gr, s)
            while left >= 2:
                with torch.cuda.stream(stream):
                    gr.graph.replay()
                s = self._next_state(self._next_state(s))
                left -= 2
                yield
            while left:
                s = self._launch_step(gr, s, st)
                left -= 1
                yield
        if times is not None:
            ev[1].record(stream)

    def _state_after(self, k: int) -> int:
        s = self.state
        for _ in range(k):
            s = self._next_state(s)
        return s

    def _capture(self, gr: "_Group", state: int) -> None:
        import torch
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                st = C.c_void_p(side.cuda_stream)
                s = self._launch_step(gr, state, st)
                self._launch_step(gr, s, st)
        torch.cuda.current_stream(self.device).wait_stream(side)
        gr.graph = g
        gr.graph_state = state

    # -- tuning and reference timings -------------------------------------------
    def signature(self) -> str:
        key = json.dumps([[c.Nx, c.Ny, c.Nz, c.N_MARKERS] for c in self.cfgs] + [self.storage])
        return hashlib.sha1(key.encode()).hexdigest()[:16]

    def reset(self) -> None:
        """Back to the initial state (equilibrium, step counter 0)."""
        import torch
        for l in self.lbms:
            l._init_feq()
            l.steps = 0
        # the IBM generations carry on: the next one's clamp reads the last
        # one's box as its previous box and clears the forces left there
        self.steps_done = 0
        self.state = 0
        self.build(self.block, self.group, self.lds)
        torch.cuda.synchronize(self.device)

    def autotune(self, steps: int = 6, cache: str = TUNE_CACHE, use_cache: bool = True):
        """Time each (block, group, lds) candidate over ``steps`` eager steps,
        keep the fastest, restore the initial state; returns {"block/group/lds":
        ms per step}."""
        import torch
        sig = self.signature()
        table = {}
        if use_cache and cache and os.path.exists(cache):
            try:
                with open(cache) as fh:
                    table = json.load(fh)
            except (OSError, ValueError):
                table = {}
        if sig in table and len(table[sig]["best"]) == 3:
            self.build(*table[sig]["best"])
            return table[sig]["times"]
        times = {}
        for block, group, lds in CANDIDATES:
            self.build(block, group, lds)
            self.step(1)                      # warm-up (first-use costs)
            torch.cuda.synchronize(self.device)
            t0 = time.perf_counter()
            self.step(steps)
            torch.cuda.synchronize(self.device)
            times[f"{block}/{group}/{lds}"] = (time.perf_counter() - t0) / steps * 1e3
            self.reset()
        best = [int(v) for v in min(times, key=times.get).split("/")]
        self.build(*best)
        table[sig] = {"best": best, "times": times}
        if cache:
            try:
                os.makedirs(os.path.dirname(cache), exist_ok=True)
                with open(cache, "w") as fh:
                    json.dump(table, fh)
            except OSError:
                pass
        return times

    @property
    def cells(self) -> int:
        return sum(c.Nx * c.Ny * c.Nz for c in self.cfgs)


def sequential(cfgs, steps: int, storage: str = "f32", device=None, first_step: int = 1,
               deviation: str = "weights", inlet=True, per=None):
    """Each instance on its own (its CLBMSolver and DroneMovingIBM with their
    default per-instance launches: the overlapped split step on the large
    lattices), one after the other; returns the objects and the elapsed
    seconds (``per``: a list that receives each instance's seconds)."""
    import torch
    objs, total = [], 0.0
    for c in cfgs:
        lbm = CLBMSolver(c, storage=storage, device=device, deviation=deviation, inlet=inlet)
        ibm = DroneMovingIBM(c, device=device)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for n in range(first_step, first_step + steps):
            fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *c.motion_state(n))
            lbm.step(fx, fy, fz)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        total += dt
        if per is not None:
            per.append(dt)
        objs.append((lbm, ibm))
    return objs, total


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16", "f16c", "f64"])
    ap.add_argument("--deviation", default="weights", choices=["weights", "inflow"],
                    help="deviation base of 16-bit storage")
    ap.add_argument("--cache-mib", type=int, default=192,
                    help="population bytes per cache group in MiB (0 = one group)")
    ap.add_argument("--sequential-steps", type=int, default=200,
                    help="steps of the per-instance sequential timing (0 = skip)")
    ap.add_argument("--lanes", type=int, default=1,
                    help="groups advanced at once, each on its own stream")
    ap.add_argument("--solo-min-mcells", type=float, default=None,
                    help="lattices of this many million cells or more step on their own objects")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    import torch
    cfgs = draw_instances(a.instances, a.seed)
    kw = {} if a.solo_min_mcells is None else {"solo_min_cells": int(a.solo_min_mcells * 1e6)}
    sweep = BatchSweep(cfgs, storage=a.storage, steps=a.steps + 64, cache_bytes=a.cache_mib << 20,
                       deviation=a.deviation, lanes=a.lanes, **kw)
    tune = None if a.no_tune else sweep.autotune()
    print(f"[batch] {len(cfgs)} instances, {sweep.cells / 1e6:.1f} M cells, block {sweep.block}, "
          f"group {sweep.group}, lds {sweep.lds}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = 0
    group_ms = {}                      # group (instance indices) -> milliseconds in the run
    while done < a.steps:
        k = min(500, a.steps - done)
        times = []
        sweep.run(k, times=times)
        done += k
        torch.cuda.synchronize()
        for idx, e0, e1 in times:
            group_ms[tuple(idx)] = group_ms.get(tuple(idx), 0.0) + e0.elapsed_time(e1)
        print(f"[batch] {done} steps", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    # the group times cover all steps but the first (an eager step of the
    # whole batch when the sweep starts); per grid size: cell updates / time
    by_grid = {}
    steps_timed = a.steps - 1
    for idx, ms in group_ms.items():
        for i in idx:
            c = cfgs[i]
            key = f"{c.Nx}x{c.Ny}x{c.Nz}"
            share = c.Nx * c.Ny * c.Nz / sum(cfgs[j].Nx * cfgs[j].Ny * cfgs[j].Nz for j in idx)
            cells, sec = by_grid.get(key, (0, 0.0))
            by_grid[key] = (cells + c.Nx * c.Ny * c.Nz, sec + share * ms / 1e3)
    res = {"instances": len(cfgs), "cells": sweep.cells, "steps": a.steps, "storage": a.storage,
           "deviation": a.deviation,
           "block": sweep.block, "group": sweep.group, "lds": sweep.lds,
           "cache_groups": len(sweep.groups),
           "solo_groups": sum(g.solo for g in sweep.groups),
           "cache_mib": a.cache_mib, "lanes": sweep.lanes, "tuning_ms_per_step": tune,
           "seconds": dt, "mlups": sweep.cells * a.steps / dt / 1e6,
           "finite": all(l.health()[1] for l in sweep.lbms),
           "by_grid": {k: {"instances": sum(f"{c.Nx}x{c.Ny}x{c.Nz}" == k for c in cfgs),
                           "mlups": cells * steps_timed / sec / 1e6}
                       for k, (cells, sec) in by_grid.items()}}
    if a.sequential_steps:
        del sweep
        torch.cuda.empty_cache()
        per = []
        _, ts = sequential(cfgs, a.sequential_steps, a.storage, deviation=a.deviation, per=per)
        res["sequential_steps"] = a.sequential_steps
        res["sequential_mlups"] = sum(c.Nx * c.Ny * c.Nz for c in cfgs) * a.sequential_steps / ts / 1e6
        res["speedup_vs_sequential"] = res["mlups"] / res["sequential_mlups"]
        seq = {}
        for c, t in zip(cfgs, per):
            key = f"{c.Nx}x{c.Ny}x{c.Nz}"
            cells, sec = seq.get(key, (0, 0.0))
            seq[key] = (cells + c.Nx * c.Ny * c.Nz, sec + t)
        for k, (cells, sec) in seq.items():
            row = res["by_grid"].setdefault(k, {})
            row["sequential_mlups"] = cells * a.sequential_steps / sec / 1e6
            if "mlups" in row:
                row["speedup"] = row["mlups"] / row["sequential_mlups"]
    text = json.dumps(res, indent=1)
    if a.json:
        with open(a.json, "w") as fh:
            fh.write(text)
    print(text, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
