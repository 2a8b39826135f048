"""Benchmark: one CLBM-IBM time step of the 600 x 200 x 50 moving NACA 0012 case.

A step is what main_moving.py:116-122 does per iteration: host kinematics,
``DroneMovingIBM.compute_force`` (marker update, interpolate + force + spread,
clamp into fields) and ``CLBMSolver.step`` (fused collide + stream, inlet and
outlet planes).  Populations, markers and forces are resident in HBM before
the timed region starts.  Prints one JSON line.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--storage f32|f16|bf16|f16c|f64]
"""

from __future__ import annotations

import argparse
import ctypes as C
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
PKG = "clbm_ibm-moving-airfoil_amd"
METRIC = "MLUPS + % HBM roofline, 600×200×50 CLBM-IBM; FP16-store vs FP32 @1e-3 CL err"
HBM_PEAK_GBS = 8000.0
ELEM = {"f32": 4, "f16": 2, "bf16": 2, "f64": 8, "f16c": 2}


def parse(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--storage", default="f32", choices=sorted(ELEM))
    ap.add_argument("--deviation", default="inflow", choices=["weights", "inflow"],
                    help="deviation base of a 16-bit --storage")
    ap.add_argument("--case", default="c1", choices=["c0", "c1", "c2"])
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer 10 times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes)")
    ap.add_argument("--events-every", type=int, default=None,
                    help="time the step kernel with events on every N-th step (0: none)")
    ap.add_argument("--spread-sort-probe", action="store_true",
                    help="only time a sort-based spread of the markers' stencil contributions "
                         "beside the library's marker computation")
    ap.add_argument("--stream-probe", action="store_true",
                    help="only time clbm_stream_probe (the step's data movement without the "
                         "collision) on the --case grid with --storage populations")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target duration of the CPU baseline sample (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=8)
    ap.add_argument("--device-warmup", type=float, default=0.25, metavar="SECONDS",
                    help="before each storage's warm-up steps, keep the GPU busy with device "
                         "copies for this long (its clocks leave their idle state; nothing of "
                         "it is timed or counted as a step; 0 = off)")
    ap.add_argument("--group", type=int, default=None,
                    help="markers per workgroup of the marker kernel (4, 8, 16)")
    ap.add_argument("--outside-lds", type=int, default=None,
                    help="LDS bytes reserved per workgroup by the outside part of the step")
    ap.add_argument("--no-overlap", dest="overlap", action="store_false",
                    help="keep the marker kernels on the step's stream")
    ap.add_argument("--no-16bit", dest="with_16bit", action="store_false",
                    help="skip the FP16/BF16-storage runs of the same workload")
    ap.add_argument("--two-launch", dest="fused", action="store_const", const=False,
                    default=True, help="split steps as two launches around a join of the "
                    "marker stream (clbm_step_split) instead of one (clbm_step_fused)")
    ap.add_argument("--fused-joined", dest="fused", action="store_const", const="joined",
                    help="the one-launch split step after a join of the marker stream each "
                    "step (for rocprofv3 --pmc, which runs kernels one at a time)")
    ap.add_argument("--profile-only", action="store_true",
                    help="run warmup + timed steps only (for rocprofv3)")
    ap.add_argument("--coupled", default=None, metavar="IL,SPREAD,SPINS",
                    help="step through coupled.step_coupled (markers and lattice in one launch "
                         "on one stream) with these interleave, inside-spread and wait-spins "
                         "values instead of the two-stream split step")
    return ap.parse_args(argv)


def cpu_baseline(cfg, seconds: float, threads: int):
    """The reference-structured C restatement (oracle/, AoS double buffer,
    12 kernels per step, float32) timed on the host on a bounded sample of the
    same workload: the same grid and markers for as many steps as fit in
    ``seconds``.  It runs the build with FMA contraction (libclbm_ref_fast.so,
    flags in the record) when present, else the bit-exact test build."""
    if seconds <= 0:
        return None
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import clbm_ref_c
    except Exception as e:  # noqa: BLE001
        return {"value": None, "unit": "MLUPS", "cores": threads, "kind": "port",
                "sample": f"unavailable: {e}"}
    fast = clbm_ref_c.available(fast=True)
    if not fast and not clbm_ref_c.available():
        return {"value": None, "unit": "MLUPS", "cores": threads, "kind": "port",
                "sample": "oracle/libclbm_ref.so not built"}
    flags = (clbm_ref_c.FAST_FLAGS if fast else
             "-O3 -march=x86-64-v3 -ffp-contract=off -fno-fast-math -fopenmp -std=c11")
    sim = clbm_ref_c.RefSim(cfg, threads=threads, fast=fast)
    t0 = time.perf_counter()
    sim.run(1, first_step=1)
    t1 = time.perf_counter() - t0
    n = int(max(2, min(500, seconds / max(t1, 1e-6))))
    t0 = time.perf_counter()
    sim.run(n, first_step=2)
    dt = time.perf_counter() - t0
    ncell = cfg.Nx * cfg.Ny * cfg.Nz
    return {"value": round(n * ncell / dt / 1e6, 2), "unit": "MLUPS", "cores": threads,
            "kind": "port",
            "sample": (f"{n} steps of the {cfg.Nx}x{cfg.Ny}x{cfg.Nz} case with its "
                       f"{cfg.N_MARKERS} markers ({dt:.1f} s), reference-structured C/OpenMP "
                       f"restatement of the 12 kernels, float32 AoS, {threads} threads"),
            "flags": "gcc " + flags}


def host_cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _traffic_file():
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        return json.load(fh)


def traffic_record(case: str, storage: str):
    """HBM bytes per launch of the fused step kernel, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs summarised in profiles/ (null when absent).
    Not measured in this run: the figure holds for the kernels of the commit
    that traffic_source names."""
    return _traffic_file().get(f"{case}/{storage}")


def traffic_source() -> str:
    commit = _traffic_file().get("_commit")
    return ("profiles/traffic.json: rocprofv3 FETCH_SIZE + WRITE_SIZE passes on the kernels of "
            f"commit {commit}, calibrated by a device copy" if commit else "none")


def copy_calibration(mib: int) -> int:
    """A plain device-to-device copy of known size, profiled with the same
    counters as the step kernel to turn counter readings into bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(10):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 10}))
    return 0


def stream_probe(pkg, cfg, storage, steps, warmup, quiet=False):
    """clbm_stream_probe over the whole lattice, parities alternating as in a
    run: every cell's 2 x 19 population accesses of a step in the fused
    kernel's sweep order, without the collision.  Timed with events on the
    stream it is launched on.  Prints the record (returns 0), or returns it
    with ``quiet``."""
    import torch
    lib = importlib.import_module(PKG + "._lib")
    lbm = pkg.CLBMSolver(cfg, storage=storage, device="cuda:0")
    st = torch.cuda.current_stream()
    h = C.c_void_p(st.cuda_stream)
    fn = lib.fn("clbm_stream_probe")
    gref, fp = C.byref(lbm.grid), C.c_void_p(lbm.state.data_ptr())
    for i in range(warmup):
        lib.check(fn(gref, fp, i & 1, h), "clbm_stream_probe")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for i in range(steps):
        lib.check(fn(gref, fp, (warmup + i) & 1, h), "clbm_stream_probe")
    e1.record(st)
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / steps
    nbytes = 2 * 19 * ELEM[storage] * lbm.ncell
    rec = {"probe": "clbm_stream_probe: data movement of one AA step, no collision",
           "workload": cfg.name, "storage": storage, "launches": steps,
           "us_per_launch": round(us, 2), "bytes_per_launch": nbytes,
           "achieved_GBs": round(nbytes / us / 1e3, 1),
           "frac_of_8TBs": round(nbytes / us / 1e3 / HBM_PEAK_GBS, 4)}
    del lbm
    torch.cuda.empty_cache()
    if quiet:
        return rec
    print(json.dumps(rec))
    return 0


def spread_sort_probe(pkg, cfg, reps: int = 200) -> int:
    """The cost of spreading by sorting, measured beside the marker kernel.

    A design that spreads by sorting (cell, contribution) pairs by cell and
    summing each cell's run in order (deterministic without atomics) sorts,
    per step, the 64 stencil cells of every marker.  With the markers of a few
    coupled steps of the case, this times on the GPU (events, ``reps``
    repetitions): the library's whole marker computation (``compute_force`` on
    one stream: the fused marker kernel placing the markers, interpolating,
    forming the forces and spreading them with fixed-point atomics, and the
    clamp kernel); a stable radix sort of the keys alone (torch.sort, rocPRIM);
    and the sort with the rest of a sort-based spread (gather of two int64
    contributions per key, prefix sums, the sum of each cell's run, one store
    per cell).  Then the library's own sort-by-cell spread
    (``spread_mode="sorted"``): its whole compute_force, and its spread pass
    alone against the separate fixed-point spread kernel (one atomic per
    stencil point, no LDS pre-sum) on the same markers.  Prints one JSON
    record."""
    import torch
    lbm = pkg.CLBMSolver(cfg, device="cuda:0")
    ibm = pkg.DroneMovingIBM(cfg, device="cuda:0", overlap=False)
    for n in range(1, 5):
        fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(n))
        lbm.step(fx, fy, fz)
    st = torch.cuda.current_stream()

    def timed(fn):
        for _ in range(5):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            fn()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) * 1e3 / reps

    kin = cfg.motion_state(5)
    t_ibm = timed(lambda: ibm.compute_force(lbm.ux, lbm.uy, *kin))
    # the 64 stencil cells of every marker (x clamped, y and z periodic)
    X, Y, Z = (t._t.float() for t in (ibm.X_marker, ibm.Y_marker, ibm.Z_marker))
    d = torch.arange(-1, 3, device="cuda:0")
    ix = (torch.floor(X)[:, None] + d).clamp(0, cfg.Nx - 1).long()
    iy = torch.remainder(torch.floor(Y)[:, None] + d, cfg.Ny).long()
    iz = torch.remainder(torch.floor(Z)[:, None] + d, cfg.Nz).long()
    cells = ((iz[:, :, None, None] * cfg.Ny + iy[:, None, :, None]) * cfg.Nx
             + ix[:, None, None, :]).reshape(-1).to(torch.int32)
    vals = torch.randint(-(1 << 40), 1 << 40, (2, cells.numel()), device="cuda:0",
                         dtype=torch.int64)
    acc = torch.zeros(2, lbm.ncell, dtype=torch.int64, device="cuda:0")

    def sort_only():
        torch.sort(cells, stable=True)

    def sort_spread():
        k, perm = torch.sort(cells, stable=True)
        v = vals[:, perm].cumsum(1)
        last = torch.ones_like(k, dtype=torch.bool)
        last[:-1] = k[1:] != k[:-1]
        ends = v[:, last]
        sums = torch.cat([ends[:, :1], ends[:, 1:] - ends[:, :-1]], 1)
        acc[:, k[last].long()] = sums

    t_sort = timed(sort_only)
    t_spread = timed(sort_spread)
    # the library's own sort-by-cell spread (spread_mode="sorted": markers,
    # interpolation, the grouped spread and the clamp as separate launches),
    # and the two spreads alone on the same markers and forces
    lib = importlib.import_module(PKG + "._lib")
    ibs = pkg.DroneMovingIBM(cfg, device="cuda:0", overlap=False, spread_mode="sorted")
    t_sorted_cf = timed(lambda: ibs.compute_force(lbm.ux, lbm.uy, *kin))
    P = C.c_void_p
    g32 = lib.ClbmGrid(cfg.Nx, cfg.Ny, cfg.Nz, lib.STORE_F32)
    marks = [P(t._t.data_ptr()) for t in (ibs.X_marker, ibs.Y_marker, ibs.Z_marker, ibs.FX_k,
                                          ibs.FY_k)]
    scratch = torch.zeros(2, lbm.ncell, dtype=torch.int64, device="cuda:0")
    acc = [P(scratch[0].data_ptr()), P(scratch[1].data_ptr())]
    work = ibs._sort_work
    t_sorted_spread = timed(lambda: lib.call(
        "ibm_spread_sorted", C.byref(g32), C.byref(ibs.ip), ibs.N, *marks, *acc,
        P(work.data_ptr()), work.numel(), P(st.cuda_stream)))
    ipf = lib.IbmParams(float(ibs.ip.f_max), int(ibs.ip.wrap_weight_quirk), lib.SPREAD_FIXED, 0)
    t_fixed_spread = timed(lambda: lib.call(
        "ibm_spread", C.byref(g32), C.byref(ipf), ibs.N, *marks, *acc, P(st.cuda_stream)))
    rec = {"probe": "spreading by sort vs the fused marker kernel", "workload": cfg.name,
           "markers": ibm.N, "keys": int(cells.numel()), "reps": reps,
           "us_compute_force": round(t_ibm, 2), "us_sort_keys": round(t_sort, 2),
           "us_sort_based_spread": round(t_spread, 2),
           "us_compute_force_sorted": round(t_sorted_cf, 2),
           "us_spread_sorted_alone": round(t_sorted_spread, 2),
           "us_spread_fixed_alone": round(t_fixed_spread, 2)}
    print(json.dumps(rec))
    return 0


def device_warmup(seconds: float) -> None:
    """Keep the GPU busy with 256 MiB device-to-device copies for ``seconds``
    of wall time.  After the idle seconds of a process start or of building
    the next storage's objects the GPU sits in a low clock state, which a run
    of a few milliseconds (the 20 steps of a short bench) does not leave: on
    MI355X the fused step kernel ran 2-3 % slower for hundreds of steps after
    an idle gap than after 0.2 s of load.  Untimed, not a step."""
    import torch
    if seconds <= 0:
        return
    a = torch.empty(64 << 20, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(8):
            b.copy_(a)
        torch.cuda.synchronize()
    del a, b


# The step kernel's events are recorded on every 64th timed step: on MI355X a
# pair of timing events recorded around a launch costs the step loop about
# 1.5 % (f32) to 3.5 % (16-bit) of its rate when done on every 4th step, and
# 0.4-0.8 % on every 16th.
EVENTS_EVERY = 64

# the dominant kernel of a step, by how the step ran (CLBMSolver.last_step_kind)
KERNEL_NAMES = {
    "fused": "k_step_fused: the lattice outside the body region (fused collide+stream), beside "
             "the marker kernels, then the region with the force once the marker stream has "
             "published it, in one launch",
    "split": "k_step outside the body region (fused collide+stream), concurrent with the "
             "marker kernels",
    "full": "k_step (fused collide+force+stream)",
    "coupled": "k_coupled_step (markers, lattice and region in one launch) and its tail",
}
KERNEL_STEPS = 200


def measure(pkg, cfg, storage, steps, warmup, dev, tune, deviation="weights", coupled=None,
            device_warmup_s=0.0, fused=True):
    """Warm-up, then ``steps`` timed coupled steps (compute_force + step) of a
    fresh solver/IBM pair with ``storage`` populations.  ``coupled``: a
    (interleave, inside_spread, wait_spins) triple to step through
    coupled.step_coupled instead (its two kernels timed as the step kernel)."""
    import torch
    lbm = pkg.CLBMSolver(cfg, storage=storage, device=dev, deviation=deviation, **tune["lbm"])
    lbm.fused_split = fused
    ibm = pkg.DroneMovingIBM(cfg, device=dev, **tune["ibm"])
    device_warmup(device_warmup_s)
    if coupled is not None:
        return measure_coupled(pkg, cfg, lbm, ibm, steps, warmup, coupled)
    step = 1
    for _ in range(warmup):
        fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(step))
        lbm.step(fx, fy, fz)
        step += 1
    # The dominant kernel is the fused step kernel over the lattice away from
    # the body (the whole lattice when the step is not split).  CLBMSolver
    # records a pair of HIP events around it on the stream it launches on, on
    # every EVERY-th timed step, alternating parities; a run of fewer than
    # 2 EVERY steps times its two middle steps instead (one of each parity, away
    # from the start-up of the timed region).  With EVERY = 0 the timed steps
    # record none, and the kernel is timed over KERNEL_STEPS further steps right
    # after them (events on every other step).
    EVERY = EVENTS_EVERY
    middle = (steps // 2 - 1, steps // 2) if 0 < EVERY and 2 <= steps < 2 * EVERY else None
    n_ev = steps // max(EVERY, 1) + 2 if EVERY > 0 else KERNEL_STEPS // 2 + 1
    pool = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            for _ in range(n_ev)]
    for e0, e1 in pool:        # created now (a first record), so the step records them itself
        e0.record()
        e1.record()
    ev, cells = [], []

    def run_steps(n, every):
        nonlocal step
        for i in range(n):
            xc, yc, th, vx, vy, om = cfg.motion_state(step)
            fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, xc, yc, th, vx, vy, om)
            if middle is not None and every == EVERY:
                timed = i in middle
            else:
                timed = every > 0 and i % every == (i // every) % 2
            if timed:
                ev.append(pool[len(ev)])
                lbm.kernel_events = ev[-1]
            lbm.step(fx, fy, fz)
            if timed:
                cells.append(lbm.last_kernel_cells)
            step += 1

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_steps(steps, EVERY)
    t_host = time.perf_counter() - t0      # the host has queued every step
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    if EVERY <= 0:
        run_steps(KERNEL_STEPS, 2)
        torch.cuda.synchronize()
    rho_min, finite = lbm.health()
    k_ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in ev]))
    out = dict(dt=dt, k_ms=k_ms, k_cells=float(np.mean(cells)), rho_min=rho_min, finite=finite,
               host_ms=t_host / steps * 1e3, kind=lbm.last_step_kind)
    del lbm, ibm
    torch.cuda.empty_cache()
    return out


def measure_coupled(pkg, cfg, lbm, ibm, steps, warmup, knobs):
    """measure() through coupled.step_coupled; HIP events on the stream
    bracket its two launches on every EVENTS_EVERY-th timed step (as
    measure() times the split step's outside part)."""
    import torch
    cm = importlib.import_module(PKG + ".coupled")
    il, spread, spins = knobs
    step = 1
    for _ in range(warmup):
        cm.step_coupled(lbm, ibm, *cfg.motion_state(step), interleave=il, inside_spread=spread,
                        wait_spins=spins)
        step += 1
    st = torch.cuda.current_stream()
    ev = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        timed = i % max(EVENTS_EVERY, 1) == 0
        if timed:
            e = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            e[0].record(st)
        cm.step_coupled(lbm, ibm, *cfg.motion_state(step), interleave=il, inside_spread=spread,
                        wait_spins=spins)
        if timed:
            e[1].record(st)
            ev.append(e)
        step += 1
    t_host = time.perf_counter() - t0
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    rho_min, finite = lbm.health()
    k_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    out = dict(dt=dt, k_ms=k_ms, k_cells=float(lbm.ncell), rho_min=rho_min, finite=finite,
               host_ms=t_host / steps * 1e3, kind="coupled")
    del lbm, ibm
    torch.cuda.empty_cache()
    return out


def main(argv=None) -> int:
    a = parse(argv)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the multi-GPU x-slab "
              f"decomposition is a design, not a measured path.", file=sys.stderr)
        return 2
    if a.copy_calibration:
        return copy_calibration(a.copy_calibration)
    global EVENTS_EVERY
    if a.events_every is not None:
        EVENTS_EVERY = a.events_every
    import torch
    pkg = importlib.import_module(PKG)
    cfgmod = pkg.config
    cfg = {"c0": cfgmod.case_c0, "c1": cfgmod.case_c1, "c2": cfgmod.case_c2_large}[a.case]()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if a.stream_probe:
        return stream_probe(pkg, cfg, a.storage, a.steps, a.warmup)
    if a.spread_sort_probe:
        return spread_sort_probe(pkg, cfg)
    ncell = cfg.Nx * cfg.Ny * cfg.Nz
    tune = {"lbm": {}, "ibm": {"overlap": a.overlap}}
    if a.outside_lds is not None:
        tune["lbm"]["outside_lds"] = a.outside_lds
    if a.group is not None:
        tune["ibm"]["group"] = a.group
    knobs = tuple(int(v) for v in a.coupled.split(",")) if a.coupled else None
    r = measure(pkg, cfg, a.storage, a.steps, a.warmup, dev, tune, deviation=a.deviation,
                coupled=knobs, device_warmup_s=a.device_warmup, fused=a.fused)
    if knobs is not None:
        print(json.dumps({"coupled": list(knobs), "workload": cfg.name, "storage": a.storage,
                          "steps": a.steps,
                          "mlups": round(a.steps * ncell / r["dt"] / 1e6, 1),
                          "ms_per_step": round(r["dt"] / a.steps * 1e3, 5),
                          "kernels_ms": round(r["k_ms"], 5), "finite": r["finite"]}))
        return 0
    if a.profile_only:
        print(json.dumps({"steps": a.steps, "seconds": r["dt"], "finite": r["finite"]}))
        return 0
    dt, k_ms, k_cells = r["dt"], r["k_ms"], r["k_cells"]
    rho_min, finite = r["rho_min"], r["finite"]
    b_alg = 2 * 19 * ELEM[a.storage] * k_cells
    achieved = b_alg / (k_ms * 1e-3) / 1e9
    # the same workload with 16-bit storage of the deviation from the inflow
    # equilibrium (the metric's FP16 line; f16c is the 1-4-11 format)
    others = {}
    if a.storage == "f32" and a.with_16bit:
        for st in ("f16c", "f16", "bf16"):
            o = measure(pkg, cfg, st, a.steps, a.warmup, dev, tune, deviation="inflow",
                        device_warmup_s=a.device_warmup, fused=a.fused)
            m = a.steps * ncell / o["dt"] / 1e6
            others[st] = {"value": round(m, 1), "deviation": "inflow",
                          "ms_per_step": round(o["dt"] / a.steps * 1e3, 5),
                          "kernel_ms": round(o["k_ms"], 5),
                          "host_ms_per_step": round(o["host_ms"], 5),
                          "kernel_frac": round(2 * 19 * 2 * o["k_cells"] / (o["k_ms"] * 1e-3)
                                               / 1e9 / HBM_PEAK_GBS, 4),
                          "traffic": traffic_record(a.case, st),
                          "finite": o["finite"]}

    mlups = a.steps * ncell / dt / 1e6
    # the dominant kernel against its data-movement ceiling: the same loads and
    # stores in the same order without the collision (whole lattice)
    probe = stream_probe(pkg, cfg, a.storage, 200, 10, quiet=True)
    ceiling = {"kernel": "k_stream_probe", "achieved": probe["achieved_GBs"], "unit": "GB/s",
               "us_per_launch": probe["us_per_launch"],
               "kernel_frac_of_probe": round(achieved / probe["achieved_GBs"], 4)}
    cpu = cpu_baseline(cfg, a.cpu_seconds, a.cpu_threads)
    if cpu and cpu.get("value"):
        # the same restatement on every core this process may use (capped by
        # OMP_NUM_THREADS where set), on half the sample time
        host = len(os.sched_getaffinity(0))
        host = min(host, int(os.environ.get("OMP_NUM_THREADS", host)))
        cpu["cpu_model"] = host_cpu_model()
        if host != a.cpu_threads:
            wide = cpu_baseline(cfg, a.cpu_seconds / 2, host)
            if wide and wide.get("value"):
                cpu["all_cores"] = {"value": wide["value"], "unit": "MLUPS", "cores": host,
                                    "sample": wide["sample"]}
    out = {
        "metric": METRIC, "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(dt / a.steps * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"f32": "f32", "f16": "f16-store/f32", "bf16": "bf16-store/f32",
                  "f16c": "f16c-store/f32", "f64": "f64"}[a.storage],
        "data": "synthetic (feq initial state, prescribed kinematics; no dataset)",
        "config": {"workload": f"{cfg.name}: D3Q19 cascaded LBM + Peskin IBM step, "
                               f"{cfg.N_MARKERS} markers, Re {cfg.Re:g}, tau {cfg.tau:.3f}",
                   "grid": [cfg.Nx, cfg.Ny, cfg.Nz], "markers": cfg.N_MARKERS,
                   "storage": a.storage, "spread": "fixed-point", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm",
                     "kernel": KERNEL_NAMES.get(r.get("kind"), "k_step (fused collide+force+stream)"),
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_record(a.case, a.storage),
                     "traffic_source": traffic_source(),
                     "kernel_ms": round(k_ms, 5),
                     "cells_per_launch": round(k_cells),
                     "bytes_per_launch": round(b_alg),
                     "step_frac": round(mlups * 1e6 * 2 * 19 * ELEM[a.storage] / 1e9 / HBM_PEAK_GBS, 4)},
        "host_ms_per_step": round(r["host_ms"], 5),
        "device_warmup_s": a.device_warmup,
        "ceiling_probe": ceiling,
        "cpu_baseline": cpu,
        "health": {"rho_min": rho_min, "finite": finite},
        "storage_16bit": {k: dict(v, ratio_vs_f32=round(v["value"] / mlups, 3))
                          for k, v in others.items()},
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
