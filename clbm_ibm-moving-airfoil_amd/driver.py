"""Driver loop: the counterpart of main_moving.py / main_highRe.py / fish/main_fish.py.

Per step (main_moving.py:116-122): motion state on the host in float64,
``ibm.compute_force(lbm.ux, lbm.uy, ...)``, ``lbm.step(fx, fy, fz)``.  Every
``vis_interval`` steps (main_moving.py:124-161): divergence guard, a frame
``frame_XXXXXX.npz`` with float32 ``ux, uy, uz, rho`` of shape (Nx, Ny, Nz),
and at the end ``meta.npz`` with the reference's keys (main_moving.py:173-200),
so the reference's renderers can read the output.  Added: the FP64 force
history ``forces.npz`` (sums, C_D, C_L per step).

Usage: python -m <package>.driver --case c0|c1|airplane|highre|fish [--steps N]
"""

from __future__ import annotations

import argparse
import math
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import config as cfgmod
from .airfoil_moving import DroneMovingIBM
from .clbm_moving import CLBMSolver

CASES = {"c0": cfgmod.case_c0, "c1": cfgmod.case_c1, "c2": cfgmod.case_c2_large,
         "airplane": cfgmod.reference_airplane, "highre": cfgmod.reference_highre,
         "fish": cfgmod.reference_fish}


def grid_meta(cfg) -> np.ndarray:
    """The 24-entry float64 record of main_moving.py:185-199."""
    e = cfg.extra
    fuse = e.get("fuse", (0.0, 0.0, 0.0))
    z0 = e.get("z0", cfg.Nz / 2.0)
    return np.array([
        cfg.Nx, cfg.Ny, cfg.Nz, cfg.U_inf, cfg.Re, cfg.x0, cfg.y0, z0,
        fuse[0], fuse[1], e.get("b_half", 0.0), e.get("z_wing_start", 0.0), e.get("z_wing_end", 0.0),
        float(int(round(z0))), float(int(round(cfg.x0))), cfg.vis_interval, fuse[2],
        e.get("x_tail_arm", 0.0), e.get("c_htail", 0.0), e.get("b_half_htail", 0.0),
        e.get("z_htail_start", 0.0), e.get("z_htail_end", 0.0), e.get("c_vtail", 0.0),
        e.get("h_vtail", 0.0)], dtype=np.float64)


FRAME_KEYS = ("rho", "ux", "uy", "uz")


def save_frame(path: str, level, **arrays) -> None:
    """An .npz as np.savez / np.savez_compressed write it (one .npy member per
    array), deflated at zlib ``level`` (None: stored).  np.load reads every
    level alike; level 1 deflates float fields several times faster than
    np.savez_compressed's default level 6, at a slightly larger file."""
    import zipfile
    mode = zipfile.ZIP_STORED if level is None else zipfile.ZIP_DEFLATED
    with zipfile.ZipFile(path, "w", compression=mode, compresslevel=level,
                         allowZip64=True) as zf:
        for name, a in arrays.items():
            with zf.open(name + ".npy", "w", force_zip64=True) as fh:
                np.lib.format.write_array(fh, np.asanyarray(a), allow_pickle=False)


class FrameWriter:
    """Frames leave the device without stopping the step loop.

    The reference reads four fields to the host and compresses them before
    its next step (main_moving.py:125-155).  Here a frame is the force-free
    macroscopic fields written by the library into one of ``depth`` device
    snapshot buffers on the solver's stream; a copy stream moves the snapshot
    into a pinned host buffer, and a worker thread turns it into the
    reference's (Nx, Ny, Nz) arrays and writes the .npz, while the GPU steps
    on.  A slot is reused only after its file is written, so a writer that
    falls behind holds the loop back instead of dropping frames.  ``depth``
    slots each hold a frame (4 float32 fields, 16 bytes a cell) on the device
    and in pinned host memory.  By default: one per two usable cores, 2 to 8,
    and no more than fit in ``pinned_budget`` bytes of page-locked host
    memory (1 GiB; at least one slot): 8 slots at 600 x 200 x 50 (96 MB a
    frame), 1 at the 800 x 260 x 220 airplane (732 MB a frame)."""

    PINNED_BUDGET = 1 << 30

    @staticmethod
    def default_depth(ncell: int, cores: int, budget: int) -> int:
        """Compression takes about 3 s of one core per 96 MB frame: one slot
        (and one writer thread) per two cores, from 2 to 8, and no more slots
        than fit in ``budget`` bytes of pinned memory (at least one)."""
        depth = min(8, max(2, cores // 2))
        return max(1, min(depth, budget // (16 * ncell)))

    def __init__(self, lbm, depth=None, compress=1, pinned_budget=None):
        import torch
        if depth is None:
            budget = self.PINNED_BUDGET if pinned_budget is None else int(pinned_budget)
            depth = self.default_depth(lbm.Nx * lbm.Ny * lbm.Nz, len(os.sched_getaffinity(0)),
                                       budget)
        self.lbm = lbm
        self.compress = compress
        shape = (4, lbm.Nz, lbm.Ny, lbm.Nx)
        self._dev = [torch.empty(shape, dtype=torch.float32, device=lbm.device)
                     for _ in range(depth)]
        self._host = [torch.empty(shape, dtype=torch.float32, pin_memory=True)
                      for _ in range(depth)]
        self._done = [None] * depth
        self._copy = torch.cuda.Stream(device=lbm.device)
        self._pool = ThreadPoolExecutor(max_workers=depth)
        self._k = 0

    def submit(self, path: str) -> None:
        import torch
        k = self._k % len(self._dev)
        self._k += 1
        if self._done[k] is not None:
            self._done[k].result()          # the slot's previous frame is on disk
        dev, host = self._dev[k], self._host[k]
        self.lbm.macro_into(dev)
        self._copy.wait_stream(torch.cuda.current_stream(self.lbm.device))
        with torch.cuda.stream(self._copy):
            host.copy_(dev, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self._copy)
        self._done[k] = self._pool.submit(self._write, ev, host, path)

    def _write(self, ev, host, path: str) -> None:
        ev.synchronize()
        a = host.numpy()
        arrays = {key: np.ascontiguousarray(a[i].transpose(2, 1, 0))
                  for i, key in enumerate(FRAME_KEYS)}
        save_frame(path, _level(self.compress), ux=arrays["ux"], uy=arrays["uy"],
                   uz=arrays["uz"], rho=arrays["rho"])

    def close(self) -> None:
        for d in self._done:
            if d is not None:
                d.result()
        self._pool.shutdown()


def _level(compress):
    """zlib level of a ``compress`` argument: False -> stored, True -> 6
    (np.savez_compressed's), an int -> that level."""
    if compress is False or compress is None:
        return None
    return 6 if compress is True else int(compress)


def run(cfg, steps=None, out_dir=None, storage="f32", write_frames=True, record_forces=True,
        verbose=True, spread_mode="fixed", deviation="weights", async_frames=True,
        compress=1, inlet=True, checkpoint=None, checkpoint_every=None, resume=None):
    """Run the coupled loop; returns a dict with the force history and timing.
    ``forces.npz`` is written whenever the forces are recorded, frames and
    ``meta.npz`` only with ``write_frames``.  ``async_frames`` writes frames
    through ``FrameWriter`` (False: read and write each frame before the next
    step, as the reference does); ``compress`` is the frames' zlib level (1 by
    default, True for np.savez_compressed's 6, False for stored members as
    np.savez writes them; np.load returns the same arrays for all).  ``inlet``
    as CLBMSolver's (True: the reference's rule, "hecht-harting").

    ``checkpoint``: a path that receives the solver state (CLBMSolver.save_state:
    populations as held, state word, step count) with the force history and
    the frame records every ``checkpoint_every`` steps (default the frame
    interval) and at the end; ``resume``: such a file to continue from (its
    step count + 1 up to ``steps``).  The resumed run's populations, forces
    and frames equal those of one uninterrupted run, bit for bit.  No
    reference counterpart: the reference cannot resume."""
    import torch
    steps = cfg.N_steps if steps is None else steps
    out_dir = cfg.DATA_DIR if out_dir is None else out_dir
    if write_frames or record_forces:
        os.makedirs(out_dir, exist_ok=True)
    lbm = CLBMSolver(cfg, storage=storage, deviation=deviation, inlet=inlet)
    ibm = DroneMovingIBM(cfg, spread_mode=spread_mode)
    snaps = dict(xc=[], yc=[], theta=[], step=[])
    frame = 0
    first = 1
    if resume is not None:
        extra = lbm.load_state(resume)
        first = lbm.steps + 1
        if "force_history" in extra:
            ibm.restore_history(extra["force_history"])
        frame = int(extra.get("frame", 0))
        for k in snaps:
            if f"snap_{k}" in extra:
                snaps[k] = list(extra[f"snap_{k}"].tolist())
    every = checkpoint_every or cfg.vis_interval

    def save_checkpoint():
        lbm.save_state(checkpoint, force_history=ibm.force_history(), frame=np.int64(frame),
                       **{f"snap_{k}": np.asarray(v, np.float64) for k, v in snaps.items()})

    writer = FrameWriter(lbm, compress=compress) if (write_frames and async_frames) else None
    diverged = False
    torch.cuda.synchronize()
    t0 = last = time.perf_counter()
    step = first - 1
    for step in range(first, steps + 1):
        xc, yc, th, vx, vy, om = cfg.motion_state(step)
        fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, xc, yc, th, vx, vy, om)
        lbm.step(fx, fy, fz)
        if record_forces:
            ibm.record_force()
        if step % cfg.vis_interval == 0 or step == steps:
            rho_min, finite = lbm.health()
            now = time.perf_counter()
            if verbose:
                rate = cfg.vis_interval / max(now - last, 1e-9)
                print(f"  step {step:6d} | rho_min={rho_min:.4f} yc={yc:.2f} "
                      f"theta={math.degrees(th):.2f} deg | {rate:.0f} steps/s", flush=True)
            last = now
            if not finite or rho_min < 0.3:
                print(f"*** Diverged at step {step}! rho_min={rho_min:.4f} ***", flush=True)
                diverged = True
                break
            if write_frames and step % cfg.vis_interval == 0:
                frame += 1
                path = os.path.join(out_dir, f"frame_{frame:06d}.npz")
                if writer is not None:
                    writer.submit(path)
                else:
                    save_frame(path, _level(compress), ux=lbm.ux.to_numpy(),
                               uy=lbm.uy.to_numpy(), uz=lbm.uz.to_numpy(),
                               rho=lbm.rho.to_numpy())
                for k, v in zip(("xc", "yc", "theta", "step"), (xc, yc, th, step)):
                    snaps[k].append(v)
        if checkpoint is not None and (step % every == 0 or step == steps):
            save_checkpoint()
    if writer is not None:
        writer.close()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    hist = ibm.force_history()
    cd, cl = ibm.coefficients(hist) if len(hist) else (np.zeros(0), np.zeros(0))
    if write_frames:
        e = cfg.extra
        if snaps["step"]:
            np.savez_compressed(
                os.path.join(out_dir, "meta.npz"),
                xc=np.array(snaps["xc"], np.float32), yc=np.array(snaps["yc"], np.float32),
                theta=np.array(snaps["theta"], np.float32), step=np.array(snaps["step"], np.int32),
                X_template=e.get("X_template", cfg.X0), Y_template=e.get("Y_template", cfg.Y0),
                X_htail_tmpl=e.get("X_htail_tmpl", np.zeros(0, np.float32)),
                Y_htail_tmpl=e.get("Y_htail_tmpl", np.zeros(0, np.float32)),
                grid_meta=grid_meta(cfg))
    if record_forces:
        np.savez_compressed(os.path.join(out_dir, "forces.npz"), sums=hist, cd=cd, cl=cl)
    ncell = cfg.Nx * cfg.Ny * cfg.Nz
    ran = step - first + 1
    res = dict(steps=step, elapsed=elapsed, mlups=ran * ncell / max(elapsed, 1e-9) / 1e6,
               history=hist, cd=cd, cl=cl, diverged=diverged)
    if verbose:
        print(f"Done. {ran} steps in {elapsed:.1f} s ({ran / max(elapsed, 1e-9):.0f} steps/s, "
              f"{res['mlups']:.0f} MLUPS)", flush=True)
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--case", default="c0", choices=sorted(CASES))
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--out", default=None)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16", "f16c", "f64"])
    ap.add_argument("--deviation", default="weights", choices=["weights", "inflow"],
                    help="deviation base of 16-bit storage")
    ap.add_argument("--no-frames", action="store_true")
    ap.add_argument("--sync-frames", action="store_true",
                    help="read and write each frame before the next step (as the reference)")
    ap.add_argument("--no-compress", action="store_true", help="frames stored, as np.savez")
    ap.add_argument("--compress-level", type=int, default=1,
                    help="zlib level of the frames (6 = np.savez_compressed's)")
    ap.add_argument("--vis-interval", type=int, default=None, help="steps between frames")
    ap.add_argument("--inlet", default="reference", choices=["reference", "hecht-harting"],
                    help="inlet rule at x = 0 (the reference's, or the opt-in Hecht-Harting one)")
    ap.add_argument("--checkpoint", default=None, help="write the run's state to this .npz file")
    ap.add_argument("--checkpoint-every", type=int, default=None,
                    help="steps between checkpoints (default: the frame interval)")
    ap.add_argument("--resume", default=None, help="continue from this checkpoint file")
    a = ap.parse_args(argv)
    cfg = CASES[a.case]()
    if a.vis_interval is not None:
        cfg = cfg.with_(vis_interval=a.vis_interval)
    res = run(cfg, a.steps, a.out, a.storage, write_frames=not a.no_frames,
              deviation=a.deviation, async_frames=not a.sync_frames,
              compress=False if a.no_compress else a.compress_level,
              inlet=True if a.inlet == "reference" else a.inlet, checkpoint=a.checkpoint,
              checkpoint_every=a.checkpoint_every, resume=a.resume)
    return 1 if res["diverged"] else 0


if __name__ == "__main__":
    sys.exit(main())
