"""Precision study: the same coupled run with FP64, FP32, FP16 and BF16 populations.

FP16/BF16 store each population as its deviation from the lattice weight
(f_q - w_q) and compute in FP32; FP64 stores and computes in FP64 and is the
reference.  For each storage the study records the FP64 marker force sums of
every step, from which the mean lift and drag coefficients and the RMS lift
over a window are taken, and the final velocity field.  It reports, against
FP64:

* relative error of the mean C_L and of the mean C_D over the window,
* relative error of the RMS of C_L about its mean,
* the relative L2 error of the final velocity field (ux, uy, uz),
* throughput (MLUPS) of each storage in the same run.

Usage: python -m <package>.precision --case c1 --steps 20000 [--skip 2000] [--json out.json]
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import config as cfgmod
from .airfoil_moving import DroneMovingIBM
from .clbm_moving import CLBMSolver

CASES = {"c0": cfgmod.case_c0, "c1": cfgmod.case_c1, "c2": cfgmod.case_c2_large}


def run_storage(cfg, storage: str, steps: int, device=None):
    """Coupled run of ``steps`` steps; returns (C_D, C_L) histories, the final
    (ux, uy, uz) as float64 arrays and the elapsed time.  ``storage`` may end
    in "-inflow" for 16-bit deviations from the inflow equilibrium."""
    import torch
    base, _, dev = storage.partition("-")
    lbm = CLBMSolver(cfg, storage=base, device=device, deviation=dev or "weights")
    ibm = DroneMovingIBM(cfg, device=device)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for n in range(1, steps + 1):
        fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(n))
        lbm.step(fx, fy, fz)
        ibm.record_force()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    finite = lbm.health()[1]
    cd, cl = ibm.coefficients(ibm.force_history())
    u = [getattr(lbm, k).to_numpy().astype(np.float64) for k in ("ux", "uy", "uz")]
    del lbm, ibm
    torch.cuda.empty_cache()
    return np.asarray(cd), np.asarray(cl), u, dt, finite


def window_stats(cd, cl, skip: int):
    w = slice(min(skip, max(len(cl) - 1, 0)), None)
    mcl, mcd = float(np.mean(cl[w])), float(np.mean(cd[w]))
    rms = float(np.sqrt(np.mean((cl[w] - mcl) ** 2)))
    return mcl, mcd, rms


def rel(a, b):
    return abs(a - b) / abs(b) if b != 0 else float("inf")


def study(cfg, steps: int, skip: int, storages=("f64", "f32", "f16", "bf16"), device=None):
    ncell = cfg.Nx * cfg.Ny * cfg.Nz
    runs = {}
    for s in storages:
        runs[s] = run_storage(cfg, s, steps, device)
        print(f"[precision] {s}: {steps} steps in {runs[s][3]:.1f} s", file=sys.stderr, flush=True)
    cd64, cl64, u64, _, _ = runs["f64"]
    mcl64, mcd64, rms64 = window_stats(cd64, cl64, skip)
    norm64 = np.sqrt(sum(float(np.sum(a * a)) for a in u64))
    out = {"case": cfg.name, "grid": [cfg.Nx, cfg.Ny, cfg.Nz], "markers": cfg.N_MARKERS,
           "steps": steps, "window": [min(skip, steps - 1), steps], "results": {}}
    for s, (cd, cl, u, dt, finite) in runs.items():
        mcl, mcd, rms = window_stats(cd, cl, skip)
        du = np.sqrt(sum(float(np.sum((a - b) ** 2)) for a, b in zip(u, u64)))
        out["results"][s] = {
            "mean_CL": mcl, "mean_CD": mcd, "rms_CL": rms,
            "rel_err_mean_CL": rel(mcl, mcl64), "rel_err_mean_CD": rel(mcd, mcd64),
            "rel_err_rms_CL": rel(rms, rms64), "rel_L2_velocity": du / norm64 if norm64 else 0.0,
            "max_abs_CL_diff": float(np.max(np.abs(cl - cl64))),
            "mlups": steps * ncell / dt / 1e6, "finite": bool(finite)}
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--case", default="c1", choices=sorted(CASES))
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--skip", type=int, default=2000, help="steps left out of the window")
    ap.add_argument("--json", default=None)
    ap.add_argument("--storages", default="f64,f32,f16,bf16,f16-inflow,f16c,f16c-inflow",
                    help="comma list; f64 is the reference and must be first")
    a = ap.parse_args(argv)
    res = study(CASES[a.case](), a.steps, a.skip, tuple(a.storages.split(",")))
    text = json.dumps(res, indent=1)
    if a.json:
        with open(a.json, "w") as fh:
            fh.write(text)
    print(text, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
