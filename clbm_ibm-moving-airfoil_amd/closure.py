"""Learned closure: a per-cell network that corrects the higher-order rates.

No reference counterpart.  The reference's cascade relaxes third-order
central moments to 0 and fourth-order ones to rho/9 in every cell
(s_q = s_pi = 1, clbm_moving.py:188-196).  On under-resolved high-Re grids
this closure models a network 19 -> 32 -> 32 -> 4 (GELU) that maps each
cell's 19 normalised central moments to corrections y of four rates,
s = 1 + y, for (k210, k120), (k201, k021, k102, k012), (k220) and
(k202, k022).  A cell whose rates leave [0, 2] falls back to the plain
cascade.

Everything runs on the GPU through the library's kernels:

* ``clbm_step_closure``: the fused step with the network evaluated per
  wavefront on bf16 MFMA;
* ``closure_grad`` and ``closure_adam``: training, forward and backward on
  bf16 MFMA with float32 accumulation and a fixed-order gradient reduction.

The training data come from the method itself.  A 2x-refined run of the
same case (same Re, so twice the chord, and two fine steps per coarse step)
is box-filtered onto the coarse lattice.  A cell's target rates are those
that turn its filtered populations at one coarse step into the filtered
populations its neighbours hold one coarse step later
(f*_q(x) = f_q(x + e_q, t + 1)), by least squares over each rate's moments.
"""

from __future__ import annotations

import ctypes as C
import json
import math
import sys
import time

import numpy as np

from . import _lib

NP = 1828
_SHAPES = [(19, 32), (32,), (32, 32), (32,), (32, 4), (4,)]
EX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
EY = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
EZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])
PQR = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (2, 0, 0),
       (0, 2, 0), (0, 0, 2), (2, 1, 0), (1, 2, 0), (2, 0, 1), (0, 2, 1), (1, 0, 2), (0, 1, 2),
       (2, 2, 0), (2, 0, 2), (0, 2, 2)]
FEATURE_EQ = np.array([1.0, 0, 0, 0, 0, 0, 0, 1 / 3, 1 / 3, 1 / 3, 0, 0, 0, 0, 0, 0, 1 / 9, 1 / 9,
                       1 / 9])
GROUPS = [[10, 11], [12, 13, 14, 15], [16], [17, 18]]   # feature indices of each rate


def split_params(p):
    out, o = [], 0
    for s in _SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(p[o:o + n]).reshape(s))
        o += n
    return out


def join_params(parts):
    return np.concatenate([np.asarray(a, dtype=np.float64).ravel() for a in parts])


def init_params(seed: int = 0, out_scale: float = 0.0) -> np.ndarray:
    """He-style random weights; the last layer starts at ``out_scale`` (0 gives
    the plain cascade exactly: y = 0, s = 1)."""
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((19, 32)) * math.sqrt(2.0 / 19)
    W2 = rng.standard_normal((32, 32)) * math.sqrt(2.0 / 32)
    W3 = rng.standard_normal((32, 4)) * math.sqrt(2.0 / 32) * out_scale
    return join_params([W1, np.zeros(32), W2, np.zeros(32), W3, np.zeros(4)])


def fold_standardization(params, mean, std) -> np.ndarray:
    """Parameters acting on raw features x from parameters trained on
    (x - mean) / std: W1' = W1 / std, b1' = b1 - (mean / std) W1."""
    W1, b1, W2, b2, W3, b3 = split_params(np.asarray(params, dtype=np.float64))
    s = np.asarray(std, dtype=np.float64)
    m = np.asarray(mean, dtype=np.float64)
    W1f = W1 / s[:, None]
    b1f = b1 - (m / s) @ W1
    return join_params([W1f, b1f, W2, b2, W3, b3])


class ClosureNet:
    """A network on the device: float32 parameters and their packed bf16 operands."""

    def __init__(self, params=None, device=None, lo: float = 0.0, hi: float = 2.0):
        import torch
        self.device = torch.device(device if device is not None else "cuda")
        lib = _lib.load()
        if lib.closure_param_count() != NP:
            raise _lib.ClbmError("closure parameter count mismatch")
        p = init_params() if params is None else np.asarray(params, dtype=np.float32)
        self.params = torch.from_numpy(np.ascontiguousarray(p, dtype=np.float32)).to(self.device)
        self.packed = torch.empty(lib.closure_packed_bytes(), dtype=torch.uint8, device=self.device)
        self.lo, self.hi = float(lo), float(hi)
        self.repack()

    def _st(self):
        import torch
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def repack(self) -> None:
        _lib.call("closure_pack", C.c_void_p(self.params.data_ptr()),
                  C.c_void_p(self.packed.data_ptr()), self._st())

    def numpy(self) -> np.ndarray:
        return self.params.cpu().numpy().astype(np.float64)

    def infer(self, X):
        """(n, 19) features -> (n, 4) outputs y on the device."""
        import torch
        X = torch.as_tensor(np.ascontiguousarray(X, dtype=np.float32)).to(self.device) \
            if not isinstance(X, torch.Tensor) else X.contiguous().float()
        Y = torch.empty((X.shape[0], 4), dtype=torch.float32, device=self.device)
        _lib.call("closure_infer", C.c_void_p(self.packed.data_ptr()),
                  C.c_void_p(self.params.data_ptr()), int(X.shape[0]), C.c_void_p(X.data_ptr()),
                  C.c_void_p(Y.data_ptr()), self._st())
        return Y

    def save(self, path: str, **meta) -> None:
        np.savez(path, params=self.numpy(), lo=self.lo, hi=self.hi,
                 meta=json.dumps(meta))

    @classmethod
    def load(cls, path: str, device=None):
        d = np.load(path, allow_pickle=False)
        return cls(d["params"], device=device, lo=float(d["lo"]), hi=float(d["hi"]))


class Trainer:
    """Adam on the weighted mean squared rate error, all on the device.  W is
    one weight per sample (n,) or one per output (n, 4)."""

    def __init__(self, X, T, W, device=None, blocks: int = 256, lr: float = 3e-3,
                 params=None, seed: int = 0):
        import torch
        self.device = torch.device(device if device is not None else "cuda")
        X = np.asarray(X, dtype=np.float64)
        self.mean = X.mean(0)
        self.std = X.std(0)
        self.std[self.std < 1e-12] = 1.0
        Xs = (X - self.mean) / self.std
        dev = self.device
        self.X = torch.from_numpy(np.ascontiguousarray(Xs, dtype=np.float32)).to(dev)
        self.T = torch.from_numpy(np.ascontiguousarray(T, dtype=np.float32)).to(dev)
        W = np.asarray(W, dtype=np.float32)
        if W.shape not in ((X.shape[0],), (X.shape[0], 4)):
            raise ValueError("W must have shape (n,) or (n, 4)")
        self.wdim = 1 if W.ndim == 1 else 4
        self.W = torch.from_numpy(np.ascontiguousarray(W)).to(dev)
        self.wsum = float(np.sum(W))
        self.net = ClosureNet(init_params(seed) if params is None else params, device=dev)
        self.m = torch.zeros(NP, dtype=torch.float32, device=dev)
        self.v = torch.zeros(NP, dtype=torch.float32, device=dev)
        self.blocks = int(blocks)
        lib = _lib.load()
        self.partial = torch.empty(lib.closure_partial_bytes(self.blocks) // 4, dtype=torch.float32,
                                   device=dev)
        self.grad = torch.empty(NP + 1, dtype=torch.float32, device=dev)
        self.lr = lr
        self.t = 0

    def gradient(self, lo: int = 0, hi=None):
        """(grad, loss) over samples lo..hi with the current parameters."""
        hi = self.X.shape[0] if hi is None else hi
        n = hi - lo
        wsum = float(self.W[lo:hi].sum().item())
        _lib.call("closure_grad", C.c_void_p(self.net.packed.data_ptr()),
                  C.c_void_p(self.net.params.data_ptr()), n,
                  C.c_void_p(self.X[lo:hi].data_ptr()), C.c_void_p(self.T[lo:hi].data_ptr()),
                  C.c_void_p(self.W[lo:hi].data_ptr()), self.wdim, C.c_float(wsum),
                  C.c_void_p(self.partial.data_ptr()), self.blocks,
                  C.c_void_p(self.grad.data_ptr()), self.net._st())
        return self.grad

    def step(self, lo: int = 0, hi=None) -> None:
        self.gradient(lo, hi)
        self.t += 1
        _lib.call("closure_adam", C.c_void_p(self.net.params.data_ptr()),
                  C.c_void_p(self.m.data_ptr()), C.c_void_p(self.v.data_ptr()),
                  C.c_void_p(self.grad.data_ptr()), C.c_float(self.lr), C.c_float(0.9),
                  C.c_float(0.999), C.c_float(1e-8), self.t, self.net._st())
        self.net.repack()

    def loss(self) -> float:
        return float(self.gradient()[NP].item())

    def fit(self, epochs: int = 50, batch: int = 16384, verbose: bool = False):
        """Minibatch epochs over a fixed shuffled order (deterministic)."""
        import torch
        n = self.X.shape[0]
        order = torch.from_numpy(np.random.default_rng(0).permutation(n)).to(self.device)
        self.X, self.T, self.W = self.X[order].contiguous(), self.T[order].contiguous(), \
            self.W[order].contiguous()
        hist = []
        for e in range(epochs):
            for lo in range(0, n, batch):
                self.step(lo, min(n, lo + batch))
            hist.append(self.loss())
            if verbose:
                print(f"[closure] epoch {e + 1}: loss {hist[-1]:.4e}", file=sys.stderr, flush=True)
        return hist

    def export(self) -> ClosureNet:
        """The trained network on raw features (standardisation folded in)."""
        p = fold_standardization(self.net.numpy(), self.mean, self.std)
        return ClosureNet(p, device=self.device)


# ---------------------------------------------------------------------------
# training data from the method's own refined runs (host-side preprocessing)
# ---------------------------------------------------------------------------
def box_filter(f_fine: np.ndarray) -> np.ndarray:
    """(2X, 2Y, 2Z, 19) -> (X, Y, Z, 19) float64: mean over each 2x2x2 block
    (accumulated in float64 without a float64 copy of the input)."""
    a = np.asarray(f_fine)
    X, Y, Z = a.shape[0] // 2, a.shape[1] // 2, a.shape[2] // 2
    a = a[:2 * X, :2 * Y, :2 * Z].reshape(X, 2, Y, 2, Z, 2, 19)
    return a.mean(axis=(1, 3, 5), dtype=np.float64)


# raw moments m_abc, a, b, c in {0, 1, 2}: one product with this matrix
_E27 = [(a, b, c) for a in range(3) for b in range(3) for c in range(3)]
_M27 = np.array([[EX[q] ** a * EY[q] ** b * EZ[q] ** c for q in range(19)] for (a, b, c) in _E27],
                dtype=np.float64)


def _central_moments(fT, u, v, w):
    """Central moments sum_q f_q (ex-u)^p (ey-v)^q (ez-w)^r of populations fT
    (19, n) about per-cell (u, v, w): (19, n) in PQR order, from the raw
    moments by the binomial shift (on D3Q19 the raw moments with all three
    exponents >= 1 vanish; the matrix above carries them as zero rows)."""
    R = _M27 @ fT
    col = {e: i for i, e in enumerate(_E27)}
    pu, pv, pw = [None, -u, u * u], [None, -v, v * v], [None, -w, w * w]
    out = np.empty((19, fT.shape[1]))
    for i, (p, q, r) in enumerate(PQR):
        acc = np.zeros(fT.shape[1])
        for a in range(p + 1):
            for b in range(q + 1):
                for c in range(r + 1):
                    if not _M27[col[(a, b, c)]].any():
                        continue
                    t = float(math.comb(p, a) * math.comb(q, b) * math.comb(r, c)) * R[col[(a, b, c)]]
                    for powers, k in ((pu, p - a), (pv, q - b), (pw, r - c)):
                        if k:
                            t = t * powers[k]
                    acc += t
        out[i] = acc
    return out


def _noneq_moments(fT):
    """Populations (19, n) -> (features X, non-equilibrium moments D,
    (rho, u, v, w)): D is the normalised central moments minus their
    equilibrium values (PQR order); X equals D with X[0] = rho - 1 (the
    network's input, no force)."""
    rho = fT.sum(0)
    u, v, w = (EX @ fT) / rho, (EY @ fT) / rho, (EZ @ fT) / rho
    D = _central_moments(fT, u, v, w) / rho - FEATURE_EQ[:, None]
    X = D.copy()
    X[0] = rho - 1.0
    return X, D, (rho, u, v, w)


def features_and_targets(f_now: np.ndarray, f_next: np.ndarray, mask=None,
                         per_output: bool = False):
    """Features (n, 19), target outputs y = s - 1 (n, 4) and weights from two
    consecutive filtered states on the coarse lattice (x, y, z periodic
    neighbours; cells listed by ``mask``).  Force is taken as zero, so the
    cells should lie away from the body.

    A rate's target is the least-squares fit over its group's moments, whose
    squared error grows with den_j = sum of the group's squared pre-collision
    non-equilibrium moments times (s_j - s*_j)^2.  The weights are those den_j
    normalised to a mean total of 1: per output (n, 4) with ``per_output``,
    else their sum per cell (n,)."""
    f = np.asarray(f_now, dtype=np.float64)
    fn = np.asarray(f_next, dtype=np.float64)
    if mask is None:
        mask = np.ones(f.shape[:3], dtype=bool)
    # populations as (19, n) rows over the listed cells; the ideal
    # post-collision populations are f*_q(x) = f_next_q(x + e_q)
    fT = np.stack([f[..., q][mask] for q in range(19)])
    fsT = np.stack([np.roll(fn[..., q], shift=(-EX[q], -EY[q], -EZ[q]), axis=(0, 1, 2))[mask]
                    for q in range(19)])
    X, Dpre, (rho, u, v, w) = _noneq_moments(fT)
    Dpost = _central_moments(fsT, u, v, w) / rho - FEATURE_EQ[:, None]
    T = np.zeros((4, len(rho)))
    D = np.zeros((4, len(rho)))
    for j, idx in enumerate(GROUPS):
        num = (Dpre[idx] * Dpost[idx]).sum(0)
        D[j] = (Dpre[idx] ** 2).sum(0)
        s = 1.0 - num / np.maximum(D[j], 1e-300)
        T[j] = np.clip(s, 0.0, 2.0) - 1.0
    W = D.sum(0)
    ok = np.all(np.isfinite(X), 0) & np.all(np.isfinite(T), 0) & np.isfinite(W)
    X, T = np.ascontiguousarray(X[:, ok].T), np.ascontiguousarray(T[:, ok].T)
    W, D = W[ok], np.ascontiguousarray(D[:, ok].T)
    scale = max(W.mean(), 1e-300) if len(W) else 1.0
    return X, T, (D if per_output else W) / scale


def one_step_error(cfg, f_now, f_next, net=None, mask=None, device=None):
    """A-priori check: relative L2 error, over ``mask``, of one coarse step
    (plain, or with ``net``) from the filtered state f_now against the filtered
    state one coarse step later.  x is periodic here (no planes)."""
    import torch
    from .clbm_moving import CLBMSolver
    s = CLBMSolver(cfg, device=device, inlet=False)
    s.f.from_numpy(f_now)
    if net is not None:
        s.closure = net
    s.step()
    got = s.f.to_numpy().astype(np.float64)
    ref = np.asarray(f_next, dtype=np.float64)
    m = np.ones(ref.shape[:3], bool) if mask is None else mask
    return float(np.sqrt(((got - ref)[m] ** 2).sum() / ((ref - 0)[m] ** 2).sum()))


# ---------------------------------------------------------------------------
# evaluation diagnostics
# ---------------------------------------------------------------------------
ORDERS = {2: list(range(4, 10)), 3: list(range(10, 16)), 4: list(range(16, 19))}
GROUP_NAMES = ["k210,k120", "k201,k021,k102,k012", "k220", "k202,k022"]


def diagnostics(f, net=None, mask=None, chunk: int = 1 << 20) -> dict:
    """What the closure does to populations f (X, Y, Z, 19), over the cells
    of ``mask``:

    * ``energy``: per order (2, 3, 4) the mean over cells of the summed
      squares of the normalised non-equilibrium central moments (those the
      rates of that order relax; order 2 also holds the trace);
    * with ``net``: ``on_network``, the fraction of cells whose four rates
      s = 1 + y all lie in [net.lo, net.hi] (the others fall back to s = 1);
      per rate group the mean and max |s - 1| over all cells (fallback cells
      count as 0); and ``correction_energy``, the mean summed squares of
      what the network changes in the post-collision moments, (s - 1) times
      the pre-collision non-equilibrium moment of each group (the plain
      cascade relaxes them to equilibrium: post-collision non-equilibrium 0).

    The features are computed as the fused closure step computes them from
    a cell's loaded populations (no force: far from the body they equal it)."""
    a = np.asarray(f)
    flat = a.reshape(-1, 19)
    cells = np.flatnonzero(np.ones(a.shape[:3], bool) if mask is None else np.asarray(mask).ravel())
    e = {k: 0.0 for k in ORDERS}
    n = len(cells)
    n_ok, corr = 0, 0.0
    abs_sum, abs_max = np.zeros(4), np.zeros(4)
    for lo in range(0, n, chunk):
        fT = flat[cells[lo:lo + chunk]].T.astype(np.float64)
        X, D, _ = _noneq_moments(fT)
        for k, idx in ORDERS.items():
            e[k] += float((D[idx] ** 2).sum())
        if net is None:
            continue
        Y = net.infer(np.ascontiguousarray(X.T)).cpu().numpy().astype(np.float64)
        S = 1.0 + Y
        ok = np.all(np.isfinite(S), 1) & np.all((S >= net.lo) & (S <= net.hi), 1)
        n_ok += int(ok.sum())
        Ye = np.where(ok[:, None], Y, 0.0)
        abs_sum += np.abs(Ye).sum(0)
        abs_max = np.maximum(abs_max, np.abs(Ye).max(0) if len(Ye) else 0.0)
        for j, idx in enumerate(GROUPS):
            corr += float(((Ye[:, j] * D[idx]) ** 2).sum())
    out = {"cells": int(n), "energy": {str(k): v / max(n, 1) for k, v in e.items()}}
    if net is not None:
        out["on_network"] = n_ok / max(n, 1)
        out["mean_abs_s_minus_1"] = {g: float(v / max(n, 1)) for g, v in zip(GROUP_NAMES, abs_sum)}
        out["max_abs_s_minus_1"] = {g: float(v) for g, v in zip(GROUP_NAMES, abs_max)}
        out["correction_energy"] = corr / max(n, 1)
    return out


def body_mask(cfg, ibm_box, margin: int = 3):
    """Cells within ``margin`` of the force box (all z), the complement of the
    far-field mask inside the x planes."""
    far = _far_mask(cfg, ibm_box, margin)
    m = ~far
    m[:4] = False
    m[-5:] = False
    return m


# ---------------------------------------------------------------------------
# the study: refined runs -> data -> training -> evaluation
# ---------------------------------------------------------------------------
def case_for(seed: int, coarse, chord: float, Re: float, refine: int = 1):
    """Seeded kinematics (as the batched sweeps draw them) on the coarse grid
    ``coarse`` (refine = 1) or its 2x refinement (refine = 2, twice the chord,
    same Re, kinematics at the same physical times)."""
    from . import config as cfgmod
    rng = np.random.default_rng(7919 * seed + 17)
    kw = dict(k=float(rng.uniform(0.25, 1.0)), heave_frac=float(rng.uniform(0.1, 0.5)),
              aoa_mean_deg=float(rng.uniform(0, 12)), pitch_amp_deg=float(rng.uniform(0, 20)),
              phase_deg=float(rng.uniform(0, 180)))
    nx, ny, nz = coarse
    return cfgmod.naca_case(refine * nx, refine * ny, refine * nz, refine * chord, Re=Re,
                            le_x=refine * nx / 4.0, name=f"seed{seed}-x{refine}", **kw)


def _far_mask(cfg_coarse, ibm_box, margin: int = 3):
    nx, ny, nz = cfg_coarse.Nx, cfg_coarse.Ny, cfg_coarse.Nz
    m = np.ones((nx, ny, nz), dtype=bool)
    m[:4] = False
    m[-5:] = False
    x0, y0, _, x1, y1, _ = ibm_box
    m[max(0, x0 - margin):x1 + margin + 1, max(0, y0 - margin):y1 + margin + 1, :] = False
    return m


def fine_pairs(seed: int, coarse, chord: float, Re: float, warm: int, pairs: int, gap: int,
               storage: str = "f32", device=None, log=None, region: str = "far"):
    """Run the 2x case for ``warm`` fine steps, then take ``pairs`` filtered
    (now, one coarse step later) pairs every ``gap`` fine steps.  Returns the
    pairs and a mask of coarse cells away from the x planes and, with
    ``region`` "far", away from the body ("all": the body's cells too, whose
    targets then absorb the immersed boundary's forcing)."""
    if region not in ("far", "all"):
        raise ValueError("region must be 'far' or 'all'")
    from .airfoil_moving import DroneMovingIBM
    from .clbm_moving import CLBMSolver
    cf = case_for(seed, coarse, chord, Re, refine=2)
    cc = case_for(seed, coarse, chord, Re, refine=1)
    lbm = CLBMSolver(cf, storage=storage, device=device)
    ibm = DroneMovingIBM(cf, device=device)
    out, n = [], 0

    def advance(k):
        nonlocal n
        for _ in range(k):
            n += 1
            fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cf.motion_state(n))
            lbm.step(fx, fy, fz)

    advance(warm)
    masks = []
    for _ in range(pairs):
        advance(gap)
        a = box_filter(lbm.f.to_numpy())
        box = ibm.force_box()
        advance(2)
        b = box_filter(lbm.f.to_numpy())
        out.append((a, b))
        coarse_box = tuple(int(v) // 2 for v in box)
        masks.append(_far_mask(cc, coarse_box) if region == "far" else
                     _far_mask(cc, coarse_box) | body_mask(cc, coarse_box))
        if log:
            log(f"seed {seed}: pair at fine step {n}, finite {lbm.health()[1]}")
    return out, masks


def run_forces(cfg, steps: int, net=None, storage: str = "f32", device=None, skip: int = 0):
    """Coupled run; returns (mean C_D, mean C_L over steps after ``skip``,
    seconds, finite)."""
    import torch
    from .airfoil_moving import DroneMovingIBM
    from .clbm_moving import CLBMSolver
    lbm = CLBMSolver(cfg, storage=storage, device=device)
    ibm = DroneMovingIBM(cfg, device=device)
    lbm.closure = net
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for n in range(1, steps + 1):
        fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(n))
        lbm.step(fx, fy, fz)
        ibm.record_force()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    cd, cl = ibm.coefficients(ibm.force_history())
    w = slice(skip, None)
    return float(np.mean(cd[w])), float(np.mean(cl[w])), dt, bool(lbm.health()[1])


def study(coarse=(300, 100, 24), chord=None, Re=1000.0, train_seeds=range(8), test_seeds=range(100, 104),
          warm=4000, pairs=3, gap=400, samples=600000, epochs=40, eval_steps=4000, device=None,
          log=None, net=None, apriori=True, aposteriori=True, train_region: str = "far"):
    """The full pipeline at a given scale; returns (dict of results, network).
    With ``net`` given the training stage is skipped; ``apriori`` and
    ``aposteriori`` switch the two evaluation stages; ``train_region``: the
    cells the training data come from (fine_pairs' ``region``)."""
    chord = coarse[1] / 2.5 if chord is None else chord
    log = log or (lambda s: print(f"[closure] {s}", file=sys.stderr, flush=True))
    res = {"coarse": list(coarse), "fine": [2 * v for v in coarse], "chord": chord, "Re": Re}
    if net is None:
        res["train_region"] = train_region
        net = _train_stage(res, coarse, chord, Re, train_seeds, warm, pairs, gap, samples, epochs,
                           device, log, train_region)
    if apriori:
        _apriori_stage(res, net, coarse, chord, Re, test_seeds, warm, gap, device, log)
    if aposteriori:
        _aposteriori_stage(res, net, coarse, chord, Re, test_seeds, eval_steps, device, log)
    return res, net


def _train_stage(res, coarse, chord, Re, train_seeds, warm, pairs, gap, samples, epochs, device,
                 log, region="far"):
    Xs, Ts, Ws = [], [], []
    rng = np.random.default_rng(0)
    for s in train_seeds:
        pr, masks = fine_pairs(s, coarse, chord, Re, warm, pairs, gap, device=device, log=log,
                               region=region)
        for (a, b), m in zip(pr, masks):
            X, T, W = features_and_targets(a, b, m, per_output=True)
            Xs.append(X)
            Ts.append(T)
            Ws.append(W)
    X, T, W = np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ws)
    if len(X) > samples:
        # keep cells in proportion to their non-equilibrium signal; each
        # output then keeps its share of its cell's weight
        tot = W.sum(1)
        idx = rng.choice(len(X), samples, replace=False, p=tot / tot.sum())
        X, T, W = X[idx], T[idx], W[idx] / tot[idx, None]
    res["train_samples"] = int(len(X))
    res["weights"] = "per output"
    log(f"training on {len(X)} samples")
    tr = Trainer(X, T, W, device=device, lr=3e-3)
    hist = tr.fit(epochs=epochs, batch=16384, verbose=False)
    res["loss_first_last"] = [float(hist[0]), float(hist[-1])]
    res["loss_plain"] = float((W * T ** 2).sum() / W.sum())
    return tr.export()


def _apriori_stage(res, net, coarse, chord, Re, test_seeds, warm, gap, device, log):
    # a-priori on held-out kinematics: the rate loss of the targets and one
    # coarse step from filtered states
    errs = {"plain": [], "closure": []}
    rate_loss = {"plain": [], "closure": []}
    for s in test_seeds:
        pr, masks = fine_pairs(s, coarse, chord, Re, warm, 1, gap, device=device, log=log)
        cfg = case_for(s, coarse, chord, Re)
        for (a, b), m in zip(pr, masks):
            errs["plain"].append(one_step_error(cfg, a, b, None, m, device))
            errs["closure"].append(one_step_error(cfg, a, b, net, m, device))
            Xh, Th, Wh = features_and_targets(a, b, m, per_output=True)
            Yh = net.infer(Xh).cpu().numpy().astype(np.float64)
            rate_loss["plain"].append(float((Wh * Th ** 2).sum() / Wh.sum()))
            rate_loss["closure"].append(float((Wh * (Yh - Th) ** 2).sum() / Wh.sum()))
    res["one_step_rel_error"] = {k: float(np.mean(v)) for k, v in errs.items()}
    res["heldout_rate_loss"] = {k: float(np.mean(v)) for k, v in rate_loss.items()}
    log(f"one-step error: {res['one_step_rel_error']}")


def _aposteriori_stage(res, net, coarse, chord, Re, test_seeds, eval_steps, device, log):
    # a-posteriori: coarse with / without the network against the fine run
    post = []
    for s in test_seeds:
        cc = case_for(s, coarse, chord, Re)
        cf = case_for(s, coarse, chord, Re, refine=2)
        fd, fl, ft, fok = run_forces(cf, 2 * eval_steps, device=device, skip=eval_steps)
        pd, pl_, pt, pok = run_forces(cc, eval_steps, device=device, skip=eval_steps // 2)
        nd, nl, nt, nok = run_forces(cc, eval_steps, net=net, device=device, skip=eval_steps // 2)
        post.append(dict(seed=s, fine=[fd, fl, ft, fok], plain=[pd, pl_, pt, pok],
                         closure=[nd, nl, nt, nok]))
        log(f"seed {s}: fine CD {fd:.4f} CL {fl:.4f} ({ft:.1f} s) | plain {pd:.4f} {pl_:.4f} "
            f"| closure {nd:.4f} {nl:.4f} ({nt:.1f} s)")
    res["a_posteriori"] = post

    def relerr(key, i):
        return float(np.mean([abs(r[key][i] - r["fine"][i]) / max(abs(r["fine"][i]), 1e-12)
                              for r in post]))
    res["mean_rel_error_vs_fine"] = {k: {"CD": relerr(k, 0), "CL": relerr(k, 1)}
                                    for k in ("plain", "closure")}
    res["wall_time_ratio_fine_over_closure"] = float(
        np.mean([r["fine"][2] / r["closure"][2] for r in post]))


def diagnose(net, coarse=(600, 200, 50), chord=None, Re=1000.0, seed: int = 100, warm: int = 4000,
             gap: int = 400, steps: int = 3000, device=None, log=None) -> dict:
    """Evaluation diagnostics (``diagnostics``) of a held-out case: on the
    filtered fine state the network was trained on (far field), and on the
    coarse run with the network after ``steps`` steps, near the body and in
    the far field."""
    from .airfoil_moving import DroneMovingIBM
    from .clbm_moving import CLBMSolver
    chord = coarse[1] / 2.5 if chord is None else chord
    log = log or (lambda s: print(f"[closure] {s}", file=sys.stderr, flush=True))
    res = {"seed": seed, "coarse": list(coarse), "Re": Re}
    pr, masks = fine_pairs(seed, coarse, chord, Re, warm, 1, gap, device=device, log=log)
    res["filtered_fine_far"] = diagnostics(pr[0][0], net, masks[0])
    log(f"filtered fine: {res['filtered_fine_far']}")
    cfg = case_for(seed, coarse, chord, Re)
    lbm = CLBMSolver(cfg, device=device)
    ibm = DroneMovingIBM(cfg, device=device)
    lbm.closure = net
    for n in range(1, steps + 1):
        fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(n))
        lbm.step(fx, fy, fz)
    res["coarse_finite"] = bool(lbm.health()[1])
    f = lbm.f.to_numpy()
    box = tuple(int(v) for v in ibm.force_box())
    res["coarse_far"] = diagnostics(f, net, _far_mask(cfg, box))
    res["coarse_body"] = diagnostics(f, net, body_mask(cfg, box))
    log(f"coarse far: {res['coarse_far']}")
    log(f"coarse body: {res['coarse_body']}")
    return res


def stability_probe(coarse=(600, 200, 50), chord=None, Re_list=(5000.0, 1000.0), steps: int = 10000,
                    check: int = 100, seed: int = 100, refine=(1, 2), net=None, device=None,
                    log=None) -> list:
    """The plain coupled method (or, on the coarse grid, the closure ``net``)
    at each Re on the coarse grid and its 2x refinement (``steps`` coarse
    steps, twice as many fine ones): the step at which the divergence guard
    of main_moving.py:143-145 (non-finite populations or rho_min < 0.3) first
    trips, checked every ``check`` coarse steps, or None if it never does."""
    import torch
    from .airfoil_moving import DroneMovingIBM
    from .clbm_moving import CLBMSolver
    chord = coarse[1] / 2.5 if chord is None else chord
    log = log or (lambda s: print(f"[closure] {s}", file=sys.stderr, flush=True))
    rows = []
    for Re in Re_list:
        for r in refine:
            cfg = case_for(seed, coarse, chord, Re, refine=r)
            lbm = CLBMSolver(cfg, device=device)
            ibm = DroneMovingIBM(cfg, device=device)
            lbm.closure = net if r == 1 else None
            t0 = time.perf_counter()
            diverged, rho_min = None, 1.0
            for n in range(1, r * steps + 1):
                fx, fy, fz = ibm.compute_force(lbm.ux, lbm.uy, *cfg.motion_state(n))
                lbm.step(fx, fy, fz)
                if n % (r * check) == 0 or n == r * steps:
                    rho_min, finite = lbm.health()
                    if not finite or rho_min < 0.3:
                        diverged = n
                        break
            torch.cuda.synchronize()
            row = dict(Re=Re, refine=r, grid=[cfg.Nx, cfg.Ny, cfg.Nz], chord=cfg.chord,
                       tau=cfg.tau, closure=lbm.closure is not None, steps=r * steps,
                       diverged_at=diverged, rho_min=rho_min, seconds=time.perf_counter() - t0)
            rows.append(row)
            log(f"probe {row}")
            del lbm, ibm
    return rows


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--coarse", default="300,100,24")
    ap.add_argument("--re", type=float, default=1000.0)
    ap.add_argument("--train-seeds", type=int, default=8)
    ap.add_argument("--test-seeds", type=int, default=4)
    ap.add_argument("--warm", type=int, default=4000)
    ap.add_argument("--pairs", type=int, default=3)
    ap.add_argument("--gap", type=int, default=400)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--eval-steps", type=int, default=4000)
    ap.add_argument("--json", default=None)
    ap.add_argument("--save", default=None, help="write the trained network (.npz)")
    ap.add_argument("--load", default=None, help="use this network (.npz) instead of training one")
    ap.add_argument("--stage", default="all", choices=["all", "train", "eval", "diagnose", "probe"],
                    help="train: data, training and the a-priori check; eval: the a-posteriori "
                         "runs only (needs --load); diagnose: the evaluation diagnostics of the "
                         "first held-out case (needs --load); probe: the stability probe "
                         "(with --load, the coarse runs use the network)")
    ap.add_argument("--train-region", default="far", choices=["far", "all"],
                    help="cells the training data come from: away from the body, or all")
    ap.add_argument("--probe-re", default="5000,4000,3000,2000,1000")
    ap.add_argument("--probe-steps", type=int, default=10000)
    ap.add_argument("--probe-refine", default="1,2", help="grids of the probe: 1 coarse, 2 fine")
    a = ap.parse_args(argv)
    coarse = tuple(int(v) for v in a.coarse.split(","))
    if a.stage in ("eval", "diagnose") and not a.load:
        ap.error(f"--stage {a.stage} needs --load")
    net = ClosureNet.load(a.load) if a.load else None
    if a.stage in ("diagnose", "probe"):
        if a.stage == "diagnose":
            res = diagnose(net, coarse, Re=a.re, seed=100, warm=a.warm, gap=a.gap,
                           steps=a.eval_steps)
        else:
            res = {"probe": stability_probe(coarse, Re_list=[float(v) for v in
                                                             a.probe_re.split(",")],
                                            steps=a.probe_steps, net=net,
                                            refine=[int(v) for v in a.probe_refine.split(",")])}
        res["stage"] = a.stage
        text = json.dumps(res, indent=1)
        if a.json:
            with open(a.json, "w") as fh:
                fh.write(text)
        print(text, flush=True)
        return 0
    res, net = study(coarse, Re=a.re, train_seeds=range(a.train_seeds),
                     test_seeds=range(100, 100 + a.test_seeds), warm=a.warm, pairs=a.pairs,
                     gap=a.gap, epochs=a.epochs, eval_steps=a.eval_steps, net=net,
                     train_region=a.train_region,
                     apriori=a.stage != "eval", aposteriori=a.stage != "train")
    res["stage"] = a.stage
    if a.save:
        net.save(a.save, coarse=list(coarse), Re=a.re)
    text = json.dumps(res, indent=1)
    if a.json:
        with open(a.json, "w") as fh:
            fh.write(text)
    print(text, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
