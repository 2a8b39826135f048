"""ctypes wrapper of clbm_ref.c, the reference-structured C restatement.

TEST INFRASTRUCTURE AND CPU BASELINE ONLY (never imported by the product).
Arrays use the reference's layouts: populations (Nx, Ny, Nz, 19), fields
(Nx, Ny, Nz), float32.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libclbm_ref.so")
# the same source built with FMA contraction (the CPU baseline; not bit-exact)
LIB_FAST = os.path.join(HERE, "libclbm_ref_fast.so")
FAST_FLAGS = "-O3 -march=x86-64-v3 -mtune=generic -ffp-contract=fast -fopenmp -std=gnu11"
_libs = {}


def available(fast: bool = False) -> bool:
    return os.path.exists(LIB_FAST if fast else LIB)


def _load(fast: bool = False):
    path = LIB_FAST if fast else LIB
    if path not in _libs:
        lib = C.CDLL(path)
        P, F, D, I = C.c_void_p, C.c_float, C.c_double, C.c_int
        lib.ref_create.restype = P
        lib.ref_create.argtypes = [I, I, I, D, D, D, D, P, I, P, P, P, F, I]
        lib.ref_destroy.argtypes = [P]
        lib.ref_set_threads.argtypes = [I]
        lib.ref_init_feq.argtypes = [P]
        lib.ref_step.argtypes = [P]
        lib.ref_set_inlet.argtypes = [P, I]
        lib.ref_compute_force.argtypes = [P, F, F, F, F, F, F]
        for name in ("ref_get_f", "ref_set_f"):
            getattr(lib, name).argtypes = [P, P]
        lib.ref_get_macro.argtypes = [P, P, P, P, P]
        lib.ref_get_forces.argtypes = [P, P, P]
        lib.ref_get_markers.argtypes = [P, P]
        lib.ref_force_sum.argtypes = [P, P, P]
        _libs[path] = lib
    return _libs[path]


def _minv() -> np.ndarray:
    import clbm_oracle
    return np.ascontiguousarray(clbm_oracle.m_raw_inv_f32())


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


class RefSim:
    """One case (a CaseConfig-like object with Nx, Ny, Nz, s_v, s_e, ux_in,
    uy_in, X0, Y0, Z0, dV, f_max, wrap_weight_quirk, motion_state)."""

    def __init__(self, cfg, threads: int = 8, inlet=True, fast: bool = False):
        self._fast = fast
        lib = self._lib()
        self.cfg = cfg
        self.shape = (cfg.Nx, cfg.Ny, cfg.Nz)
        self._keep = [np.ascontiguousarray(a, dtype=np.float32) for a in (cfg.X0, cfg.Y0, cfg.Z0)]
        self._minv = _minv()
        lib.ref_set_threads(threads)
        self.h = lib.ref_create(cfg.Nx, cfg.Ny, cfg.Nz, cfg.s_v, cfg.s_e, cfg.ux_in, cfg.uy_in,
                                _ptr(self._minv), len(self._keep[0]), *[_ptr(a) for a in self._keep],
                                float(getattr(cfg, "f_max", 0.05)),
                                int(getattr(cfg, "wrap_weight_quirk", True)))
        self.dV = (None if getattr(cfg, "dV", None) is None
                   else np.ascontiguousarray(cfg.dV, dtype=np.float32))
        self.n = len(self._keep[0])
        # inlet: True for the reference's rule, "hecht-harting" for the opt-in one
        lib.ref_set_inlet(self.h, 1 if inlet == "hecht-harting" else 0)
        lib.ref_init_feq(self.h)

    def _lib(self):
        return _load(self._fast)

    def __del__(self):
        if getattr(self, "h", None):
            self._lib().ref_destroy(self.h)
            self.h = None

    def step_coupled(self, step: int):
        lib = self._lib()
        kin = self.cfg.motion_state(step)
        lib.ref_compute_force(self.h, *[float(k) for k in kin])
        lib.ref_step(self.h)

    def run(self, steps: int, first_step: int = 1, record: bool = False):
        lib = self._lib()
        hist = []
        out = np.zeros(2, np.float64)
        for n in range(first_step, first_step + steps):
            kin = self.cfg.motion_state(n)
            lib.ref_compute_force(self.h, *[float(k) for k in kin])
            if record:
                lib.ref_force_sum(self.h, _ptr(self.dV) if self.dV is not None else None, _ptr(out))
                hist.append(out.copy())
            lib.ref_step(self.h)
        return np.array(hist)

    def f(self) -> np.ndarray:
        a = np.empty(self.shape + (19,), np.float32)
        self._lib().ref_get_f(self.h, _ptr(a))
        return a

    def set_f(self, f) -> None:
        a = np.ascontiguousarray(f, dtype=np.float32)
        self._lib().ref_set_f(self.h, _ptr(a))

    def macro(self):
        out = [np.empty(self.shape, np.float32) for _ in range(4)]
        self._lib().ref_get_macro(self.h, *[_ptr(a) for a in out])
        return out

    def forces(self):
        out = [np.empty(self.shape, np.float32) for _ in range(2)]
        self._lib().ref_get_forces(self.h, *[_ptr(a) for a in out])
        return out

    def markers(self) -> dict:
        a = np.empty((9, self.n), np.float32)
        self._lib().ref_get_markers(self.h, _ptr(a))
        return dict(zip(("X", "Y", "Z", "VX", "VY", "U", "V", "FX", "FY"), a))
