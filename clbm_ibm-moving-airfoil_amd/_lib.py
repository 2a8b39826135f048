"""ctypes binding of libclbm_amd.so (the C ABI declared in include/clbm_amd.h).

The library is loaded from this package's directory.  There is no fallback:
if the shared object is missing or does not load, every entry point raises.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclbm_amd.so")

STORE_F32, STORE_F16, STORE_BF16, STORE_F64, STORE_F16C = 0, 1, 2, 3, 4
STORAGE = {"f32": STORE_F32, "f16": STORE_F16, "bf16": STORE_BF16, "f64": STORE_F64,
           "f16c": STORE_F16C}
ELEM_BYTES = {STORE_F32: 4, STORE_F16: 2, STORE_BF16: 2, STORE_F64: 8, STORE_F16C: 2}
SPREAD_FIXED, SPREAD_ATOMIC_F32, SPREAD_SORTED = 0, 1, 2
SPREAD_MODES = {"fixed": SPREAD_FIXED, "atomic": SPREAD_ATOMIC_F32, "sorted": SPREAD_SORTED}
BOX_SUBSLOTS = 64
BOX_WORDS = 2 * 6 * BOX_SUBSLOTS + 32       # uint64 words of a force-box record
FORCE_BOX_OFFSET = 2 * 6 * BOX_SUBSLOTS     # word offset of the decoded int32[6] box
REGION_FLAG_OFFSET = 2 * 6 * BOX_SUBSLOTS + 3   # word offset of the region-violation flag
SPLIT_READY_OFFSET = 2 * 6 * BOX_SUBSLOTS + 16  # last generation published for clbm_step_fused
SPLIT_EXPIRED_OFFSET = 2 * 6 * BOX_SUBSLOTS + 17   # generations whose region wait ran out
ABI_VERSION = 15
CLOSURE_GRAD_NOALLOC_MAX_BLOCKS = 1024   # rows of the library's own partial-gradient array


class ClbmGrid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("storage", C.c_int32),
                ("dev_ux", C.c_float), ("dev_uy", C.c_float)]


class ClbmParams(C.Structure):
    _fields_ = [("s_v", C.c_double), ("s_e", C.c_double), ("ux_in", C.c_double),
                ("uy_in", C.c_double), ("inlet", C.c_int32), ("outside_lds", C.c_int32)]


class IbmParams(C.Structure):
    _fields_ = [("f_max", C.c_float), ("wrap_weight_quirk", C.c_int32), ("spread_mode", C.c_int32),
                ("fixed_shift", C.c_int32), ("group", C.c_int32), ("check_region", C.c_int32),
                ("region", C.c_int32 * 6)]


class IbmSplitPlan(C.Structure):
    _fields_ = [("grid", ClbmGrid), ("params", ClbmParams), ("ip", IbmParams), ("f", C.c_void_p),
                ("n", C.c_int32), ("order", C.c_void_p)] + [
        (k, C.c_void_p) for k in ("X0", "Y0", "Z0", "X", "Y", "Z", "VX", "VY", "U", "V", "FX",
                                  "FY", "stencil_base", "acc_x", "acc_y", "box", "fx", "fy",
                                  "hx", "hy")] + [
        ("nh", C.c_int32), ("z0", C.c_int32), ("z1", C.c_int32), ("side", C.c_void_p),
        ("fork_ev", C.c_void_p)]


class ClbmBatch(C.Structure):
    _fields_ = [("n", C.c_int32), ("storage", C.c_int32), ("block", C.c_int32),
                ("blocks", C.c_int32), ("args", C.c_void_p), ("inst", C.c_void_p),
                ("first", C.c_void_p), ("lds", C.c_int32)]


class IbmBuffers(C.Structure):
    _fields_ = [("n", C.c_int32), ("order", C.c_void_p)] + [
        (k, C.c_void_p) for k in ("X0", "Y0", "Z0", "X", "Y", "Z", "VX", "VY", "U", "V", "FX",
                                  "FY", "acc_x", "acc_y", "box", "fx", "fy", "kin")] + [
        ("gen0", C.c_uint32), ("regions", C.c_void_p)]


class IbmBatch(C.Structure):
    _fields_ = [("n", C.c_int32), ("storage", C.c_int32), ("spread_mode", C.c_int32),
                ("group", C.c_int32), ("blocks", C.c_int32), ("clip_blocks", C.c_int32),
                ("args", C.c_void_p), ("clip_args", C.c_void_p), ("inst", C.c_void_p),
                ("first", C.c_void_p), ("counter", C.c_void_p), ("regions", C.c_int32)]


_P = C.c_void_p
_F = C.c_float
_I = C.c_int
_G = C.POINTER(ClbmGrid)
_LP = C.POINTER(ClbmParams)
_IP = C.POINTER(IbmParams)
_SIGNATURES = {
    "clbm_abi_version": (C.c_int, []),
    "clbm_last_error": (C.c_char_p, []),
    "clbm_state_bytes": (C.c_size_t, [_G]),
    "clbm_state_after_step": (_I, [_I, _LP]),
    "clbm_init_feq": (_I, [_G, _P, C.c_double, C.c_double, C.c_double, C.c_double, _P]),
    "clbm_step": (_I, [_G, _LP, _P, _I, _P, _P, _P, _P, _P]),
    "clbm_step_part": (_I, [_G, _LP, _P, _I, _P, _P, _P, _P, _P, _I, _P]),
    "clbm_step_region": (_I, [_G, _LP, _P, _P]),
    "clbm_step_split": (_I, [_G, _LP, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "ibm_compute_force_split": (_I, [C.POINTER(IbmSplitPlan), _I, C.c_uint32, _F, _F, _F, _F, _F,
                                     _F, _P, _I, _P]),
    "clbm_step_fused": (_I, [_G, _LP, _P, _I, _P, _P, _P, _P, C.c_uint32, C.c_double, _P, _P, _P,
                             _P]),
    "ibm_step_region": (_I, [_G, _LP, _P, _P, _I, _I, _I, C.c_double, C.c_double, C.c_double,
                             _P]),
    "clbm_event_create": (_I, [C.POINTER(C.c_void_p)]),
    "clbm_event_destroy": (_I, [_P]),
    "clbm_stream_after": (_I, [_P, _P, _P]),
    "clbm_streams_concurrent": (_I, [_P, _P, C.POINTER(C.c_int)]),
    "clbm_side_stream": (_I, [_P, C.POINTER(C.c_void_p), C.POINTER(C.c_int)]),
    "clbm_macro": (_I, [_G, _LP, _P, _I, _P, _P, _P, _P, _P]),
    "clbm_export_state": (_I, [_G, _LP, _P, _I, _P, _P]),
    "clbm_export_state_f64": (_I, [_G, _LP, _P, _I, _P, _P]),
    "clbm_import_state": (_I, [_G, _P, _P, _P]),
    "clbm_import_state_f64": (_I, [_G, _P, _P, _P]),
    "clbm_health": (_I, [_G, _LP, _P, _I, _P, _P]),
    "clbm_stream_probe": (_I, [_G, _P, _I, _P]),
    "ibm_compute_force": (_I, [_G, _IP, _LP, _P, _I, _P, _P, _I, _P, _P, _P, _P, _F, _F, _F, _F,
                               _F, _F, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                               C.c_uint32, _P]),
    "ibm_update_markers": (_I, [_G, _I, _P, _P, _P, _F, _F, _F, _F, _F, _F, _P, _P, _P, _P, _P, _P,
                                C.c_uint32, _P]),
    "ibm_band_velocity": (_I, [_G, _LP, _P, _I, _P, C.c_uint32, _P, _P, _P]),
    "ibm_interp_force": (_I, [_G, _IP, _LP, _P, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                              _P, _P]),
    "ibm_spread": (_I, [_G, _IP, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "ibm_spread_sorted_bytes": (C.c_size_t, [_G, _I]),
    "ibm_spread_sorted": (_I, [_G, _IP, _I, _P, _P, _P, _P, _P, _P, _P, _P, C.c_size_t, _P]),
    "ibm_clip_forces": (_I, [_G, _IP, _P, _P, _P, _P, _P, C.c_uint32, _P]),
    "ibm_force_sum_f64": (_I, [_I, _P, _P, _P, _P, _P]),
    "clbm_ibm_step": (_I, [_G, _LP, _P, _I, _IP, _I, _P, _P, _P, _P, _F, _F, _F, _F, _F, _F,
                           _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_uint32, _P,
                           _I, _I, _I, _P]),
    "clbm_batch_bytes": (C.c_size_t, [_I, _G, _I]),
    "clbm_batch_build": (_I, [_I, _G, _LP, _P, _P, _P, _P, _I, _P, C.c_size_t,
                              C.POINTER(ClbmBatch)]),
    "clbm_batch_step": (_I, [C.POINTER(ClbmBatch), _I, _P]),
    "ibm_batch_bytes": (C.c_size_t, [_I, C.POINTER(IbmBuffers), _I]),
    "ibm_batch_build": (_I, [_I, _G, _LP, _P, _IP, C.POINTER(IbmBuffers), _P, C.c_size_t,
                             C.POINTER(IbmBatch)]),
    "ibm_batch_compute": (_I, [C.POINTER(IbmBatch), _I, _P]),
    "ibm_batch_markers": (_I, [C.POINTER(IbmBatch), _I, _P]),
    "clbm_batch_step_acc": (_I, [C.POINTER(ClbmBatch), C.POINTER(IbmBatch), _I, _P]),
    "closure_param_count": (_I, []),
    "closure_packed_bytes": (C.c_size_t, []),
    "closure_pack": (_I, [_P, _P, _P]),
    "closure_infer": (_I, [_P, _P, _I, _P, _P, _P]),
    "clbm_step_closure": (_I, [_G, _LP, _P, _I, _P, _P, _P, _P, _P, _P, _F, _F, _P]),
    "closure_partial_bytes": (C.c_size_t, [_I]),
    "closure_grad": (_I, [_P, _P, _I, _P, _P, _P, _I, _F, _P, _I, _P, _P]),
    "closure_grad_noalloc": (_I, [_P, _P, _I, _P, _P, _P, _I, _F, _I, _P, _P]),
    "closure_adam": (_I, [_P, _P, _P, _P, _F, _F, _F, _F, _I, _P]),
}
STATE_PARITY, STATE_PLANES = 1, 2

_lock = threading.Lock()
_lib = None


class ClbmError(RuntimeError):
    pass


def load():
    """Load the shared library once; raise if it is absent (no fallback path)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ClbmError(f"{LIB_PATH} is missing: build it with "
                            f"`python {os.path.join(HERE, 'build.py')}` (hipcc, gfx950)")
        # PyTorch brings its own HIP runtime.  Loaded first, it is the one the
        # library binds to (the same shared-object name), and the package's
        # device buffers and the library's launches share one runtime.  Loaded
        # after the library, torch's libraries would bind to the system HIP
        # runtime the library brought in, and the library's launches then found
        # no device (measured on the GPU host).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.clbm_abi_version() != ABI_VERSION:
            raise ClbmError("libclbm_amd.so ABI version mismatch")
        _lib = lib
        return lib


_fns = {}


def stream_of(device) -> int:
    """The hipStream_t (as an int) of torch's current stream on ``device``,
    as ``torch.cuda.current_stream(device).cuda_stream`` gives it, read with
    torch's raw accessor where it has one (a few microseconds less per call
    on the per-step host path, which queues a 16-bit step in about as long
    as the GPU takes to run it)."""
    import torch
    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if raw is None:
        return torch.cuda.current_stream(device).cuda_stream
    idx = device.index
    return raw(torch.cuda.current_device() if idx is None else idx)


_side = {}


def side_stream(device):
    """(handle, concurrent): the library's marker stream on ``device``
    (clbm_side_stream; one per device and process, shared by every body),
    as a ctypes pointer, and whether it passed the check that its kernels
    run beside those of torch's current stream rather than behind them on
    one hardware queue."""
    import torch
    idx = torch.cuda.current_device() if device.index is None else device.index
    got = _side.get(idx)
    if got is None:
        h, ok = C.c_void_p(), C.c_int()
        with torch.cuda.device(idx):
            call("clbm_side_stream", C.c_void_p(stream_of(device)), C.byref(h), C.byref(ok))
        got = _side[idx] = (h, bool(ok.value))
    return got


def fn(name: str):
    """The ctypes function object of an entry point (looked up once)."""
    f = _fns.get(name)
    if f is None:
        f = _fns[name] = getattr(load(), name)
    return f


def check(rc: int, name: str) -> None:
    """Raise ClbmError with the library's message when an entry point returned rc != 0."""
    if rc != 0:
        msg = load().clbm_last_error().decode(errors="replace")
        raise ClbmError(f"{name} failed ({rc}): {msg}")


def call(name: str, *args) -> None:
    """Call an entry point and raise ClbmError with the library's message on failure."""
    rc = fn(name)(*args)
    if rc != 0:
        check(rc, name)


def exported_symbols() -> list:
    return list(_SIGNATURES)
