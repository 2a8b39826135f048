"""The C ABI: the shared library loads on a CPU-only host, exports every
function include/clbm_amd.h declares, and validates arguments before it
launches anything (no GPU call is made here)."""

import ctypes as C
import os
import re

import pytest

from conftest import ROOT, load_sub

HEADER = os.path.join(ROOT, "include", "clbm_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*([a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n.startswith(("clbm_", "ibm_", "closure_"))))


def test_header_declares_the_expected_surface():
    names = declared_functions()
    for n in ("clbm_step", "clbm_init_feq", "clbm_macro", "ibm_update_markers", "ibm_compute_force",
              "ibm_spread", "ibm_clip_forces", "ibm_force_sum_f64", "ibm_band_velocity"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = load_sub("_lib")
    so = lib.load()
    names = declared_functions()
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing
    # the ctypes table binds exactly the declared functions
    assert sorted(lib.exported_symbols()) == names


def test_abi_version_and_state_bytes():
    lib = load_sub("_lib")
    so = lib.load()
    assert so.clbm_abi_version() == lib.ABI_VERSION
    g = lib.ClbmGrid(600, 200, 50, lib.STORE_F32)
    assert so.clbm_state_bytes(C.byref(g)) == 19 * 600 * 200 * 50 * 4
    g = lib.ClbmGrid(600, 200, 50, lib.STORE_F16)
    assert so.clbm_state_bytes(C.byref(g)) == 19 * 600 * 200 * 50 * 2


def test_state_after_step():
    lib = load_sub("_lib")
    so = lib.load()
    p = lib.ClbmParams(1.7, 1.0, 0.05, 0.0, 1, 0)
    q = lib.ClbmParams(1.7, 1.0, 0.05, 0.0, 0, 0)
    assert so.clbm_state_after_step(0, C.byref(p)) == lib.STATE_PARITY | lib.STATE_PLANES
    assert so.clbm_state_after_step(lib.STATE_PARITY | lib.STATE_PLANES, C.byref(p)) == lib.STATE_PLANES
    assert so.clbm_state_after_step(lib.STATE_PARITY | lib.STATE_PLANES, C.byref(q)) == 0
    hh = lib.ClbmParams(1.7, 1.0, 0.05, 0.0, 2, 0)      # Hecht-Harting inlet: same planes
    assert so.clbm_state_after_step(0, C.byref(hh)) == lib.STATE_PARITY | lib.STATE_PLANES


def test_solver_rejects_an_unknown_inlet(pkg):
    with pytest.raises(ValueError, match="inlet"):
        pkg.CLBMSolver(pkg.config.naca_case(16, 8, 4, 4.0), inlet="zou-he")


@pytest.mark.parametrize("call,match", [
    (lambda lib: lib.call("clbm_step", None, None, None, 0, None, None, None, None, None), "params"),
    (lambda lib: lib.call("clbm_init_feq", C.byref(lib.ClbmGrid(0, 4, 4, 0)), C.c_void_p(1),
                          1.0, 0.0, 0.0, 0.0, None), "positive"),
    (lambda lib: lib.call("clbm_init_feq", C.byref(lib.ClbmGrid(4, 4, 4, 9)), C.c_void_p(1),
                          1.0, 0.0, 0.0, 0.0, None), "storage"),
    (lambda lib: lib.call("clbm_init_feq", C.byref(lib.ClbmGrid(2048, 2048, 1024, 0)),
                          C.c_void_p(1), 1.0, 0.0, 0.0, 0.0, None), "2\\^31"),
    (lambda lib: lib.call("clbm_macro", C.byref(lib.ClbmGrid(8, 4, 4, 0)), None, None, 0, None,
                          None, None, None, None), "NULL"),
    (lambda lib: lib.call("ibm_update_markers", C.byref(lib.ClbmGrid(8, 4, 4, 0)), -1,
                          *([None] * 3), *([0.0] * 6), *([None] * 6), 1, None), "n < 0"),
    (lambda lib: lib.call("ibm_force_sum_f64", 5, None, None, None, None, None), "NULL"),
    # one float32 population plane of 1.2 G cells exceeds the 4 GiB of 32-bit offsets
    (lambda lib: lib.call("clbm_init_feq", C.byref(lib.ClbmGrid(1200, 1000, 1000, 0)),
                          C.c_void_p(1), 1.0, 0.0, 0.0, 0.0, None), "4 GiB"),
    (lambda lib: lib.call("ibm_step_region", C.byref(lib.ClbmGrid(8, 4, 4, 0)), None, None, None,
                          3, 0, 3, 1.0, 1.0, 0.0, None), "NULL"),
    (lambda lib: lib.call("ibm_step_region", C.byref(lib.ClbmGrid(8, 4, 4, 0)),
                          C.byref(lib.ClbmParams(1.7, 1.0, 0.05, 0.0, 1, 0)), None, None, 0, 3,
                          1, 1.0, 1.0, 0.0, (C.c_int32 * 6)()), "z range"),
    # inlet rules: 0, 1 and 2 only, and an inlet needs ux_in < 1 (no launch happens)
    (lambda lib: lib.call("clbm_step", C.byref(lib.ClbmGrid(8, 4, 4, 0)),
                          C.byref(lib.ClbmParams(1.7, 1.0, 0.05, 0.0, 3, 0)), C.c_void_p(1), 0,
                          None, None, None, None, None), "inlet must be"),
    (lambda lib: lib.call("clbm_step", C.byref(lib.ClbmGrid(8, 4, 4, 0)),
                          C.byref(lib.ClbmParams(1.7, 1.0, 1.0, 0.0, 2, 0)), C.c_void_p(1), 0,
                          None, None, None, None, None), "ux_in < 1"),
    (lambda lib: lib.call("clbm_stream_probe", C.byref(lib.ClbmGrid(8, 4, 4, 0)), C.c_void_p(1),
                          2, None), "parity"),
    (lambda lib: lib.call("clbm_stream_after", None, None, None), "NULL event"),
    (lambda lib: lib.call("clbm_event_create", None), "NULL"),
])
def test_argument_errors_are_reported(call, match):
    lib = load_sub("_lib")
    with pytest.raises(lib.ClbmError, match=match):
        call(lib)


def test_product_does_not_import_the_oracle():
    """The package never reaches for the CPU oracle (no fallback path)."""
    pkg_dir = os.path.join(ROOT, "clbm_ibm-moving-airfoil_amd")
    for dirpath, _, files in os.walk(pkg_dir):
        for fn in files:
            if fn.endswith(".py"):
                src = open(os.path.join(dirpath, fn)).read()
                assert "clbm_oracle" not in src and "clbm_ref" not in src, fn


@pytest.mark.gpu
def test_library_loaded_before_torch_touches_the_gpu_still_launches(gpu):
    """A fresh process that loads the library before anything else (its host
    entry points need no GPU) and only then uses torch's GPU: the library's
    launches find the device (it binds to the HIP runtime torch brings)."""
    import subprocess
    import sys
    code = (
        "import importlib, sys\n"
        f"sys.path.insert(0, {ROOT!r})\n"
        "lib = importlib.import_module('clbm_ibm-moving-airfoil_amd._lib')\n"
        "lib.load()\n"
        "import torch\n"
        "pkg = importlib.import_module('clbm_ibm-moving-airfoil_amd')\n"
        "s = pkg.CLBMSolver(pkg.config.naca_case(40, 24, 4, 8.0, name='order'))\n"
        "s.step()\n"
        "rho_min, finite = s.health()\n"
        "assert finite and abs(rho_min - 1.0) < 0.01, (rho_min, finite)\n"
        "print('ok')\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr[-2000:]
