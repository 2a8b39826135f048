"""Locate the HIP-backed package from the import-compatible modules beside this file."""

import importlib
import os
import sys

_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if _ROOT not in sys.path:
    sys.path.insert(0, _ROOT)
pkg = importlib.import_module("clbm_ibm-moving-airfoil_amd")


def config_from(module_name: str):
    """The case of the config module the driver imported (or aliased in sys.modules)."""
    mod = sys.modules.get(module_name)
    if mod is None:
        mod = importlib.import_module(module_name)
    return pkg.config.from_module(mod)
