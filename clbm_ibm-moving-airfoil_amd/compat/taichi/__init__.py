"""The part of the ``taichi`` module the reference drivers touch before they
build the solver objects: ``ti.init(arch=..., default_fp=..., ...)`` and the
``gpu``/``cpu``/``f32``/``i32`` names (main_moving.py:28-37,
main_highRe.py:32-41, fish/main_fish.py:26-35).  Fields, kernels and
everything else live in the HIP-backed modules; nothing else is provided.
Put this directory on ``sys.path`` only where the real Taichi is absent.
"""

gpu = "gpu"
cpu = "cpu"
f32 = "f32"
f64 = "f64"
i32 = "i32"
_settings = {}


def init(arch=None, **kwargs):
    """Record the requested settings.  The arithmetic precision of the
    HIP-backed solver is chosen by its storage argument, not here."""
    _settings.clear()
    _settings.update(kwargs, arch=arch)
    return None
