"""MI355X-native CLBM-IBM time step (D3Q19 cascaded LBM + Peskin direct-forcing IBM).

Drop-in for the solver surface of kioxli/CLBM_IBM-moving-airfoil:
``CLBMSolver`` (clbm_moving.py), ``DroneMovingIBM`` (airfoil_moving.py) and
``FishIBM`` (fish/ibm_fish.py), backed by hand-written HIP kernels for gfx950
in ``libclbm_amd.so`` (C ABI: include/clbm_amd.h).  The directory name has
hyphens, so import it with ``importlib.import_module("clbm_ibm-moving-airfoil_amd")``
or put ``compat/`` on ``sys.path`` to get the reference's module names.
"""

from . import config
from ._lib import ClbmError, load as load_library
from .airfoil_moving import DroneMovingIBM, FishIBM
from .clbm_moving import CLBMSolver
from .config import CaseConfig

__all__ = ["CLBMSolver", "DroneMovingIBM", "FishIBM", "CaseConfig", "config", "ClbmError",
           "load_library"]
