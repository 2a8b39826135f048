"""Import-compatible ``clbm_fish`` (fish/clbm_fish.py): the solver built from ``config_fish``."""

from _pkg import config_from, pkg


class CLBMSolver(pkg.CLBMSolver):
    def __init__(self):
        super().__init__(config_from("config_fish"))
