"""Import-compatible ``clbm_moving``: ``from clbm_moving import CLBMSolver``.

The reference's CLBMSolver (clbm_moving.py:44-343) takes no arguments and
reads ``config_moving`` at import time; this one reads the same module (or
whatever main_highRe.py:26-27 aliased under that name) when it is built.
"""

from _pkg import config_from, pkg


class CLBMSolver(pkg.CLBMSolver):
    def __init__(self):
        super().__init__(config_from("config_moving"))
