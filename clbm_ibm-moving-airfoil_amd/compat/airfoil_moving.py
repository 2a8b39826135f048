"""Import-compatible ``airfoil_moving``: ``from airfoil_moving import DroneMovingIBM``.

Mirrors the no-argument constructor of airfoil_moving.py:64-95 over the
HIP-backed DroneMovingIBM; the case comes from ``config_moving``.
"""

from _pkg import config_from, pkg


class DroneMovingIBM(pkg.DroneMovingIBM):
    def __init__(self):
        super().__init__(config_from("config_moving"))
