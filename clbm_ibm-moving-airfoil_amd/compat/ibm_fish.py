"""Import-compatible ``ibm_fish`` (fish/ibm_fish.py): FishIBM built from ``config_fish``."""

from _pkg import config_from, pkg


class FishIBM(pkg.FishIBM):
    def __init__(self):
        super().__init__(config_from("config_fish"))
