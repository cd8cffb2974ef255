"""D2Q9 (mirrors xlb/velocity_set/d2q9.py:8-24)."""

import numpy as np

from xlb_amd.velocity_set.velocity_set import VelocitySet


class D2Q9(VelocitySet):
    lattice_id = 0

    def __init__(self, precision_policy=None, compute_backend=None):
        cx = [0, 0, 0, 1, -1, 1, -1, 1, -1]
        cy = [0, 1, -1, 0, 1, -1, 0, 1, -1]
        c = np.array(tuple(zip(cx, cy))).T
        w = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 9, 1 / 36, 1 / 36])
        super().__init__(2, 9, c, w, precision_policy=precision_policy, compute_backend=compute_backend)
