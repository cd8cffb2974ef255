"""D3Q27 (mirrors xlb/velocity_set/d3q27.py:9-32)."""

import itertools

import numpy as np

from xlb_amd.velocity_set.velocity_set import VelocitySet


class D3Q27(VelocitySet):
    lattice_id = 2

    def __init__(self, precision_policy=None, compute_backend=None):
        c = np.array(list(itertools.product([0, -1, 1], repeat=3))).T
        w = np.zeros(27)
        for i in range(27):
            s = np.sum(np.abs(c[:, i]))
            w[i] = {0: 8.0 / 27.0, 1: 2.0 / 27.0, 2: 1.0 / 54.0, 3: 1.0 / 216.0}[int(s)]
        super().__init__(3, 27, c, w, precision_policy=precision_policy, compute_backend=compute_backend)
