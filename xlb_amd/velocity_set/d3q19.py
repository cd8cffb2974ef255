"""D3Q19 (mirrors xlb/velocity_set/d3q19.py:9-30)."""

import itertools

import numpy as np

from xlb_amd.velocity_set.velocity_set import VelocitySet


class D3Q19(VelocitySet):
    lattice_id = 1

    def __init__(self, precision_policy=None, compute_backend=None):
        c = np.array([ci for ci in itertools.product([0, -1, 1], repeat=3) if np.sum(np.abs(ci)) <= 2]).T
        w = np.zeros(19)
        for i in range(19):
            s = np.sum(np.abs(c[:, i]))
            w[i] = {0: 1.0 / 3.0, 1: 1.0 / 18.0, 2: 1.0 / 36.0}[int(s)]
        super().__init__(3, 19, c, w, precision_policy=precision_policy, compute_backend=compute_backend)
