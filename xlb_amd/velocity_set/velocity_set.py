"""Lattice velocity sets (mirrors xlb/velocity_set/velocity_set.py:14-233).

All properties are NumPy arrays built exactly as the reference builds them (c as (d, q) int,
w (q,), opp_indices, cc (q, d(d+1)/2), qi, main/left/right index sets); the HIP kernels carry
the same tables as compile-time constants.
"""

import math

import numpy as np

from xlb_amd.compute_backend import ComputeBackend


class VelocitySet(object):
    lattice_id = None  # XLBH_D2Q9 / D3Q19 / D3Q27

    def __init__(self, d, q, c, w, precision_policy, compute_backend):
        if compute_backend is not None and compute_backend != ComputeBackend.MI355X:
            raise ValueError(f"Unsupported compute backend: {compute_backend} (this package implements ComputeBackend.MI355X)")
        self.d = d
        self.q = q
        self.precision_policy = precision_policy
        self.compute_backend = compute_backend or ComputeBackend.MI355X
        self._c = np.array(c)
        self._w = np.array(w)
        self.c = self._c
        self.w = self._w
        self.opp_indices = self._construct_opposite_indices()
        self.cc = self._construct_lattice_moment()
        self.c_float = self._c.astype(np.float64)
        self.qi = self._construct_qi()
        self.cs = np.float64(math.sqrt(3) / 3.0)
        self.cs2 = np.float64(1.0 / 3.0)
        self.inv_cs2 = np.float64(3.0)
        self.main_indices = self._construct_main_indices()
        self.right_indices = self._construct_right_indices()
        self.left_indices = self._construct_left_indices()

    def _construct_qi(self):
        # Qi = cc - cs^2 I, off-diagonal entries doubled (velocity_set.py:124-138)
        Qi = self.cc.copy()
        if self.d == 3:
            diagonal, offdiagonal = (0, 3, 5), (1, 2, 4)
        elif self.d == 2:
            diagonal, offdiagonal = (0, 2), (1,)
        else:
            raise ValueError(f"dim = {self.d} not supported")
        Qi[:, diagonal] += -1.0 / 3.0
        Qi[:, offdiagonal] *= 2.0
        return Qi

    def _construct_lattice_moment(self):
        # cc[:, k] = c_a c_b for a <= b in order xx, xy, xz, yy, yz, zz (velocity_set.py:140-164)
        c = self._c.T
        nt = self.d * (self.d + 1) // 2
        cc = np.zeros((self.q, nt))
        k = 0
        for a in range(self.d):
            for b in range(a, self.d):
                cc[:, k] = c[:, a] * c[:, b]
                k += 1
        return cc

    def _construct_opposite_indices(self):
        c = self._c.T
        return np.array([c.tolist().index((-c[i]).tolist()) for i in range(self.q)])

    def _construct_main_indices(self):
        c = self._c.T
        return np.nonzero(np.sum(np.abs(c), axis=1) == 1)[0]

    def _construct_right_indices(self):
        return np.nonzero(self._c.T[:, 0] == 1)[0]

    def _construct_left_indices(self):
        return np.nonzero(self._c.T[:, 0] == -1)[0]

    def __str__(self):
        return self.__repr__()

    def __repr__(self):
        return "D{}Q{}".format(self.d, self.q)
