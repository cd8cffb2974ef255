"""Grids and field allocation (mirrors xlb/grid/grid.py:9-86 and warp_grid.py:10-32).

Fields are torch tensors in HBM with the reference's WarpGrid layout: C order
(cardinality, nx, ny, nz), z unit-stride, and (cardinality, nx, ny, 1) for 2-D grids.
"""

from typing import Tuple

import numpy as np
import torch

from xlb_amd.compute_backend import ComputeBackend
from xlb_amd.default_config import DefaultConfig


def grid_factory(shape: Tuple[int, ...], compute_backend: ComputeBackend = None):
    compute_backend = compute_backend or DefaultConfig.default_backend or ComputeBackend.MI355X
    if compute_backend == ComputeBackend.MI355X:
        return MI355XGrid(shape)
    raise ValueError(f"Compute backend {compute_backend} is not supported")


class Grid:
    def __init__(self, shape: Tuple[int, ...], compute_backend: ComputeBackend):
        self.shape = tuple(int(s) for s in shape)
        self.dim = len(self.shape)
        self.compute_backend = compute_backend

    def bounding_box_indices(self, remove_edges=False):
        """Cells of each face of the box as index lists (d rows), the interface of grid.py:34-86:
        faces "bottom"/"top" (last axis at 0 / n-1), "left"/"right" (x) and, in 3-D,
        "front"/"back" (y), in that order.  A face holds every cell whose coordinate on the face's
        axis is fixed, enumerated in C order; with remove_edges the other coordinates run over
        1 .. n-2 only."""
        if self.dim not in _FACES:
            raise ValueError(f"grid dimension {self.dim} not supported")
        trim = 1 if remove_edges else 0
        faces = {}
        for name, axis, far in _FACES[self.dim]:
            spans = [np.arange(trim, n - trim) for n in self.shape]
            spans[axis] = np.array([self.shape[axis] - 1 if far else 0])
            cells = np.stack(np.meshgrid(*spans, indexing="ij")).reshape(self.dim, -1)
            faces[name] = cells.tolist()
        return faces


# (face name, axis held fixed, far side?) per dimension, in the order the faces are listed.
_FACES = {
    2: (("bottom", 1, False), ("top", 1, True), ("left", 0, False), ("right", 0, True)),
    3: (("bottom", 2, False), ("top", 2, True), ("left", 0, False), ("right", 0, True), ("front", 1, False),
        ("back", 1, True)),
}


class MI355XGrid(Grid):
    def __init__(self, shape, device=None):
        super().__init__(shape, ComputeBackend.MI355X)
        self.device = torch.device(device) if device is not None else None

    @property
    def field_shape(self):
        return self.shape if len(self.shape) != 2 else self.shape + (1,)

    @property
    def nxyz(self):
        s = self.field_shape
        return int(s[0]), int(s[1]), int(s[2])

    def _device(self):
        if self.device is not None:
            return self.device
        return torch.device("cuda", torch.cuda.current_device())

    def create_field(self, cardinality: int, dtype=None, fill_value=None):
        """Allocate a (cardinality, *shape) field on the GPU (warp_grid.py:17-32)."""
        from xlb_amd.precision_policy import Precision

        if dtype is None:
            policy = DefaultConfig.default_precision_policy
            dtype = policy.store_precision if policy is not None else Precision.FP32
        tdt = dtype.torch_dtype if isinstance(dtype, Precision) else dtype
        shape = (int(cardinality),) + self.field_shape
        if fill_value is None:
            return torch.zeros(shape, dtype=tdt, device=self._device())
        return torch.full(shape, fill_value, dtype=tdt, device=self._device())
