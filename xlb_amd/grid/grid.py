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
        """Index lists of the 4 (2-D) or 6 (3-D) faces; grid.py:34-86 (same keys, order and lists)."""
        origin = np.array([0, 0, 0])
        bounds = np.array(self.shape)
        if remove_edges:
            origin += 1
            bounds -= 1
        slice_x = slice(origin[0], bounds[0])
        slice_y = slice(origin[1], bounds[1])
        dim = len(bounds)
        grid = np.indices(self.shape)
        if dim == 2:
            nx, ny = self.shape
            box = {
                "bottom": grid[:, slice_x, 0],
                "top": grid[:, slice_x, ny - 1],
                "left": grid[:, 0, slice_y],
                "right": grid[:, nx - 1, slice_y],
            }
        elif dim == 3:
            nx, ny, nz = self.shape
            slice_z = slice(origin[2], bounds[2])
            box = {
                "bottom": grid[:, slice_x, slice_y, 0].reshape(3, -1),
                "top": grid[:, slice_x, slice_y, nz - 1].reshape(3, -1),
                "left": grid[:, 0, slice_y, slice_z].reshape(3, -1),
                "right": grid[:, nx - 1, slice_y, slice_z].reshape(3, -1),
                "front": grid[:, slice_x, 0, slice_z].reshape(3, -1),
                "back": grid[:, slice_x, ny - 1, slice_z].reshape(3, -1),
            }
        else:
            raise ValueError(f"grid dimension {dim} not supported")
        return {k: v.tolist() for k, v in box.items()}


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
