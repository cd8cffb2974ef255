"""Grid fields (tests/grids/test_grid_warp.py): create_field shapes for 2-D and 3-D grids (2-D
fields carry a trailing axis of 1, warp_grid.py:26), fill values, and the bounding-box index sets
(grid.py:34-86) against their definition."""

import numpy as np
import pytest
import torch

import xlb_amd as xa
from xlb_amd.grid import grid_factory
from xlb_amd.precision_policy import Precision

pytestmark = pytest.mark.gpu


def init(vs_cls):
    P = xa.PrecisionPolicy.FP32FP32
    xa.init(velocity_set=vs_cls(P, xa.ComputeBackend.MI355X), default_backend=xa.ComputeBackend.MI355X,
            default_precision_policy=P)


@pytest.mark.parametrize("grid_size", [50, 100, 150])
def test_grid_create_field(grid_size):
    init(xa.velocity_set.D3Q19)
    for shape in [(grid_size, grid_size), (grid_size, grid_size, grid_size)]:
        f = grid_factory(shape).create_field(cardinality=9, dtype=Precision.FP32)
        want = (9,) + shape + ((1,) if len(shape) == 2 else ())
        assert tuple(f.shape) == want
        assert isinstance(f, torch.Tensor) and f.is_cuda and f.dtype == torch.float32


def test_grid_create_field_fill_value():
    init(xa.velocity_set.D2Q9)
    f = grid_factory((100, 100)).create_field(cardinality=9, dtype=Precision.FP32, fill_value=3.14)
    assert np.allclose(f.cpu().numpy(), 3.14)


@pytest.mark.parametrize("shape", [(7, 5), (6, 5, 4)])
def test_bounding_box_indices(shape):
    """Each face holds exactly the cells on it; remove_edges drops the cells that also lie on
    another face's boundary line (grid.py:55-57)."""
    init(xa.velocity_set.D3Q19 if len(shape) == 3 else xa.velocity_set.D2Q9)
    g = grid_factory(shape)
    d = len(shape)
    axis = {"left": (0, 0), "right": (0, shape[0] - 1), "bottom": (d - 1, 0), "top": (d - 1, shape[d - 1] - 1)}
    if d == 3:
        axis.update({"front": (1, 0), "back": (1, shape[1] - 1)})
    for remove_edges in (False, True):
        box = g.bounding_box_indices(remove_edges=remove_edges)
        assert set(box) == set(axis)
        for name, (ax, val) in axis.items():
            cells = set(zip(*[list(v) for v in box[name]]))
            want = set()
            for idx in np.ndindex(*shape):
                if idx[ax] != val:
                    continue
                if remove_edges and any(idx[a] in (0, shape[a] - 1) for a in range(d) if a != ax):
                    continue
                want.add(idx)
            assert cells == want, f"{name} remove_edges={remove_edges}"
