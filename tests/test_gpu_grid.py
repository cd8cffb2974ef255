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


@pytest.mark.parametrize("policy", [xa.PrecisionPolicy.FP32FP32, xa.PrecisionPolicy.FP32FP16_SHIFTED,
                                    xa.PrecisionPolicy.FP32I16_BLOCK])
def test_padded_population_fields_index_like_contiguous_ones(policy):
    """Population fields from create_nse_fields use the library's plane stride (a (q, nx, ny, nz)
    view with strides (plane, ny*nz, nz, 1)); indexing, assignment and .cpu().numpy() behave as for
    a contiguous field, and a step on the padded layout equals the step on a dense copy bit for bit."""
    from xlb_amd import _lib
    from xlb_amd.workloads import periodic_box_3d

    wl = periodic_box_3d(16, policy)
    f = wl.f_0
    q, nx, ny, nz = f.shape
    plane = int(_lib.lib().xlbh_plane_stride(1, policy.store_format, nx, ny, nz))
    assert f.stride() == (plane, ny * nz, nz, 1) and plane >= nx * ny * nz
    dense = f.clone(memory_format=torch.contiguous_format)
    assert dense.stride() == (nx * ny * nz, ny * nz, nz, 1)
    np.testing.assert_array_equal(f.cpu().numpy(), dense.cpu().numpy())
    np.testing.assert_array_equal(f[3, 2:5, 1, ::2].cpu().numpy(), dense[3, 2:5, 1, ::2].cpu().numpy())
    if policy.block:
        # the block format needs its exponents: a dense copy carries them only through copy_population
        # into a field of the same layout; step the padded field against a second padded field
        a, b = f, wl.f_1
        xa.grid.copy_population(b, a)
        ref0 = xa.grid.decode_populations(a, wl.stepper.velocity_set.w, block=True)
        np.testing.assert_array_equal(xa.grid.decode_populations(b, wl.stepper.velocity_set.w, block=True).cpu().numpy(),
                                      ref0.cpu().numpy())
        return
    # the step on the padded layout and on the dense copy
    a, b = wl.stepper.run(f, wl.f_1, wl.bc_mask, wl.missing_mask, wl.omega, 3)
    d0, d1 = dense.clone(), torch.empty_like(dense)
    d0, d1 = wl.stepper.run(d0, d1, wl.bc_mask, wl.missing_mask, wl.omega, 3)
    np.testing.assert_array_equal(a.cpu().numpy(), d0.cpu().numpy())
    # assignment through the view reaches the planes, not the pad
    f2 = wl.f_1
    f2[5, 1, 2, 3] = 0.25
    assert float(f2.contiguous()[5, 1, 2, 3]) == 0.25
