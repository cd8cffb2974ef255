"""Host-side behaviour of the operator API that needs no GPU."""

import pytest

import xlb_amd as xa


def _init(lat_name):
    policy = xa.PrecisionPolicy.FP32FP32
    vs = getattr(xa.velocity_set, lat_name)(policy, xa.ComputeBackend.MI355X)
    xa.init(velocity_set=vs, default_backend=xa.ComputeBackend.MI355X, default_precision_policy=policy)
    return vs


def test_grads_bc_is_3d_only_and_padded():
    """bc_grads_approximation.py:64-77: index-defined walls need padding and a mesh distance; 2-D
    raises NotImplementedError."""
    _init("D3Q19")
    bc = xa.operator.boundary_condition.GradsApproximationBC(indices=[[1, 2], [1, 2], [1, 2]])
    assert bc.needs_padding and bc.needs_mesh_distance
    assert bc.implementation_step == xa.operator.boundary_condition.ImplementationStep.STREAMING
    _init("D2Q9")
    with pytest.raises(NotImplementedError, match="2D"):
        xa.operator.boundary_condition.GradsApproximationBC(indices=[[1, 2], [1, 2]])


@pytest.mark.parametrize("shape", [(7, 5), (6, 5, 4), (2, 3), (1, 1, 2)])
def test_bounding_box_indices_contents_and_order(shape):
    """grid.py:34-86: every face lists the cells with its axis coordinate fixed at 0 or n-1, in C
    order, faces in the reference's order; remove_edges keeps 1 <= other coordinates <= n-2."""
    import itertools

    from xlb_amd.grid.grid import Grid

    d = len(shape)
    names = ["bottom", "top", "left", "right"] + (["front", "back"] if d == 3 else [])
    fixed = {"left": (0, 0), "right": (0, shape[0] - 1), "bottom": (d - 1, 0), "top": (d - 1, shape[d - 1] - 1),
             "front": (1, 0), "back": (1, shape[1] - 1)}
    for remove_edges in (False, True):
        box = Grid(shape, xa.ComputeBackend.MI355X).bounding_box_indices(remove_edges=remove_edges)
        assert list(box) == names
        for name in names:
            ax, val = fixed[name]
            want = [idx for idx in itertools.product(*[range(n) for n in shape]) if idx[ax] == val
                    and not (remove_edges and any(idx[a] in (0, shape[a] - 1) for a in range(d) if a != ax))]
            got = list(zip(*box[name])) if len(box[name][0]) else []
            assert len(box[name]) == d and got == want, (name, remove_edges)
            assert all(isinstance(v, int) for row in box[name] for v in row)


def test_bc_overlap_check():
    """check_boundary_overlaps.py:5-24 (Warp behaviour): a cell named twice by one boundary
    condition or by two of them raises ValueError; disjoint lists and mesh-defined BCs pass."""
    from xlb_amd.helper.check_boundary_overlaps import check_bc_overlaps

    class Bc:
        def __init__(self, indices):
            self.indices = indices

    check_bc_overlaps([Bc([[0, 1], [2, 2], [3, 3]]), Bc([[0], [2], [4]]), Bc(None)], 3, None)
    check_bc_overlaps([], 3, None)
    with pytest.raises(ValueError, match="Bc has duplicate indices"):
        check_bc_overlaps([Bc([[0, 1, 0], [2, 2, 2], [3, 3, 3]])], 3, None)
    with pytest.raises(ValueError, match="list contains duplicate indices"):
        check_bc_overlaps([Bc([[0, 1], [2, 2]]), Bc([[5, 1], [0, 2]])], 2, None)
