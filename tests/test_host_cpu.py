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
