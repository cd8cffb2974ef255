"""MeshBoundaryMasker (xlbh_mesh_mask) against the oracle's voxel-parallel restatement of
mesh_boundary_masker.py:49-236: masks bit-exact for box, sphere and random-soup meshes on D3Q19
and D3Q27, alone and on top of an indices-masked state; and a wind tunnel with a mesh sphere
stepping like the oracle from the same masks."""

import numpy as np
import pytest
import torch

import xlb_amd as xa
from lbm_cases import oracle_aux, oracle_masks
from mesh_cases import box_mesh, icosphere, random_soup
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SHAPE = (24, 20, 18)
MESHES = {
    "box": lambda: box_mesh((4.3, 5.6, 3.2), (15.4, 13.7, 12.1)),
    "box_on_grid": lambda: box_mesh((4.0, 5.0, 3.0), (15.0, 13.0, 12.0)),
    "sphere": lambda: icosphere((11.7, 9.6, 8.9), 6.3, 3),
    "soup": lambda: random_soup(SHAPE, 300, seed=5),
}


def setup(lat_name, policy=None):
    policy = policy or xa.PrecisionPolicy.FP32FP32
    vs = getattr(xa.velocity_set, lat_name)(policy, xa.ComputeBackend.MI355X)
    xa.init(velocity_set=vs, default_backend=xa.ComputeBackend.MI355X, default_precision_policy=policy)
    xa.operator.boundary_condition.boundary_condition_registry.reset()
    return vs


@pytest.mark.parametrize("lat_name", ["D3Q19", "D3Q27"])
@pytest.mark.parametrize("mesh", sorted(MESHES))
def test_mesh_masker_matches_oracle(lat_name, mesh):
    vs = setup(lat_name)
    tris = MESHES[mesh]()
    bc = xa.operator.boundary_condition.HalfwayBounceBackBC(mesh_vertices=tris.reshape(-1, 3))
    bc_mask = torch.zeros((1,) + SHAPE, dtype=torch.uint8, device="cuda")
    missing = torch.zeros((vs.q,) + SHAPE, dtype=torch.bool, device="cuda")
    masker = xa.operator.boundary_masker.MeshBoundaryMasker(vs, xa.PrecisionPolicy.FP32FP32, xa.ComputeBackend.MI355X)
    bc_mask, missing = masker(bc, bc_mask, missing)
    o_mask = np.zeros((1,) + SHAPE, np.uint8)
    o_missing = np.zeros((vs.q,) + SHAPE, np.uint8)
    orc.mesh_mask(vs.lattice_id, tris, bc.id, o_mask, o_missing)
    assert (o_mask == 255).sum() > 0 and (o_mask == bc.id).sum() > 0
    np.testing.assert_array_equal(bc_mask.cpu().numpy(), o_mask)
    np.testing.assert_array_equal(missing.cpu().numpy().astype(np.uint8), o_missing)
    assert bc.mesh_vertices is None


def test_mesh_on_top_of_indices_masks():
    vs = setup("D3Q19")
    grid = xa.grid.grid_factory(SHAPE, compute_backend=xa.ComputeBackend.MI355X)
    BC = xa.operator.boundary_condition
    box = grid.bounding_box_indices()
    walls = np.unique(np.array([box["bottom"][i] + box["top"][i] for i in range(3)]), axis=-1).tolist()
    wall_bc = BC.FullwayBounceBackBC(indices=walls)
    tris = MESHES["sphere"]()
    mesh_bc = BC.HalfwayBounceBackBC(mesh_vertices=tris.reshape(-1, 3))
    st = xa.operator.stepper.IncompressibleNavierStokesStepper(grid, [wall_bc, mesh_bc], collision_type="BGK")
    _, _, bc_mask, missing = st.prepare_fields()
    o_mask, o_missing = oracle_masks(vs.lattice_id, vs.q, vs.d, SHAPE, [wall_bc])
    orc.mesh_mask(vs.lattice_id, tris, mesh_bc.id, o_mask, o_missing)
    np.testing.assert_array_equal(bc_mask.cpu().numpy(), o_mask)
    np.testing.assert_array_equal(missing.cpu().numpy().astype(np.uint8), o_missing)


def test_wind_tunnel_with_mesh_sphere_steps_like_oracle():
    """Regularized inlet, extrapolation outlet, fullway walls and a halfway-bounce-back sphere given
    as a mesh (flow_past_sphere_3d.py / windtunnel_3d.py:95 geometry), FP64, 20 steps."""
    policy = xa.PrecisionPolicy.FP64FP64
    vs = setup("D3Q19", policy)
    shape = (32, 20, 20)
    grid = xa.grid.grid_factory(shape, compute_backend=xa.ComputeBackend.MI355X)
    BC = xa.operator.boundary_condition
    box = grid.bounding_box_indices()
    ne = grid.bounding_box_indices(remove_edges=True)
    walls = np.unique(np.array([box["bottom"][i] + box["top"][i] + box["front"][i] + box["back"][i]
                                for i in range(3)]), axis=-1).tolist()
    tris = icosphere((10.3, 9.7, 10.1), 4.2, 2)
    bcs = [BC.FullwayBounceBackBC(indices=walls),
           BC.RegularizedBC("velocity", prescribed_value=(0.04, 0.0, 0.0), indices=ne["left"]),
           BC.ExtrapolationOutflowBC(indices=ne["right"]),
           BC.HalfwayBounceBackBC(mesh_vertices=tris.reshape(-1, 3))]
    descs = [dict(id=int(b.id), kind=int(b.kind)) for b in bcs]
    st = xa.operator.stepper.IncompressibleNavierStokesStepper(grid, bcs, collision_type="BGK")
    f0, f1, bc_mask, missing = st.prepare_fields()
    o_mask, o_missing = oracle_masks(vs.lattice_id, vs.q, vs.d, shape, bcs[:3])
    orc.mesh_mask(vs.lattice_id, tris, bcs[3].id, o_mask, o_missing)
    np.testing.assert_array_equal(bc_mask.cpu().numpy(), o_mask)
    omega = 1.7
    o_f0 = f0.cpu().numpy().copy()
    o_f1 = o_f0.copy()
    orc.aux_init(orc.F64, o_f1, o_mask, bcs[1].id, oracle_aux(bcs[1], o_mask, vs.d, policy))
    for t in range(20):
        f0, f1 = st(f0, f1, bc_mask, missing, omega, t)
        f0, f1 = f1, f0
        orc.step(vs.lattice_id, 0, orc.F64, 1, o_f0, o_f1, o_mask, o_missing, descs, omega)
        o_f0, o_f1 = o_f1, o_f0
    solid = o_mask[0] == 255
    err = np.abs(f0.cpu().numpy() - o_f0)[:, ~solid].max()
    assert err <= 1e-12, f"max|df| {err:.3e}"


def test_mesh_masker_rejects_bad_input():
    vs = setup("D3Q19")
    masker = xa.operator.boundary_masker.MeshBoundaryMasker(vs, xa.PrecisionPolicy.FP32FP32, xa.ComputeBackend.MI355X)
    bc_mask = torch.zeros((1,) + SHAPE, dtype=torch.uint8, device="cuda")
    missing = torch.zeros((vs.q,) + SHAPE, dtype=torch.bool, device="cuda")
    outside = box_mesh((-1.0, 2.0, 2.0), (5.0, 5.0, 5.0)).reshape(-1, 3)
    # the operator dispatch re-raises as a generic Exception carrying the message (operator.py:54-74)
    with pytest.raises(Exception, match="exceed domain dimensions"):
        masker(xa.operator.boundary_condition.HalfwayBounceBackBC(mesh_vertices=outside), bc_mask, missing)
