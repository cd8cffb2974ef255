"""MeshBoundaryMasker (mirrors xlb/operator/boundary_masker/mesh_boundary_masker.py:12-236).

Surface voxelisation of a triangle mesh given as ``bc.mesh_vertices`` (an (3T, 3) array, every
three consecutive rows one triangle, in lattice units; the reference builds its Warp mesh with
indices = arange, mesh_boundary_masker.py:219-223).  A voxel the mesh touches becomes solid
(bc_mask = 255); a voxel next to a solid voxel in direction c_l receives the BC's id and
missing[opp l].  One HIP launch sequence (xlbh_mesh_mask); same argument checks and errors as the
reference, which also removes ``mesh_vertices`` from the BC once used.
"""

import numpy as np
import torch

from xlb_amd import _lib
from xlb_amd.compute_backend import ComputeBackend
from xlb_amd.operator.operator import Operator, expect_field, field_extents


class MeshBoundaryMasker(Operator):
    def __init__(self, velocity_set=None, precision_policy=None, compute_backend=None):
        super().__init__(velocity_set, precision_policy, compute_backend)
        if self.velocity_set.d == 2:
            raise NotImplementedError("This Operator is not implemented in 2D!")
        self._ws = None

    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, bc, bc_mask, missing_mask):
        assert bc.mesh_vertices is not None, (
            f'Please provide the mesh vertices for {bc.__class__.__name__} BC using keyword "mesh_vertices"!'
        )
        assert bc.indices is None, (
            f"Please use IndicesBoundaryMasker operator if {bc.__class__.__name__} is imposed on known indices of the grid!"
        )
        mesh_vertices = np.asarray(bc.mesh_vertices)
        assert mesh_vertices.ndim == 2 and mesh_vertices.shape[1] == self.velocity_set.d, (
            "Mesh points must be reshaped into an array (N, 3) where N indicates number of points!"
        )
        if mesh_vertices.shape[0] % 3 != 0:
            raise ValueError("mesh_vertices must hold three vertices per triangle (a triangle soup)")
        _lib.require_cuda(bc_mask, missing_mask)
        nx, ny, nz = field_extents(bc_mask)
        expect_field(bc_mask, 1, (nx, ny, nz), "bc_mask", torch.uint8)
        expect_field(missing_mask, self.velocity_set.q, (nx, ny, nz), "missing_mask", torch.bool)
        domain_shape = (nx, ny, nz)
        # the vertices go to the GPU as they are; the extent check (exact in any float type) and the
        # cast to float32 run there instead of on the host
        verts = torch.from_numpy(np.ascontiguousarray(mesh_vertices)).to(bc_mask.device)
        if mesh_vertices.shape[0]:
            lo, hi = torch.aminmax(verts, dim=0)
            ext = torch.stack([lo, hi]).cpu().numpy()
            mesh_min, mesh_max = ext[0], ext[1]
            if any(mesh_min < 0) or any(mesh_max >= domain_shape):
                raise ValueError(
                    f"Mesh extents ({mesh_min}, {mesh_max}) exceed domain dimensions {domain_shape}. "
                    "The mesh must be fully contained within the domain."
                )
        bc.__dict__.pop("mesh_vertices", None)
        bc.mesh_vertices = None
        assert not bc.needs_mesh_distance, 'Please use "MeshDistanceBoundaryMasker" if this BC needs mesh distance!'
        if not (1 <= int(bc.id) <= 254):
            raise ValueError(f"boundary id {bc.id} does not fit the uint8 bc_mask (1..254)")
        L = _lib.lib()
        nbytes = int(L.xlbh_mesh_workspace_bytes(nx, ny, nz))
        if self._ws is None or self._ws.numel() < nbytes or self._ws.device != bc_mask.device:
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=bc_mask.device)
        tri = verts.to(torch.float32).reshape(-1, 9).contiguous()
        del verts
        _lib.check(
            L.xlbh_mesh_mask(self.lattice_id, _lib.ptr(tri), int(tri.shape[0]), int(bc.id), _lib.ptr(bc_mask),
                             _lib.ptr(missing_mask), _lib.ptr(self._ws), nx, ny, nz, _lib.stream_of(bc_mask)),
            "MeshBoundaryMasker",
        )
        del tri
        _lib.mark_written(bc_mask, missing_mask)
        return bc_mask, missing_mask
