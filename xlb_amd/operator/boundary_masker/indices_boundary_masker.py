"""IndicesBoundaryMasker (mirrors xlb/operator/boundary_masker/indices_boundary_masker.py:12-224).

Warp semantics: for every in-domain index p of a BC, bc_mask[p] = id and missing[l, p] is set
for every direction l whose pull source p - c_l lies outside the domain; a BC that needs
padding and whose index is strictly interior also marks its push neighbours p + c_l
(missing[l, p + c_l], bc_mask[p + c_l] = id).  The BCs are applied one launch each in list
order, so where two BCs touch the same cell the later one's id wins (the reference's single
launch over all indices leaves that order to thread scheduling).
"""

import numpy as np
import torch

from xlb_amd import _lib
from xlb_amd.compute_backend import ComputeBackend
from xlb_amd.operator.boundary_masker.packing import remember_bits
from xlb_amd.operator.operator import Operator, expect_field, field_extents


class IndicesBoundaryMasker(Operator):
    def are_indices_in_interior(self, indices, shape):
        """Per index: inside the domain and not on its boundary (indices_boundary_masker.py:29-40)."""
        d = self.velocity_set.d
        shape_array = np.array(shape)
        return np.all((indices[:d] > 0) & (indices[:d] < shape_array[:d, np.newaxis] - 1), axis=0)

    def pack_bc(self, bc, bc_mask_shape, start_index=None):
        """Host packing of one BC: (3, n) int32 indices and (n,) uint8 interior flags."""
        assert bc.indices is not None, f'Please specify indices associated with the {bc.__class__.__name__} BC using keyword "indices"!'
        assert bc.mesh_vertices is None, f"Please use MeshBoundaryMasker operator if {bc.__class__.__name__} is imposed on a mesh (e.g. STL)!"
        bc_indices = np.asarray(bc.indices)
        if bc_indices.ndim != 2:
            bc_indices = bc_indices.reshape(len(bc.indices), -1)
        num = bc_indices.shape[1]
        if start_index is not None:
            bc_indices = bc_indices - np.array(start_index)[:, np.newaxis]
        if bc_indices.shape[0] == 2:
            bc_indices = np.vstack([bc_indices, np.zeros(num, dtype=int)])
        if bc.needs_padding:
            interior = self.are_indices_in_interior(bc_indices, bc_mask_shape)
        else:
            interior = np.zeros(num, dtype=bool)
        return np.ascontiguousarray(bc_indices, dtype=np.int32), interior.astype(np.uint8)

    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, bclist, bc_mask, missing_mask, start_index=None):
        _lib.require_cuda(bc_mask, missing_mask)
        nx, ny, nz = field_extents(bc_mask)
        expect_field(bc_mask, 1, (nx, ny, nz), "bc_mask", torch.uint8)
        expect_field(missing_mask, self.velocity_set.q, (nx, ny, nz), "missing_mask", torch.bool)
        from xlb_amd.operator.boundary_masker.packing import packed_missing

        bits = packed_missing(missing_mask, self.lattice_id).clone()
        dev = bc_mask.device
        L = _lib.lib()
        for bc in bclist:
            if not (1 <= int(bc.id) <= 254):
                raise ValueError(f"boundary id {bc.id} of {bc.__class__.__name__} does not fit the uint8 bc_mask (1..254)")
            idx, interior = self.pack_bc(bc, tuple(bc_mask[0].shape), start_index)
            n = idx.shape[1]
            if n == 0:
                continue
            d_idx = torch.from_numpy(idx).to(dev)
            d_int = torch.from_numpy(interior).to(dev)
            _lib.check(
                L.xlbh_mask_indices(self.lattice_id, _lib.ptr(d_idx), _lib.ptr(d_int), n, int(bc.id), _lib.ptr(bc_mask),
                                    _lib.ptr(missing_mask), _lib.ptr(bits), nx, ny, nz, _lib.stream_of(bc_mask)),
                "IndicesBoundaryMasker",
            )
            # keep the host copies alive until the copies above are enqueued
            del d_idx, d_int
        _lib.mark_written(bc_mask, missing_mask)
        remember_bits(missing_mask, self.lattice_id, bits)
        return bc_mask, missing_mask
