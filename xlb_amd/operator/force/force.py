"""Forces (mirrors xlb/operator/force/momentum_transfer.py:14-176 and
exact_difference_force.py:13-123)."""

import numpy as np
import torch

from xlb_amd import _lib
from xlb_amd.compute_backend import ComputeBackend
from xlb_amd.operator.boundary_masker.packing import packed_missing
from xlb_amd.operator.operator import Operator, expect_field, expect_population, field_extents


class MomentumTransfer(Operator):
    """Momentum-exchange force on the cells of one no-slip BC (momentum_transfer.py:92-176).

    Call with the newest post-collision field as f_0 (after the caller's buffer swap), as the
    reference does.  Returns a float64 numpy array of length d.  The sum runs in FP64 with a
    fixed reduction order (the reference accumulates FP32 atomics in arbitrary order).
    """

    def __init__(self, no_slip_bc_instance, velocity_set=None, precision_policy=None, compute_backend=None):
        self.no_slip_bc_instance = no_slip_bc_instance
        super().__init__(velocity_set, precision_policy, compute_backend)
        self._ws = None
        self._out = None

    def force_tensor(self, f_0, f_1, bc_mask, missing_mask):
        """Same as calling the operator but returns a (d,) float64 device tensor without a host sync."""
        _lib.require_cuda(f_0, bc_mask, missing_mask)
        nx, ny, nz = field_extents(f_0)
        q, ext = self.velocity_set.q, (nx, ny, nz)
        fmt = self.format_of(f_0)
        plane = expect_population(f_0, q, ext, "f_0", f_0.dtype, self.lattice_id, fmt)
        expect_field(bc_mask, 1, ext, "bc_mask", torch.uint8)
        expect_field(missing_mask, q, ext, "missing_mask", torch.bool)
        N = nx * ny * nz
        L = _lib.lib()
        nbytes = int(L.xlbh_reduce_workspace_bytes(N))
        if self._ws is None or self._ws.numel() < nbytes or self._ws.device != f_0.device:
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=f_0.device)
        out = torch.zeros(4, dtype=torch.float64, device=f_0.device)
        bc = self.no_slip_bc_instance
        bits = packed_missing(missing_mask, self.lattice_id)
        aux = getattr(bc, "aux", None)
        if bc.needs_aux_init and (aux is None or aux.numel() != N):
            bc.aux_data_init(None, None, bc_mask, missing_mask)
            aux = bc.aux
        _lib.check(
            L.xlbh_momentum_transfer(
                self.lattice_id, fmt, self.compute_id, bc.kind, int(bc.id), bc.bc_desc(),
                _lib.ptr(f_0), _lib.ptr(bc_mask), _lib.ptr(bits), _lib.ptr(aux), _lib.ptr(out), _lib.ptr(self._ws),
                nx, ny, nz, plane, _lib.stream_of(f_0),
            ),
            "MomentumTransfer",
        )
        return out[: self.velocity_set.d]

    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, f_0, f_1, bc_mask, missing_mask):
        return self.force_tensor(f_0, f_1, bc_mask, missing_mask).cpu().numpy()


class ExactDifference(Operator):
    """f_post += feq(rho, u + F) - feq(rho, u)  (exact_difference_force.py:45-86)."""

    def __init__(self, force_vector, equilibrium=None, velocity_set=None, precision_policy=None, compute_backend=None):
        self.force_vector = np.asarray(force_vector, dtype=np.float64)
        super().__init__(velocity_set, precision_policy, compute_backend)

    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, f_postcollision, feq, fout, rho, u):
        # A BGK collision with omega = 0 is the identity, so this is the forcing term alone.
        from xlb_amd.operator.collision.collision import _collide

        return _collide(self, _lib.BGK, f_postcollision, feq, fout, rho, u, 0.0, _lib.FORCE_UNIFORM, self.force_vector)
