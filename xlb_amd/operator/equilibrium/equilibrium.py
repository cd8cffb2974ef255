"""Equilibrium operators (mirrors xlb/operator/equilibrium/equilibrium.py and
quadratic_equilibrium.py:12-98)."""

import torch

from xlb_amd import _lib
from xlb_amd.compute_backend import ComputeBackend
from xlb_amd.operator.operator import Operator, expect_field, expect_population, field_extents
from xlb_amd.precision_policy import Precision


class Equilibrium(Operator):
    """Base class of equilibrium distributions."""


class QuadraticEquilibrium(Equilibrium):
    """feq_l = rho w_l (1 + 3 c_l.u + 4.5 (c_l.u)^2 - 1.5 |u|^2) (quadratic_equilibrium.py:35-60)."""

    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, rho, u, f):
        _lib.require_cuda(rho, u, f)
        if u.dtype != rho.dtype:
            u = u.to(rho.dtype)
        rho = rho.contiguous()
        u = u.contiguous()
        nx, ny, nz = field_extents(f)
        fmt = self.format_of(f)
        plane = expect_population(f, self.velocity_set.q, (nx, ny, nz), "f", f.dtype, self.lattice_id, fmt)
        expect_field(rho, 1, (nx, ny, nz), "rho")
        expect_field(u, self.velocity_set.d, (nx, ny, nz), "u")
        _lib.check(
            _lib.lib().xlbh_equilibrium(
                self.lattice_id, fmt, self.compute_id, Precision.from_torch(rho.dtype).field_id,
                _lib.ptr(rho), _lib.ptr(u), _lib.ptr(f), nx, ny, nz, plane, _lib.stream_of(f),
            ),
            "QuadraticEquilibrium",
        )
        _lib.mark_written(f)
        return f
