"""Moments of the populations (mirrors xlb/operator/macroscopic/macroscopic.py:13-66,
zero_moment.py:11-49, first_moment.py:11-67, second_moment.py:13-106)."""

from xlb_amd import _lib
from xlb_amd.compute_backend import ComputeBackend
from xlb_amd.operator.operator import Operator, expect_field, expect_population, field_extents
from xlb_amd.precision_policy import Precision


def _moments(op, which, f, rho, u, pi):
    _lib.require_cuda(f, rho, u, pi)
    ref = next(t for t in (rho, u, pi) if t is not None)
    for t in (rho, u, pi):
        if t is not None and t.dtype != ref.dtype:
            raise ValueError("rho, u and pi must share one dtype")
    nx, ny, nz = field_extents(f)
    vs = op.velocity_set
    fmt = op.format_of(f)
    plane = expect_population(f, vs.q, (nx, ny, nz), "f", f.dtype, op.lattice_id, fmt)
    if rho is not None:
        expect_field(rho, 1, (nx, ny, nz), "rho")
    if u is not None:
        expect_field(u, vs.d, (nx, ny, nz), "u")
    if pi is not None:
        expect_field(pi, vs.d * (vs.d + 1) // 2, (nx, ny, nz), "pi")
    _lib.check(
        _lib.lib().xlbh_moments(
            op.lattice_id, fmt, op.compute_id, Precision.from_torch(ref.dtype).field_id, which,
            _lib.ptr(f), _lib.ptr(rho), _lib.ptr(u), _lib.ptr(pi), nx, ny, nz, plane, _lib.stream_of(f),
        ),
        type(op).__name__,
    )
    _lib.mark_written(rho, u, pi)


class ZeroMoment(Operator):
    """rho = sum_l f_l."""

    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, f, rho):
        _moments(self, 1, f, rho, None, None)
        return rho


class FirstMoment(Operator):
    """u = sum_l c_l f_l / rho (rho given)."""

    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, f, rho, u):
        _moments(self, 2, f, rho, u, None)
        return u


class Macroscopic(Operator):
    """(rho, u) from f; outputs written in the dtype of the given rho/u fields."""

    def __init__(self, *args, **kwargs):
        self.zero_moment = ZeroMoment(*args, **kwargs)
        self.first_moment = FirstMoment(*args, **kwargs)
        super().__init__(*args, **kwargs)

    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, f, rho, u):
        _moments(self, 0, f, rho, u, None)
        return rho, u


class SecondMoment(Operator):
    """Pi_t = sum_l cc_lt f_l, components ordered xx, xy, xz, yy, yz, zz (xx, xy, yy in 2-D)."""

    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, f, pi):
        _moments(self, 3, f, None, None, pi)
        return pi
