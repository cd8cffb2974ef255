"""Collision operators (mirrors xlb/operator/collision/collision.py, bgk.py:12-81,
kbc.py:18-347, forced_collision.py:13-104)."""

import numpy as np

from xlb_amd import _lib
from xlb_amd.compute_backend import ComputeBackend
from xlb_amd.operator.operator import Operator, expect_field, field_extents, shared_layout
from xlb_amd.precision_policy import Precision


def _collide(op, collision, f, feq, fout, rho, u, omega, force_mode=_lib.FORCE_NONE, force_vector=None, force_field=None,
             collision_param=0.0):
    _lib.require_cuda(f, feq, fout, rho, u, force_field)
    nx, ny, nz = field_extents(f)
    vs = op.velocity_set
    ext = (nx, ny, nz)
    for t, name in ((f, "f"), (feq, "feq"), (fout, "fout")):
        if t is None or tuple(t.shape) != (vs.q,) + ext or t.dtype != f.dtype:
            raise ValueError(f"{name} must be a {(vs.q,) + ext} field of dtype {f.dtype}")
    plane, (f_, feq_, fout_) = shared_layout([f, feq, fout])
    if collision == _lib.KBC or force_mode != _lib.FORCE_NONE:
        expect_field(rho, 1, ext, "rho")
        expect_field(u, vs.d, ext, "u")
    else:
        rho, u = None, None
    if force_mode == _lib.FORCE_FIELD:
        expect_field(force_field, vs.d, ext, "force_field")
    field = rho if rho is not None else f
    field_id = Precision.from_torch(field.dtype).field_id if rho is not None else _lib.FIELD_F32
    if u is not None and rho is not None and u.dtype != rho.dtype:
        u = u.to(rho.dtype)
    fu_ptr, _keep = _lib.doubles(force_vector if force_vector is not None else (0.0, 0.0, 0.0))
    if force_field is not None and force_field.dtype != op.compute_dtype:
        force_field = force_field.to(op.compute_dtype)
    _lib.check(
        _lib.lib().xlbh_collide(
            op.lattice_id, collision, op.format_of(f), op.compute_id, field_id,
            _lib.ptr(f_), _lib.ptr(feq_), _lib.ptr(fout_), _lib.ptr(rho), _lib.ptr(u), float(omega),
            float(collision_param), force_mode, fu_ptr, _lib.ptr(force_field), nx, ny, nz, plane, _lib.stream_of(f),
        ),
        type(op).__name__,
    )
    if fout_ is not fout:
        fout.copy_(fout_)
    _lib.mark_written(fout)
    return fout


class Collision(Operator):
    """Base class of collision operators."""

    collision_id = None
    collision_param = 0.0


class BGK(Collision):
    """fout = f - omega (f - feq)  (bgk.py:30-34)."""

    collision_id = _lib.BGK

    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, f, feq, fout, rho, u, omega):
        return _collide(self, _lib.BGK, f, feq, fout, None, None, omega)


class KBC(Collision):
    """Karlin-Boesch-Chikatamarla entropic collision (kbc.py:192-296); D2Q9 and D3Q27 only."""

    collision_id = _lib.KBC

    def __init__(self, velocity_set=None, precision_policy=None, compute_backend=None):
        self.epsilon = 1e-32
        super().__init__(velocity_set=velocity_set, precision_policy=precision_policy, compute_backend=compute_backend)
        if self.velocity_set is not None and self.velocity_set.lattice_id == _lib.D3Q19:
            raise NotImplementedError("Velocity set not supported: {}".format(type(self.velocity_set)))

    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, f, feq, fout, rho, u, omega):
        return _collide(self, _lib.KBC, f, feq, fout, rho, u, omega)


class SmagorinskyLESBGK(Collision):
    """BGK with a Smagorinsky eddy relaxation time (smagorinsky_les_bgk.py:14-94):
    tau = tau0 + (sqrt(tau0^2 + 36 Cs^2 sqrt(strain)) - tau0) / 2 with the reference's "strain"
    (fneq^2 over the directions whose component sum is 1, 2 fneq^2 where it is >= 2)."""

    collision_id = _lib.SMAGORINSKY_BGK

    def __init__(self, velocity_set=None, precision_policy=None, compute_backend=None, smagorinsky_coef: float = 0.17):
        self.smagorinsky_coef = smagorinsky_coef
        super().__init__(velocity_set=velocity_set, precision_policy=precision_policy, compute_backend=compute_backend)

    @property
    def collision_param(self):
        return float(self.smagorinsky_coef)

    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, f, feq, fout, rho, u, omega):
        return _collide(self, _lib.SMAGORINSKY_BGK, f, feq, fout, None, None, omega,
                        collision_param=self.smagorinsky_coef)


class ForcedCollision(Collision):
    """Collision followed by exact-difference forcing (forced_collision.py:13-104).

    ``force_vector``: uniform body force (d,).  The MI355X backend also accepts a per-cell force
    field (d, nx, ny, nz) through ``force_field`` (the reference is limited to a uniform vector,
    exact_difference_force.py:34).
    """

    def __init__(self, collision_operator, forcing_scheme="exact_difference", force_vector=None, force_field=None):
        assert collision_operator is not None
        self.collision_operator = collision_operator
        super().__init__(collision_operator.velocity_set, collision_operator.precision_policy, collision_operator.compute_backend)
        assert forcing_scheme == "exact_difference", NotImplementedError(f"Force model {forcing_scheme} not implemented!")
        if force_vector is None and force_field is None:
            raise ValueError("ForcedCollision needs force_vector or force_field")
        if force_vector is not None:
            force_vector = np.asarray(force_vector, dtype=np.float64)
            assert force_vector.shape[0] == self.velocity_set.d, "Check the dimensions of the input force!"
        self.force_vector = force_vector
        self.force_field = force_field
        self.collision_id = collision_operator.collision_id
        self.collision_param = getattr(collision_operator, "collision_param", 0.0)

    @property
    def force_mode(self):
        return _lib.FORCE_FIELD if self.force_field is not None else _lib.FORCE_UNIFORM

    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, f, feq, fout, rho, u, omega):
        return _collide(
            self, self.collision_id, f, feq, fout, rho, u, omega, self.force_mode, self.force_vector, self.force_field,
            self.collision_param,
        )
