"""Boundary conditions (mirrors xlb/operator/boundary_condition/: boundary_condition.py:22-175,
boundary_condition_registry.py:6-30, and the bc_*.py classes).

Each class keeps the reference's constructor signature, implementation step and flags
(needs_padding, needs_aux_init, needs_aux_recovery, num_of_aux_data).  Its arithmetic lives
in the HIP kernels, selected by ``kind``; the stepper fuses it into the step, and calling a BC
directly runs the standalone kernel (BoundaryCondition._construct_kernel semantics).

Auxiliary data (the prescribed normal speed or density of Zou-He / regularized BCs) is held in
a per-cell side table ``aux`` in the compute precision, filled once by ``aux_data_init`` from
``profile`` -- a Python callable evaluated on the host over the BC's cells.  The reference keeps
the value in f_1[0] and copies it back every step (boundary_condition.py:130-161,
nse_stepper.py:318-342); the value itself is the same (rounded to the store precision, as the
reference stores it, except for the shifted-storage policies, which keep it in FP32).
"""

from collections import Counter
from enum import Enum, auto

import numpy as np
import torch

from xlb_amd import _lib
from xlb_amd.compute_backend import ComputeBackend
from xlb_amd.operator.boundary_masker.packing import cell_coordinates, cells_with_id, packed_missing
from xlb_amd.operator.operator import Operator, expect_field, field_extents, shared_layout


class ImplementationStep(Enum):
    COLLISION = auto()
    STREAMING = auto()


class BoundaryConditionRegistry:
    """Process-global id registry (boundary_condition_registry.py:6-30): ids start at 1 in
    construction order; 0 is fluid and 255 is reserved for solid cells."""

    def __init__(self):
        self.id_to_bc = {}
        self.bc_to_id = {}
        self.next_id = 1

    def register_boundary_condition(self, boundary_condition):
        _id = self.next_id
        self.next_id += 1
        self.id_to_bc[_id] = boundary_condition
        self.bc_to_id[boundary_condition] = _id
        return _id

    def reset(self):
        """Forget every registration (ids restart at 1); for long-lived processes and tests."""
        self.__init__()


boundary_condition_registry = BoundaryConditionRegistry()


class BoundaryCondition(Operator):
    kind = _lib.BC_NONE

    def __init__(
        self,
        implementation_step: ImplementationStep,
        velocity_set=None,
        precision_policy=None,
        compute_backend=None,
        indices=None,
        mesh_vertices=None,
    ):
        self.id = boundary_condition_registry.register_boundary_condition(self.__class__.__name__ + "_" + str(hash(self)))
        super().__init__(velocity_set, precision_policy, compute_backend)
        self.indices = indices
        self.mesh_vertices = mesh_vertices
        self.implementation_step = implementation_step
        self.needs_padding = False
        self.needs_mesh_distance = False
        self.needs_aux_init = False
        self.is_initialized_with_aux_data = False
        self.num_of_aux_data = 0
        self.needs_aux_recovery = False
        self.aux = None

    # ---- description handed to the kernels ----
    def bc_desc(self):
        d = _lib.BcDesc()
        d.kind = self.kind
        d.rho = 0.0
        for i in range(3):
            d.u[i] = 0.0
        return d

    def update_bc_auxilary_data(self, f_pre, f_post, bc_mask, missing_mask):
        """Placeholder of the reference (boundary_condition.py:73-79); the post-collision update of
        ExtrapolationOutflowBC is fused into the step kernel."""
        return f_post

    # ---- auxiliary data ----
    def _aux_values(self, index):
        """Prescribed values at cells `index` ((d, n) int array) as float64, (n,)."""
        vals = np.asarray(self.profile(index), dtype=np.float64).reshape(-1)
        n = index.shape[1]
        if vals.size == 1:
            vals = np.full(n, vals[0])
        if vals.size != n:
            raise ValueError(f"profile returned {vals.size} values for {n} cells")
        return vals

    def aux_data_init(self, f_0, f_1, bc_mask, missing_mask, aux=None):
        """Fill the side table at the cells with bc_mask == id (boundary_condition.py:163-175).

        ``aux``: (N,) tensor in the compute dtype shared by all BCs of a stepper; allocated here
        when not given.  f_0 and f_1 are returned unchanged (the reference writes into f_1)."""
        nx, ny, nz = field_extents(bc_mask)
        if aux is None:
            aux = torch.zeros(nx * ny * nz, dtype=self.compute_dtype, device=bc_mask.device)
        cells = cells_with_id(bc_mask, self.id)
        if cells.numel() > 0:
            index = cell_coordinates(cells, (nx, ny, nz), self.velocity_set.d)
            vals = torch.from_numpy(self._aux_values(index))
            if not (self.precision_policy.shifted or self.precision_policy.block):
                vals = vals.to(self.store_dtype)  # the reference stores it in the store dtype
            vals = vals.to(self.compute_dtype).to(bc_mask.device)
            _lib.check(
                _lib.lib().xlbh_scatter_aux(self.compute_id, _lib.ptr(cells), _lib.ptr(vals), int(cells.numel()),
                                            _lib.ptr(aux), _lib.stream_of(bc_mask)),
                "aux_data_init",
            )
        self.aux = aux
        self.is_initialized_with_aux_data = True
        return f_0, f_1

    # ---- standalone application ----
    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, f_pre, f_post, bc_mask, missing_mask):
        _lib.require_cuda(f_pre, f_post, bc_mask, missing_mask)
        nx, ny, nz = field_extents(f_post)
        q, ext = self.velocity_set.q, (nx, ny, nz)
        for t, name in ((f_post, "f_post"), (f_pre, "f_pre")):
            if tuple(t.shape) != (q,) + ext or t.dtype != f_post.dtype:
                raise ValueError(f"{name} must be a {(q,) + ext} field of dtype {f_post.dtype}")
        plane, (pre_, post_) = shared_layout([f_pre, f_post])
        expect_field(bc_mask, 1, ext, "bc_mask", torch.uint8)
        expect_field(missing_mask, q, ext, "missing_mask", torch.bool)
        bits = packed_missing(missing_mask, self.lattice_id)
        aux = None
        if self.needs_aux_init:
            if self.aux is None or self.aux.numel() != nx * ny * nz:
                self.aux_data_init(None, None, bc_mask, missing_mask)
            aux = self.aux
        desc = self.bc_desc()
        _lib.check(
            _lib.lib().xlbh_apply_bc(
                self.lattice_id, self.format_of(f_post), self.compute_id, self.kind, int(self.id), desc,
                _lib.ptr(pre_), _lib.ptr(post_), _lib.ptr(bc_mask), _lib.ptr(bits), _lib.ptr(aux), nx, ny, nz, plane,
                _lib.stream_of(f_post),
            ),
            self.__class__.__name__,
        )
        if post_ is not f_post:
            f_post.copy_(post_)
        _lib.mark_written(f_post)
        return f_post


class EquilibriumBC(BoundaryCondition):
    """f = feq(rho, u) at the boundary cells (bc_equilibrium.py:24-101; streaming step)."""

    kind = _lib.BC_EQUILIBRIUM

    def __init__(self, rho, u, equilibrium_operator=None, velocity_set=None, precision_policy=None,
                 compute_backend=None, indices=None, mesh_vertices=None):
        from xlb_amd.operator.equilibrium import Equilibrium, QuadraticEquilibrium

        self.rho = rho
        self.u = u
        self.equilibrium_operator = QuadraticEquilibrium() if equilibrium_operator is None else equilibrium_operator
        if not issubclass(type(self.equilibrium_operator), Equilibrium):
            raise ValueError("Equilibrium operator must be a subclass of Equilibrium")
        super().__init__(ImplementationStep.STREAMING, velocity_set, precision_policy, compute_backend, indices, mesh_vertices)

    def bc_desc(self):
        d = super().bc_desc()
        d.rho = float(self.rho)
        u = list(self.u) + [0.0] * (3 - len(self.u))
        for i in range(3):
            d.u[i] = float(u[i])
        return d


class HalfwayBounceBackBC(BoundaryCondition):
    """Missing populations take the cell's own pre-stream opposite population
    (bc_halfway_bounce_back.py:22-99; streaming step, needs padding)."""

    kind = _lib.BC_HALFWAY_BB

    def __init__(self, velocity_set=None, precision_policy=None, compute_backend=None, indices=None, mesh_vertices=None):
        super().__init__(ImplementationStep.STREAMING, velocity_set, precision_policy, compute_backend, indices, mesh_vertices)
        self.needs_padding = True


class FullwayBounceBackBC(BoundaryCondition):
    """Post-collision populations = flipped post-stream populations
    (bc_fullway_bounce_back.py:22-86; collision step)."""

    kind = _lib.BC_FULLWAY_BB

    def __init__(self, velocity_set=None, precision_policy=None, compute_backend=None, indices=None, mesh_vertices=None):
        super().__init__(ImplementationStep.COLLISION, velocity_set, precision_policy, compute_backend, indices, mesh_vertices)


class DoNothingBC(BoundaryCondition):
    """Skip streaming at the boundary cells (bc_do_nothing.py:21-76; streaming step)."""

    kind = _lib.BC_DO_NOTHING

    def __init__(self, velocity_set=None, precision_policy=None, compute_backend=None, indices=None, mesh_vertices=None):
        super().__init__(ImplementationStep.STREAMING, velocity_set, precision_policy, compute_backend, indices, mesh_vertices)


class ZouHeBC(BoundaryCondition):
    """Zou-He velocity / pressure BC with non-equilibrium bounce-back (bc_zouhe.py:25-362).

    Warp semantics: the normal is -c of the first missing axis direction; a velocity BC
    prescribes the normal speed only (u = -v n).  ``profile``: callable ``profile(index)`` with
    ``index`` an int array (d, n) of cell coordinates, returning n values (normal speed or
    density); evaluated once by ``aux_data_init``.  ``prescribed_value``: constant (velocity: a
    vector with at most one non-zero entry, whose signed value is used; pressure: a float).
    """

    kind_velocity = _lib.BC_ZOUHE_VELOCITY
    kind_pressure = _lib.BC_ZOUHE_PRESSURE

    def __init__(self, bc_type, profile=None, prescribed_value=None, velocity_set=None, precision_policy=None,
                 compute_backend=None, indices=None, mesh_vertices=None):
        assert bc_type in ["velocity", "pressure"], f"type = {bc_type} not supported! Use 'pressure' or 'velocity'."
        from xlb_amd.operator.equilibrium import QuadraticEquilibrium

        self.bc_type = bc_type
        self.equilibrium_operator = QuadraticEquilibrium()
        self.profile = profile
        super().__init__(ImplementationStep.STREAMING, velocity_set, precision_policy, compute_backend, indices, mesh_vertices)
        self.kind = self.kind_velocity if bc_type == "velocity" else self.kind_pressure

        if prescribed_value is not None:
            if profile is not None:
                raise ValueError("Cannot specify both profile and prescribed_value")
            if isinstance(prescribed_value, (tuple, list)):
                prescribed_value = np.array(prescribed_value, dtype=np.float64)
            elif isinstance(prescribed_value, (int, float)):
                if bc_type == "pressure":
                    prescribed_value = float(prescribed_value)
                else:
                    raise ValueError("Velocity prescribed_value must be a tuple or array")
            elif isinstance(prescribed_value, np.ndarray):
                prescribed_value = prescribed_value.astype(np.float64)
            if bc_type == "velocity":
                if not isinstance(prescribed_value, np.ndarray):
                    raise ValueError("Velocity prescribed_value must be an array-like")
                if np.count_nonzero(prescribed_value) > 1:
                    raise ValueError("This BC only supports normal prescribed values (only one non-zero element allowed)")
            # the single signed non-zero entry (or 0 for a no-slip vector), as in bc_zouhe.py:103-108
            arr = np.atleast_1d(np.asarray(prescribed_value, dtype=np.float64))
            nz_idx = np.nonzero(arr)[0]
            self.prescribed_value = float(arr[nz_idx][0]) if nz_idx.size else 0.0
            self.profile = self._create_constant_prescribed_profile()
        if self.profile is None:
            raise ValueError("ZouHeBC needs a profile or a prescribed_value")

        self.needs_aux_init = True
        self.needs_aux_recovery = True
        self.num_of_aux_data = 1
        self.needs_padding = True

    def _create_constant_prescribed_profile(self):
        value = self.prescribed_value

        def prescribed_profile(index):
            return np.full(np.asarray(index).shape[-1], value, dtype=np.float64)

        return prescribed_profile


class RegularizedBC(ZouHeBC):
    """Zou-He followed by regularisation of the non-equilibrium part (bc_regularized.py:21-220;
    Latt et al. 2008, eq. 45)."""

    kind_velocity = _lib.BC_REGULARIZED_VELOCITY
    kind_pressure = _lib.BC_REGULARIZED_PRESSURE


class ExtrapolationOutflowBC(BoundaryCondition):
    """Outflow BC: missing populations bounce back during streaming, and after collision each
    outgoing slot stores (1 - c_s) f_ps[l] + c_s f_neighbour[l] for use next step
    (bc_extrapolation_outflow.py:24-209).  Warp semantics: the normal is derived per cell
    from the missing mask."""

    kind = _lib.BC_EXTRAPOLATION_OUTFLOW

    def __init__(self, velocity_set=None, precision_policy=None, compute_backend=None, indices=None, mesh_vertices=None):
        super().__init__(ImplementationStep.STREAMING, velocity_set, precision_policy, compute_backend, indices, mesh_vertices)
        if indices is not None:
            self._get_normal_vec(indices)

    def _get_normal_vec(self, indices):
        # Most common coordinate per axis (bc_extrapolation_outflow.py:63-77); kept as an attribute.
        freq_counts = [Counter(coord).most_common(1)[0] for coord in indices]
        counts = np.array([count for _, count in freq_counts])
        elements = np.array([element for element, _ in freq_counts])
        self.normal = counts // counts.max()
        if elements[np.argmax(counts)] == 0:
            self.normal *= -1


class GradsApproximationBC(BoundaryCondition):
    """Grad's-approximation wall (bc_grads_approximation.py:28-321; streaming step, 3-D): the
    missing populations take the interpolated bounce-back with the wall weight 1/2 the reference
    uses for index-defined walls (:195-201), then every population is regularised around the
    cell's own density and velocity (:203-208, :112-135).  Index-defined walls need padding; a mesh
    definition needs a mesh-distance masker, which the reference does not provide either."""

    kind = _lib.BC_GRADS_APPROXIMATION

    def __init__(self, velocity_set=None, precision_policy=None, compute_backend=None, indices=None, mesh_vertices=None):
        self.u = (0, 0, 0)
        super().__init__(ImplementationStep.STREAMING, velocity_set, precision_policy, compute_backend, indices, mesh_vertices)
        self.needs_mesh_distance = True
        if self.mesh_vertices is None:
            assert self.indices is not None
            self.needs_padding = True
        if self.velocity_set.d == 2:
            raise NotImplementedError("This BC is not implemented in 2D!")
