"""IncompressibleNavierStokesStepper (mirrors xlb/operator/stepper/nse_stepper.py:25-392 and
stepper.py:6-34).

``__call__(f_0, f_1, bc_mask, missing_mask, omega, timestep)`` runs ONE fused HIP kernel
(libxlbhip ``xlbh_step``): pull streaming, streaming-step BCs, macroscopic, equilibrium,
BGK/KBC/Smagorinsky collision (+ exact-difference force), collision-step BCs, extrapolation-outflow update
and store, writing f_1 and returning (f_0, f_1); the caller swaps, as with the reference.

Per-mask device state (derived once and re-derived only when the mask tensors change):
packed missing bits (N uint32), the 256-entry BC table indexed by boundary id, and the aux side
table of prescribed values.  ``run`` moves the caller's step/swap loop below the C ABI
(``xlbh_run``).
"""

import ctypes

import numpy as np
import torch

from xlb_amd import _lib
from xlb_amd.compute_backend import ComputeBackend
from xlb_amd.default_config import DefaultConfig
from xlb_amd.helper.check_boundary_overlaps import check_bc_overlaps
from xlb_amd.grid.grid import copy_population
from xlb_amd.helper.nse_solver import create_nse_fields
from xlb_amd.operator.boundary_condition.boundary_condition import ImplementationStep
from xlb_amd.operator.boundary_masker import IndicesBoundaryMasker, MeshBoundaryMasker
from xlb_amd.operator.boundary_masker.packing import cells_with_id, packed_missing
from xlb_amd.operator.collision import BGK, KBC, ForcedCollision, SmagorinskyLESBGK
from xlb_amd.operator.equilibrium import QuadraticEquilibrium
from xlb_amd.operator.macroscopic import Macroscopic
from xlb_amd.operator.operator import Operator, common_plane, expect_field, expect_population, field_extents
from xlb_amd.operator.stream import Stream


class Stepper(Operator):
    """Base of steppers (stepper.py:6-34)."""

    def __init__(self, grid, boundary_conditions):
        self.grid = grid
        self.boundary_conditions = boundary_conditions
        super().__init__(DefaultConfig.velocity_set, DefaultConfig.default_precision_policy, DefaultConfig.default_backend)

    def prepare_fields(self, initializer=None):
        raise NotImplementedError("Subclasses must implement prepare_fields()")


class IncompressibleNavierStokesStepper(Stepper):
    def __init__(
        self,
        grid,
        boundary_conditions=[],
        collision_type="BGK",
        streaming_scheme="pull",
        forcing_scheme="exact_difference",
        force_vector=None,
        force_field=None,
    ):
        super().__init__(grid, list(boundary_conditions))
        if collision_type == "BGK":
            self.collision = BGK(self.velocity_set, self.precision_policy, self.compute_backend)
        elif collision_type == "KBC":
            self.collision = KBC(self.velocity_set, self.precision_policy, self.compute_backend)
        elif collision_type == "SmagorinskyLESBGK":
            self.collision = SmagorinskyLESBGK(self.velocity_set, self.precision_policy, self.compute_backend)
        else:
            raise ValueError(f"Unknown collision type {collision_type}")
        self.collision_type = collision_type
        if force_vector is not None or force_field is not None:
            self.collision = ForcedCollision(
                collision_operator=self.collision, forcing_scheme=forcing_scheme, force_vector=force_vector, force_field=force_field
            )
        self.streaming_scheme = streaming_scheme
        assert streaming_scheme == "pull", f"Unknown or unimplemented streaming scheme for backend: {self.compute_backend}"
        self.stream = Stream(self.velocity_set, self.precision_policy, self.compute_backend)
        self.equilibrium = QuadraticEquilibrium(self.velocity_set, self.precision_policy, self.compute_backend)
        self.macroscopic = Macroscopic(self.velocity_set, self.precision_policy, self.compute_backend)
        self._state_key = None
        self._bits = None
        self._table = None
        self._aux = None
        self._has_bc_cells = True
        self._has_solid = False
        self._bc_cells = None
        self._mask_ref = None

    # ------------------------------------------------------------------ fields
    def prepare_fields(self, initializer=None):
        """Allocate and initialise (f_0, f_1, bc_mask, missing_mask) (nse_stepper.py:58-98)."""
        _, f_0, f_1, missing_mask, bc_mask = create_nse_fields(
            grid=self.grid, velocity_set=self.velocity_set, compute_backend=self.compute_backend, precision_policy=self.precision_policy
        )
        if initializer is not None:
            f_0 = initializer(self.grid, self.velocity_set, self.precision_policy, self.compute_backend)
        else:
            from xlb_amd.helper.initializers import initialize_eq

            f_0 = initialize_eq(f_0, self.grid, self.velocity_set, self.precision_policy, self.compute_backend)
        if f_1.shape == f_0.shape and f_1.stride() == f_0.stride():
            copy_population(f_1, f_0)
        else:
            f_1.copy_(f_0)
        bc_mask, missing_mask = self._process_boundary_conditions(self.boundary_conditions, bc_mask, missing_mask)
        f_0, f_1 = self._initialize_auxiliary_data(self.boundary_conditions, f_0, f_1, bc_mask, missing_mask)
        self._sync_state(bc_mask, missing_mask)
        return f_0, f_1, bc_mask, missing_mask

    @classmethod
    def _process_boundary_conditions(cls, boundary_conditions, bc_mask, missing_mask):
        """Overlap check, indices masker, then the mesh masker for 3-D mesh BCs (nse_stepper.py:100-127)."""
        check_bc_overlaps(boundary_conditions, DefaultConfig.velocity_set.d, DefaultConfig.default_backend)
        indices_masker = IndicesBoundaryMasker(
            velocity_set=DefaultConfig.velocity_set,
            precision_policy=DefaultConfig.default_precision_policy,
            compute_backend=DefaultConfig.default_backend,
        )
        bc_with_vertices = [bc for bc in boundary_conditions if bc.mesh_vertices is not None]
        bc_with_indices = [bc for bc in boundary_conditions if bc.indices is not None]
        if bc_with_indices:
            bc_mask, missing_mask = indices_masker(bc_with_indices, bc_mask, missing_mask)
        if DefaultConfig.velocity_set.d == 3 and bc_with_vertices:
            mesh_masker = MeshBoundaryMasker(
                velocity_set=DefaultConfig.velocity_set,
                precision_policy=DefaultConfig.default_precision_policy,
                compute_backend=DefaultConfig.default_backend,
            )
            for bc in bc_with_vertices:
                bc_mask, missing_mask = mesh_masker(bc, bc_mask, missing_mask)
        return bc_mask, missing_mask

    def _initialize_auxiliary_data(self, boundary_conditions, f_0, f_1, bc_mask, missing_mask):
        """Fill the shared aux side table for BCs that carry prescribed values (nse_stepper.py:129-135)."""
        aux_bcs = [bc for bc in boundary_conditions if bc.needs_aux_init]
        if not aux_bcs:
            self._aux = None
            return f_0, f_1
        nx, ny, nz = field_extents(bc_mask)
        self._aux = torch.zeros(nx * ny * nz, dtype=self.compute_dtype, device=bc_mask.device)
        for bc in aux_bcs:
            bc.aux_data_init(f_0, f_1, bc_mask, missing_mask, aux=self._aux)
        return f_0, f_1

    # ------------------------------------------------------------------ device state
    def _mask_key(self, bc_mask, missing_mask):
        return (id(bc_mask), bc_mask._version, bc_mask.data_ptr(), id(missing_mask), missing_mask._version, missing_mask.data_ptr())

    def _build_table(self, device):
        descs = (_lib.BcDesc * 256)()
        for bc in self.boundary_conditions:
            if not (1 <= int(bc.id) <= 254):
                raise ValueError(f"boundary id {bc.id} does not fit the uint8 bc_mask (1..254)")
            descs[int(bc.id)] = bc.bc_desc()
        raw = np.frombuffer(bytes(descs), dtype=np.uint8).copy()
        return torch.from_numpy(raw).to(device)

    def _sync_state(self, bc_mask, missing_mask):
        key = self._mask_key(bc_mask, missing_mask)
        if key == self._state_key:
            return
        _lib.require_cuda(bc_mask, missing_mask)
        ext = field_extents(bc_mask)
        expect_field(bc_mask, 1, ext, "bc_mask", torch.uint8)
        expect_field(missing_mask, self.velocity_set.q, ext, "missing_mask", torch.bool)
        self._bits = packed_missing(missing_mask, self.lattice_id)
        self._table = self._build_table(bc_mask.device)
        self._has_bc_cells = bool(bc_mask.any().item())
        self._has_solid = self._has_bc_cells and bool((bc_mask == 255).any().item())
        self._bc_cells = None
        nx, ny, nz = field_extents(bc_mask)
        aux_bcs = [bc for bc in self.boundary_conditions if bc.needs_aux_init]
        if aux_bcs and (self._aux is None or self._aux.numel() != nx * ny * nz or self._state_key is not None):
            # masks changed after prepare_fields (or were built by the caller): re-derive aux
            self._aux = torch.zeros(nx * ny * nz, dtype=self.compute_dtype, device=bc_mask.device)
            for bc in aux_bcs:
                bc.aux_data_init(None, None, bc_mask, missing_mask, aux=self._aux)
        self._mask_ref = bc_mask
        self._state_key = self._mask_key(bc_mask, missing_mask)

    def _force_args(self):
        col = self.collision
        if isinstance(col, ForcedCollision):
            if col.force_field is not None:
                ff = col.force_field
                if ff.dtype != self.compute_dtype:
                    ff = ff.to(self.compute_dtype)
                    col.force_field = ff
                return _lib.FORCE_FIELD, None, ff
            return _lib.FORCE_UNIFORM, col.force_vector, None
        return _lib.FORCE_NONE, None, None

    def _collision_id(self):
        return {"KBC": _lib.KBC, "SmagorinskyLESBGK": _lib.SMAGORINSKY_BGK}.get(self.collision_type, _lib.BGK)

    def _collision_param(self):
        return float(getattr(self.collision, "collision_param", 0.0))

    def _launch_args(self, f_0, f_1, bc_mask, missing_mask):
        nx, ny, nz = field_extents(f_0)
        q, ext = self.velocity_set.q, (nx, ny, nz)
        fmt = self.precision_policy.store_format
        planes = [expect_population(f_0, q, ext, "f_0", self.store_dtype, self.lattice_id, fmt)]
        if f_1 is not None:
            planes.append(expect_population(f_1, q, ext, "f_1", self.store_dtype, self.lattice_id, fmt))
            if f_0.data_ptr() == f_1.data_ptr():
                raise ValueError("f_0 and f_1 must be distinct buffers")
        plane = common_plane(planes, "IncompressibleNavierStokesStepper")
        expect_field(bc_mask, 1, ext, "bc_mask", torch.uint8)
        self._sync_state(bc_mask, missing_mask)
        force_mode, force_vec, force_field = self._force_args()
        if force_field is not None:
            expect_field(force_field, self.velocity_set.d, ext, "force_field", self.compute_dtype)
        fu_ptr, keep = _lib.doubles(force_vec if force_vec is not None else (0.0, 0.0, 0.0))
        if self._has_bc_cells:
            masks = (_lib.ptr(bc_mask), _lib.ptr(self._bits), _lib.ptr(self._aux), _lib.ptr(self._table))
        else:
            masks = (None, None, None, None)
        return nx, ny, nz, plane, masks, force_mode, fu_ptr, keep, force_field

    # ------------------------------------------------------------------ stepping
    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, f_0, f_1, bc_mask, missing_mask, omega, timestep):
        nx, ny, nz, plane, masks, force_mode, fu_ptr, keep, force_field = self._launch_args(f_0, f_1, bc_mask, missing_mask)
        _lib.check(
            _lib.lib().xlbh_step(
                self.lattice_id, self._collision_id(), self.precision_policy.store_format, self.compute_id,
                _lib.ptr(f_0), _lib.ptr(f_1), *masks, force_mode, fu_ptr, _lib.ptr(force_field),
                nx, ny, nz, plane, float(omega), self._collision_param(), int(timestep), _lib.stream_of(f_0),
            ),
            "IncompressibleNavierStokesStepper",
        )
        _lib.mark_written(f_1)
        return f_0, f_1

    def step_slab(self, f_0, f_1, bc_mask, missing_mask, omega, timestep, x_begin, x_end):
        """One fused step of an x-slab (``xlbh_step_slab``): only planes x_begin <= x < x_end of
        f_1 are written, with 1 <= x_begin < x_end <= nx - 1.  Planes x_begin - 1 and x_end of
        f_0 are read as ghost planes; they must hold the adjacent slabs' populations (those with
        c_x = +1 in the lower ghost plane and c_x = -1 in the upper one are the ones read).  The
        fields, masks and boundary conditions are the slab's own, ghost planes included.  This is
        the per-shard step of an x-sharded run (the reference's distribute.py:18-47 exchanges the
        crossing populations after each shard's step; the pull scheme needs them before it)."""
        nx, ny, nz, plane, masks, force_mode, fu_ptr, keep, force_field = self._launch_args(f_0, f_1, bc_mask, missing_mask)
        _lib.check(
            _lib.lib().xlbh_step_slab(
                self.lattice_id, self._collision_id(), self.precision_policy.store_format, self.compute_id,
                _lib.ptr(f_0), _lib.ptr(f_1), *masks, force_mode, fu_ptr, _lib.ptr(force_field),
                nx, ny, nz, plane, float(omega), self._collision_param(), int(timestep), int(x_begin), int(x_end),
                _lib.stream_of(f_0),
            ),
            "IncompressibleNavierStokesStepper.step_slab",
        )
        _lib.mark_written(f_1)
        return f_0, f_1

    _ADJOINT_BC_KINDS = (
        _lib.BC_HALFWAY_BB, _lib.BC_EXTRAPOLATION_OUTFLOW, _lib.BC_DO_NOTHING, _lib.BC_ZOUHE_VELOCITY,
        _lib.BC_ZOUHE_PRESSURE, _lib.BC_REGULARIZED_VELOCITY, _lib.BC_REGULARIZED_PRESSURE,
        _lib.BC_GRADS_APPROXIMATION,
    )

    def _check_adjoint_supported(self):
        if self.collision_type not in ("BGK", "KBC", "SmagorinskyLESBGK"):
            raise NotImplementedError(f"no adjoint for collision {self.collision_type}")

    def _adjoint_cells(self):
        """(pointer, count) of the cells the adjoint's boundary pass runs over: every cell whose BC
        kind has contributions outside the pull (bounce-back, extrapolation, do-nothing, Zou-He /
        regularized, Grad's approximation); fullway bounce-back and equilibrium cells have none."""
        if not self._has_bc_cells:
            return None, 0
        if self._bc_cells is None:
            ids = [int(bc.id) for bc in self.boundary_conditions if bc.kind in self._ADJOINT_BC_KINDS]
            parts = [cells_with_id(self._mask_ref, i) for i in ids]
            parts = [p for p in parts if p.numel() > 0]
            self._bc_cells = torch.cat(parts) if parts else torch.zeros(0, dtype=torch.int32, device=self._mask_ref.device)
        if self._bc_cells.numel() == 0:
            return None, 0
        return _lib.ptr(self._bc_cells), int(self._bc_cells.numel())

    def adjoint_step(self, f_t, lam_out, lam_in, bc_mask, missing_mask, omega, grad_force=None):
        """Reverse mode of one step (xlbh_adjoint_step): given the state f_t a forward step read and
        lam_out = dL/df_{t+1}, write lam_in = dL/df_t and add dL/dg to ``grad_force`` ((d, nx, ny, nz)
        float64) when the stepper has a body force.  Every collision (BGK, KBC, SmagorinskyLESBGK)."""
        self._check_adjoint_supported()
        nx, ny, nz, plane, masks, force_mode, fu_ptr, keep, force_field = self._launch_args(f_t, None, bc_mask, missing_mask)
        q, ext = self.velocity_set.q, (nx, ny, nz)
        expect_field(lam_out, q, ext, "lam_out", self.compute_dtype)
        expect_field(lam_in, q, ext, "lam_in", self.compute_dtype)
        if lam_in.data_ptr() == lam_out.data_ptr():
            raise ValueError("lam_in and lam_out must be distinct buffers")
        if grad_force is not None:
            expect_field(grad_force, self.velocity_set.d, ext, "grad_force", torch.float64)
        if self._has_solid:
            raise NotImplementedError("the adjoint step does not support solid (255) cells")
        _lib.check(
            _lib.lib().xlbh_adjoint_step(
                self.lattice_id, self._collision_id(), self.precision_policy.store_format, self.compute_id,
                _lib.ptr(f_t), _lib.ptr(lam_out), _lib.ptr(lam_in), *masks, *self._adjoint_cells(), force_mode, fu_ptr,
                _lib.ptr(force_field),
                _lib.ptr(grad_force), nx, ny, nz, plane, float(omega), self._collision_param(), _lib.stream_of(f_t),
            ),
            "IncompressibleNavierStokesStepper.adjoint_step",
        )
        _lib.mark_written(lam_in, grad_force)
        return lam_in

    def record(self, f_start, states, bc_mask, missing_mask, omega, num_states, timestep=0):
        """``num_states`` steps from ``f_start`` keeping every state (xlbh_record): states[j] is the
        output of step ``timestep + j``.  ``states`` is a contiguous (>= num_states, q, ...) tensor
        in the store dtype."""
        nx, ny, nz, plane, masks, force_mode, fu_ptr, keep, force_field = self._launch_args(f_start, None, bc_mask, missing_mask)
        if num_states < 0:
            raise ValueError("num_states must be >= 0")
        if num_states > 0:
            self._expect_states(states, num_states, f_start, plane)
        _lib.check(
            _lib.lib().xlbh_record(
                self.lattice_id, self._collision_id(), self.precision_policy.store_format, self.compute_id,
                _lib.ptr(f_start), _lib.ptr(states), *masks, force_mode, fu_ptr, _lib.ptr(force_field),
                nx, ny, nz, plane, float(omega), self._collision_param(), int(timestep), int(num_states),
                _lib.stream_of(f_start),
            ),
            "IncompressibleNavierStokesStepper.record",
        )
        _lib.mark_written(states)
        return states

    def state_buffer(self, num_states, f_like):
        """Storage for ``num_states`` population fields laid out like ``f_like`` (its plane stride),
        one after another, as ``record`` and ``adjoint_run`` expect: a (num_states, q, nx, ny, nz)
        view with strides (q * plane, plane, ny * nz, nz, 1)."""
        from xlb_amd.operator.operator import plane_stride

        q, nx, ny, nz = (int(v) for v in f_like.shape)
        plane = plane_stride(f_like)
        if plane is None:
            raise ValueError("f_like must be C-ordered within each plane")
        per = int(_lib.lib().xlbh_population_bytes(self.lattice_id, self.precision_policy.store_format, nx, ny, nz,
                                                   plane))
        elem = f_like.element_size()
        if per < 0 or per % elem or (per // elem) % q:
            raise ValueError("state buffers are not available for this store format")
        buf = torch.empty(max(1, num_states) * per // elem, dtype=f_like.dtype, device=f_like.device)
        return buf.as_strided((max(1, num_states), q, nx, ny, nz), (per // elem, plane, ny * nz, nz, 1))

    def _expect_states(self, states, n, f_like, plane):
        q = self.velocity_set.q
        nx, ny, nz = (int(v) for v in f_like.shape[1:])
        per = int(_lib.lib().xlbh_population_bytes(self.lattice_id, self.precision_policy.store_format, nx, ny, nz,
                                                   plane))
        elem = f_like.element_size()
        want = (per // elem, plane, ny * nz, nz, 1)
        ok = (states is not None and states.dtype == self.store_dtype and states.dim() == 5 and states.shape[0] >= n
              and tuple(states.shape[1:]) == (q, nx, ny, nz) and per % elem == 0
              and all(states.shape[d] == 1 or states.stride(d) == want[d] for d in range(5))
              and states.untyped_storage().nbytes() - states.storage_offset() * elem >= n * per)
        if not ok:
            raise ValueError(f"states must hold {n} population fields laid out like the state they follow: a "
                             f"(n, q, nx, ny, nz) store-dtype tensor with strides {want} (see state_buffer)")

    def adjoint_workspace(self, bc_mask):
        """A workspace for the fused reverse step (see ``adjoint_run``)."""
        nx, ny, nz = field_extents(bc_mask)
        nbytes = int(_lib.lib().xlbh_adjoint_workspace_bytes(self.lattice_id, self.compute_id, nx, ny, nz))
        return torch.empty(nbytes, dtype=torch.uint8, device=bc_mask.device)

    def adjoint_run(self, f_ckpt, states, lam_a, lam_b, bc_mask, missing_mask, omega, num_steps, timestep=0,
                    grad_force=None, recompute=True, workspace=None, begin=True, end=True):
        """Reverse sweep over one checkpoint segment below the ABI (xlbh_adjoint_run): recompute the
        ``num_steps - 1`` states after ``f_ckpt`` into ``states`` ((num_steps - 1, q, ...) in the
        store dtype), then apply the adjoint steps in reverse starting from lam_a = dL/df at the
        segment end.  Returns the buffer (lam_a or lam_b) that holds dL/df_ckpt.  With
        ``recompute=False`` the states are taken as recorded by ``record``.  A ``workspace`` (from
        ``adjoint_workspace``) selects the fused one-kernel reverse step; a sweep over several
        segments shares it, with ``begin`` on its first call and ``end`` on its last.  Every
        collision (BGK, KBC, SmagorinskyLESBGK)."""
        self._check_adjoint_supported()
        if self._has_solid:
            raise NotImplementedError("the adjoint step does not support solid (255) cells")
        nx, ny, nz, plane, masks, force_mode, fu_ptr, keep, force_field = self._launch_args(f_ckpt, None, bc_mask, missing_mask)
        q, ext = self.velocity_set.q, (nx, ny, nz)
        expect_field(lam_a, q, ext, "lam_a", self.compute_dtype)
        expect_field(lam_b, q, ext, "lam_b", self.compute_dtype)
        if lam_a.data_ptr() == lam_b.data_ptr():
            raise ValueError("lam_a and lam_b must be distinct buffers")
        if num_steps < 1:
            raise ValueError("num_steps must be >= 1")
        if num_steps > 1:
            self._expect_states(states, num_steps - 1, f_ckpt, plane)
        if grad_force is not None:
            expect_field(grad_force, self.velocity_set.d, ext, "grad_force", torch.float64)
        flags = (_lib.ADJ_RECOMPUTE if recompute else 0)
        if workspace is not None:
            need = int(_lib.lib().xlbh_adjoint_workspace_bytes(self.lattice_id, self.compute_id, nx, ny, nz))
            if workspace.dtype != torch.uint8 or workspace.numel() < need or not workspace.is_contiguous():
                raise ValueError(f"workspace must be a contiguous uint8 tensor of >= {need} bytes")
            flags |= (_lib.ADJ_BEGIN if begin else 0) | (_lib.ADJ_END if end else 0)
        _lib.check(
            _lib.lib().xlbh_adjoint_run(
                self.lattice_id, self._collision_id(), self.precision_policy.store_format, self.compute_id,
                _lib.ptr(f_ckpt), _lib.ptr(states), _lib.ptr(lam_a), _lib.ptr(lam_b), *masks, *self._adjoint_cells(),
                force_mode, fu_ptr, _lib.ptr(force_field), _lib.ptr(grad_force), nx, ny, nz, plane, float(omega),
                self._collision_param(), int(timestep), int(num_steps), _lib.ptr(workspace), flags,
                _lib.stream_of(f_ckpt),
            ),
            "IncompressibleNavierStokesStepper.adjoint_run",
        )
        _lib.mark_written(states, lam_a, lam_b, grad_force, workspace)
        return lam_a if num_steps % 2 == 0 else lam_b

    def run(self, f_0, f_1, bc_mask, missing_mask, omega, num_steps, timestep=0):
        """``num_steps`` steps with the buffer swap in C; returns (f_0, f_1) with the newest
        populations in f_0, i.e. the state the reference's loop
        ``f_0, f_1 = stepper(...); f_0, f_1 = f_1, f_0`` leaves after ``num_steps`` passes."""
        nx, ny, nz, plane, masks, force_mode, fu_ptr, keep, force_field = self._launch_args(f_0, f_1, bc_mask, missing_mask)
        _lib.check(
            _lib.lib().xlbh_run(
                self.lattice_id, self._collision_id(), self.precision_policy.store_format, self.compute_id,
                _lib.ptr(f_0), _lib.ptr(f_1), *masks, force_mode, fu_ptr, _lib.ptr(force_field),
                nx, ny, nz, plane, float(omega), self._collision_param(), int(timestep), int(num_steps),
                _lib.stream_of(f_0),
            ),
            "IncompressibleNavierStokesStepper.run",
        )
        _lib.mark_written(f_0, f_1)
        if num_steps % 2 == 1:
            return f_1, f_0
        return f_0, f_1
