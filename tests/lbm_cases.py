"""Shared lattice-Boltzmann test cases: the same configuration built through the product API
(xlb_amd, HIP kernels) and through the CPU oracle, plus helpers to move population fields
between the two in their stored bit patterns.

The oracle side restates the reference's host packing of boundary indices
(indices_boundary_masker.py:160-208) instead of calling the product's, so mask construction
is checked end to end.
"""

import numpy as np

from oracle import oracle as orc

# ------------------------------------------------------------------ formats
STORE_NP = {
    orc.F32: np.float32,
    orc.F64: np.float64,
    orc.F16: np.uint16,
    orc.BF16: np.uint16,
    orc.F16_SHIFTED: np.uint16,
    orc.BF16_SHIFTED: np.uint16,
    orc.I16_BLOCK: np.int16,
}


# ---- block format (int16 mantissas + one int32 exponent per row segment after the planes) ----
def block_geometry(shape3):
    """(bz, segments per row, rows) of the block format for grid (nx, ny, nz)."""
    nx, ny, nz = shape3
    fast = nz if nz > 1 else ny
    bz = next((b for b in (64, 128, 256) if b >= fast), 256)
    return bz, -(-fast // bz), (nx * ny if nz > 1 else nx)


def block_layout(q, shape3):
    """(byte offset of the exponents, number of segments) of a dense block field."""
    n = int(np.prod(shape3))
    _, segs, rows = block_geometry(shape3)
    return -(-(q * n * 2) // 256) * 256, segs * rows


def pop_alloc(fmt, q, shape3):
    """Host population field in store format fmt; a block field is the (q, ...) view at the start
    of a buffer that also holds its exponents."""
    if fmt != orc.I16_BLOCK:
        return np.zeros((q,) + tuple(shape3), dtype=STORE_NP[fmt])
    off, nseg = block_layout(q, shape3)
    buf = np.zeros((off + 4 * nseg) // 2, dtype=np.int16)
    return buf[: q * int(np.prod(shape3))].reshape((q,) + tuple(shape3))


def _block_buffer(a):
    b = a
    while b.base is not None and isinstance(b.base, np.ndarray):
        b = b.base
    return b


def block_exps(a):
    """The int32 segment exponents of a host block field (a view into its buffer)."""
    off, nseg = block_layout(a.shape[0], a.shape[1:])
    return _block_buffer(a)[off // 2: off // 2 + 2 * nseg].view(np.int32)


def pop_copy(a):
    if a.dtype != np.int16:
        return a.copy()
    return _block_buffer(a).copy()[: a.size].reshape(a.shape)


def policy_codes(policy):
    """(store format, compute id) of a policy for the oracle, which restates the reference's FP32
    arithmetic also for the deviation form (it has no other)."""
    comp = policy.compute_id
    return policy.store_format, (0 if comp == 2 else comp)


def to_np(t):
    """GPU tensor -> numpy array holding the stored bit pattern (a block field with its exponents)."""
    import torch

    if t.dtype == torch.int16:
        import xlb_amd as xa

        a = pop_alloc(orc.I16_BLOCK, t.shape[0], tuple(t.shape[1:]))
        a[...] = t.detach().cpu().numpy()
        block_exps(a)[...] = xa.grid.block_exponents(t).cpu().numpy()
        return a
    if t.dtype in (torch.float16, torch.bfloat16):
        return t.detach().view(torch.int16).cpu().numpy().view(np.uint16)
    if t.dtype == torch.bool:
        return t.detach().cpu().numpy().astype(np.uint8)
    return t.detach().cpu().numpy()


def from_np(a, like):
    """numpy stored bit pattern -> tensor with dtype/device of `like`."""
    import torch

    if like.dtype in (torch.float16, torch.bfloat16):
        return torch.from_numpy(a.view(np.int16).copy()).to(like.device).view(like.dtype)
    return torch.from_numpy(np.ascontiguousarray(a)).to(like.device)


def upload(a, t):
    """Host stored populations (with a block field's exponents) -> the GPU field t, in place."""
    import torch

    if t.dtype == torch.int16:
        import xlb_amd as xa

        t.copy_(torch.from_numpy(np.ascontiguousarray(a)).to(t.device))
        xa.grid.block_exponents(t).copy_(torch.from_numpy(block_exps(a).copy()).to(t.device))
        return t
    return t.copy_(from_np(a, t))


def decode(a, fmt, w):
    """Stored populations -> float64 values of f (block: w (1 + m 2^-k) with w rounded to float,
    the value the kernels decode to before their single rounding)."""
    if fmt == orc.I16_BLOCK:
        q, shape3 = a.shape[0], a.shape[1:]
        bz, segs, _ = block_geometry(shape3)
        k = block_exps(a).astype(np.float64)
        nx, ny, nz = shape3
        if nz > 1:
            k = np.repeat(k.reshape(nx, ny, segs), bz, axis=2)[:, :, :nz]
        else:
            k = np.repeat(k.reshape(nx, segs), bz, axis=1)[:, :ny, None]
        wf = np.asarray(w, dtype=np.float32).astype(np.float64).reshape((-1, 1, 1, 1))
        return wf * (1.0 + a.astype(np.float64) * np.exp2(-k)[None])
    if fmt in (orc.F32, orc.F64):
        v = a.astype(np.float64)
    elif fmt in (orc.F16, orc.F16_SHIFTED):
        v = a.view(np.float16).astype(np.float64)
    else:
        v = orc.bf16_bits_to_f32(a).astype(np.float64)
    if fmt in (orc.F16_SHIFTED, orc.BF16_SHIFTED):
        v = v + np.asarray(w, dtype=np.float64).reshape((-1, 1, 1, 1))
    return v


def encode(v, fmt, w):
    """float64 values of f -> stored populations (host rounding of the oracle)."""
    v = np.asarray(v, dtype=np.float64)
    if fmt in (orc.F16_SHIFTED, orc.BF16_SHIFTED):
        v = v - np.asarray(w, dtype=np.float64).reshape((-1, 1, 1, 1))
    if fmt == orc.F32:
        return v.astype(np.float32)
    if fmt == orc.F64:
        return v.copy()
    if fmt in (orc.F16, orc.F16_SHIFTED):
        return orc.f32_to_f16_bits(v.astype(np.float32))
    return orc.f32_to_bf16_bits(v.astype(np.float32))


def moments(fv, c):
    """rho, u (float64) of decoded populations fv (q, nx, ny, nz); c (d, q)."""
    rho = fv.sum(axis=0)
    u = np.tensordot(np.asarray(c, dtype=np.float64), fv, axes=(1, 0)) / rho
    return rho, u


# ------------------------------------------------------------------ oracle-side BC handling
def oracle_bc_desc(bc):
    d = dict(id=int(bc.id), kind=int(bc.kind), rho=0.0, u=(0.0, 0.0, 0.0))
    if hasattr(bc, "rho") and hasattr(bc, "u") and bc.kind == orc.EQUILIBRIUM:
        d["rho"] = float(bc.rho)
        d["u"] = tuple(float(x) for x in bc.u)
    return d


def oracle_pack(bc, shape3, d):
    """(3, n) int32 indices and (n,) interior flags, as indices_boundary_masker.py:160-208."""
    idx = np.asarray(bc.indices)
    if idx.ndim != 2:
        idx = idx.reshape(len(bc.indices), -1)
    n = idx.shape[1]
    if idx.shape[0] == 2:
        idx = np.vstack([idx, np.zeros(n, dtype=int)])
    if bc.needs_padding:
        shape_array = np.array(shape3)
        interior = np.all((idx[:d] > 0) & (idx[:d] < shape_array[:d, np.newaxis] - 1), axis=0)
    else:
        interior = np.zeros(n, dtype=bool)
    return idx.astype(np.int32), interior.astype(np.uint8)


def oracle_masks(lat, q, d, shape3, bcs):
    bc_mask = np.zeros((1,) + tuple(shape3), dtype=np.uint8)
    missing = np.zeros((q,) + tuple(shape3), dtype=np.uint8)
    for bc in bcs:
        if bc.indices is None:
            continue
        idx, inter = oracle_pack(bc, shape3, d)
        orc.mask_indices(lat, idx, inter, bc.id, bc_mask, missing)
    return bc_mask, missing


def oracle_aux(bc, bc_mask, d, policy):
    """Per-cell prescribed values (float64, (N,)), rounded like the product/reference store them."""
    import torch

    nx, ny, nz = bc_mask.shape[1:]
    cells = np.nonzero(bc_mask.reshape(-1) == bc.id)[0]
    vals = np.zeros(nx * ny * nz)
    if cells.size:
        idx = np.stack([cells // (ny * nz), (cells // nz) % ny, cells % nz])[:d]
        v = np.asarray(bc.profile(idx), dtype=np.float64).reshape(-1)
        if v.size == 1:
            v = np.full(cells.size, v[0])
        t = torch.from_numpy(v)
        if not (policy.shifted or policy.block):
            t = t.to(policy.store_precision.torch_dtype)
        vals[cells] = t.to(policy.compute_precision.torch_dtype).double().numpy()
    return vals


# ------------------------------------------------------------------ case definitions
def sphere_indices(shape, center, radius):
    grids = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    r2 = sum((g - c) ** 2 for g, c in zip(grids, center))
    ind = np.where(r2 < radius**2)
    return [tuple(int(v) for v in ind[i]) for i in range(len(shape))]


def unique_faces(box, names, d):
    walls = [sum((box[n][i] for n in names), []) for i in range(d)]
    return np.unique(np.array(walls), axis=-1).tolist()


def make_bcs(name, grid, xa):
    """Boundary conditions of case `name` built with the product API module `xa`."""
    BC = xa.operator.boundary_condition
    box = grid.bounding_box_indices()
    ne = grid.bounding_box_indices(remove_edges=True)
    shape = grid.shape
    if name in ("periodic", "periodic_kbc", "periodic_forced", "periodic_field"):
        return []
    if name == "cavity_fullway":  # mlups_3d.py:47-58
        walls = unique_faces(box, ["bottom", "left", "right", "front", "back"], 3)
        return [BC.EquilibriumBC(rho=1.0, u=(0.02, 0.0, 0.0), indices=ne["top"]), BC.FullwayBounceBackBC(indices=walls)]
    if name == "cavity_halfway":  # lid-driven cavity, halfway walls
        walls = unique_faces(box, ["bottom", "left", "right", "front", "back"], 3)
        return [BC.HalfwayBounceBackBC(indices=walls), BC.EquilibriumBC(rho=1.0, u=(0.05, 0.0, 0.0), indices=ne["top"])]
    if name == "cavity2d":  # lid_driven_cavity_2d.py:47-57
        walls = unique_faces(box, ["bottom", "left", "right"], 2)
        return [BC.HalfwayBounceBackBC(indices=walls), BC.EquilibriumBC(rho=1.0, u=(0.05, 0.0), indices=ne["top"])]
    if name == "channel2d_zouhe":
        H = shape[1] - 1.0

        def parabola(index):
            y = index[1].astype(np.float64)
            return 4.0 * 0.05 * y * (H - y) / (H * H)

        walls = unique_faces(box, ["bottom", "top"], 2)
        return [
            BC.HalfwayBounceBackBC(indices=walls),
            BC.ZouHeBC("velocity", profile=parabola, indices=ne["left"]),
            BC.ExtrapolationOutflowBC(indices=ne["right"]),
        ]
    if name == "channel2d_regularized":
        walls = unique_faces(box, ["bottom", "top"], 2)
        return [
            BC.FullwayBounceBackBC(indices=walls),
            BC.RegularizedBC("velocity", prescribed_value=(0.04, 0.0), indices=ne["left"]),
            BC.RegularizedBC("pressure", prescribed_value=1.0, indices=ne["right"]),
        ]
    if name == "channel2d_zouhe_pressure":
        walls = unique_faces(box, ["bottom", "top"], 2)
        return [
            BC.HalfwayBounceBackBC(indices=walls),
            BC.ZouHeBC("pressure", prescribed_value=1.005, indices=ne["left"]),
            BC.ZouHeBC("pressure", prescribed_value=1.0, indices=ne["right"]),
        ]
    if name == "sphere3d":  # flow_past_sphere_3d.py:41-125 / wind tunnel
        walls = unique_faces(box, ["bottom", "top", "front", "back"], 3)
        r = max(2, shape[1] // 6)
        sph = sphere_indices(shape, (shape[0] // 4, shape[1] // 2, shape[2] // 2), r)
        return [
            BC.FullwayBounceBackBC(indices=walls),
            BC.RegularizedBC("velocity", prescribed_value=(0.04, 0.0, 0.0), indices=ne["left"]),
            BC.ExtrapolationOutflowBC(indices=ne["right"]),
            BC.HalfwayBounceBackBC(indices=sph),
        ]
    if name == "do_nothing":
        return [BC.DoNothingBC(indices=ne["right"]), BC.HalfwayBounceBackBC(indices=ne["left"])]
    if name == "c2_sphere":  # C2: sphere r = 12 at (64, 64, 64), the rest as the wind tunnel above
        walls = unique_faces(box, ["bottom", "top", "front", "back"], 3)
        sph = sphere_indices(shape, (64, 64, 64), 12)
        return [
            BC.FullwayBounceBackBC(indices=walls),
            BC.RegularizedBC("velocity", prescribed_value=(0.04, 0.0, 0.0), indices=ne["left"]),
            BC.ExtrapolationOutflowBC(indices=ne["right"]),
            BC.HalfwayBounceBackBC(indices=sph),
        ]
    if name == "channel2d_zouhe_interior":  # a Zou-He inlet on an interior column (padded: missing[0] set)
        walls = unique_faces(box, ["bottom", "top"], 2)
        col = [[2] * (shape[1] - 2), list(range(1, shape[1] - 1))]
        return [
            BC.HalfwayBounceBackBC(indices=walls),
            BC.ZouHeBC("velocity", prescribed_value=(0.03, 0.0), indices=col),
            BC.ExtrapolationOutflowBC(indices=ne["right"]),
        ]
    if name == "cavity_grads":  # lid-driven cavity with Grad's-approximation walls
        walls = unique_faces(box, ["bottom", "left", "right", "front", "back"], 3)
        return [BC.GradsApproximationBC(indices=walls), BC.EquilibriumBC(rho=1.0, u=(0.05, 0.0, 0.0), indices=ne["top"])]
    if name == "sphere3d_grads":  # wind tunnel with a Grad's-approximation sphere (interior indices, padded)
        walls = unique_faces(box, ["bottom", "top", "front", "back"], 3)
        r = max(2, shape[1] // 6)
        sph = sphere_indices(shape, (shape[0] // 4, shape[1] // 2, shape[2] // 2), r)
        return [
            BC.FullwayBounceBackBC(indices=walls),
            BC.RegularizedBC("velocity", prescribed_value=(0.04, 0.0, 0.0), indices=ne["left"]),
            BC.ExtrapolationOutflowBC(indices=ne["right"]),
            BC.GradsApproximationBC(indices=sph),
        ]
    raise KeyError(name)


CASES = {
    # name: (lattice class name, shape, collision, omega, init kind)
    "periodic": ("D3Q19", (24, 20, 16), "BGK", 1.6, "random"),
    "periodic_kbc": ("D3Q27", (16, 12, 10), "KBC", 1.9, "random"),
    "periodic_q27": ("D3Q27", (16, 12, 10), "BGK", 1.7, "random"),
    "periodic_forced": ("D3Q19", (12, 10, 8), "BGK", 1.2, "random"),
    "periodic_field": ("D2Q9", (32, 24), "BGK", 1.3, "random"),
    "cavity_fullway": ("D3Q19", (20, 20, 20), "BGK", 1.0, "rest"),
    "cavity_halfway": ("D3Q19", (16, 16, 16), "BGK", 1.6779, "rest"),
    "cavity2d": ("D2Q9", (32, 32), "BGK", 1.5, "rest"),
    "channel2d_zouhe": ("D2Q9", (48, 24), "BGK", 1.2, "rest"),
    "channel2d_regularized": ("D2Q9", (40, 20), "KBC", 1.7, "rest"),
    "channel2d_zouhe_pressure": ("D2Q9", (40, 20), "BGK", 1.1, "rest"),
    "sphere3d": ("D3Q27", (40, 24, 24), "KBC", 1.9, "rest"),
    "sphere3d_bgk": ("D3Q19", (40, 24, 24), "BGK", 1.6, "rest"),
    "do_nothing": ("D3Q19", (12, 10, 10), "BGK", 1.0, "random"),
    "cavity_smagorinsky": ("D3Q19", (16, 16, 16), "SmagorinskyLESBGK", 1.95, "random"),
    "channel2d_smagorinsky": ("D2Q9", (48, 24), "SmagorinskyLESBGK", 1.9, "rest"),
    "sphere3d_smagorinsky": ("D3Q27", (40, 24, 24), "SmagorinskyLESBGK", 1.97, "rest"),
    "cavity_grads": ("D3Q19", (16, 16, 16), "BGK", 1.6, "random"),
    "sphere3d_grads": ("D3Q27", (40, 24, 24), "KBC", 1.9, "rest"),
    "sphere3d_grads_bgk": ("D3Q19", (40, 24, 24), "BGK", 1.7, "rest"),
    # smallest grids: every pull wraps onto the cell itself or its only neighbour
    "tiny_periodic3d": ("D3Q19", (1, 1, 2), "BGK", 1.5, "random"),
    "tiny_periodic2d": ("D2Q9", (1, 1), "BGK", 1.3, "random"),
    "tiny_periodic_kbc": ("D3Q27", (2, 1, 3), "KBC", 1.8, "random"),
    "tiny_cavity2d": ("D2Q9", (3, 3), "BGK", 1.5, "rest"),
    "tiny_cavity3d": ("D3Q19", (3, 3, 3), "BGK", 1.0, "rest"),
}


# Full-size configurations, used by name only (not part of the all-case sweeps).
LARGE_CASES = {
    # C1: 64^3 lid-driven cavity, halfway walls, lid u = (0.05, 0, 0), Re = 100 (nu = 0.05*64/100)
    "c1_cavity64": ("D3Q19", (64, 64, 64), "BGK", 1.0 / (3 * 0.05 * 64 / 100 + 0.5), "rest"),
    # C2: D3Q27 KBC wind tunnel 256x128x128, sphere r = 12, inlet 0.04, Re = 200 on the diameter
    "c2_sphere": ("D3Q27", (256, 128, 128), "KBC", 1.0 / (3 * 0.04 * 24 / 200 + 0.5), "rest"),
    # C3: periodic D3Q19 BGK Taylor-Green vortex 512x256x256, u0 = 0.04, nu = 0.01
    "c3_tgv": ("D3Q19", (512, 256, 256), "BGK", 1.0 / (3 * 0.01 + 0.5), "tgv"),
    # the benchmark workload (mlups_3d.py:45-90) at edge 256, and at 64
    "mlups_cavity256": ("D3Q19", (256, 256, 256), "BGK", 1.0, "rest"),
    "mlups_cavity64": ("D3Q19", (64, 64, 64), "BGK", 1.0, "rest"),
    # rows longer than one 256-cell tile (several block-format segments per row)
    "long_periodic3d": ("D3Q19", (4, 6, 600), "BGK", 1.7, "random"),
    "long_periodic2d_kbc": ("D2Q9", (6, 520), "KBC", 1.6, "random"),
    "long_cavity3d": ("D3Q19", (6, 5, 300), "BGK", 1.2, "rest"),
    "long_cavity2d": ("D2Q9", (5, 300), "BGK", 1.5, "rest"),
}
BC_SET = {"mlups_cavity256": "cavity_fullway", "mlups_cavity64": "cavity_fullway", "sphere3d_bgk": "sphere3d", "c1_cavity64": "cavity_halfway", "c3_tgv": "periodic", "cavity_smagorinsky": "cavity_halfway",
          "channel2d_smagorinsky": "channel2d_zouhe", "sphere3d_smagorinsky": "sphere3d",
          "sphere3d_grads_bgk": "sphere3d_grads", "tiny_periodic3d": "periodic", "tiny_periodic2d": "periodic",
          "tiny_periodic_kbc": "periodic", "periodic_q27": "periodic", "tiny_cavity2d": "cavity2d", "tiny_cavity3d": "cavity_fullway",
          "long_periodic3d": "periodic", "long_periodic2d_kbc": "periodic", "long_cavity3d": "cavity_halfway",
          "long_cavity2d": "cavity2d"}

ORACLE_COLLISION = {"BGK": 0, "KBC": 1, "SmagorinskyLESBGK": 2}


def random_init(vs, shape, seed=0, du=0.02):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.01 * rng.standard_normal(shape)
    u = du * rng.standard_normal((vs.d,) + tuple(shape))
    return rho, u


class Pair:
    """One case set up on both the product (GPU) and the oracle (CPU)."""

    def __init__(self, name, policy, seed=0):
        import xlb_amd as xa
        import torch

        lat_name, shape, coll, omega, init = CASES[name] if name in CASES else LARGE_CASES[name]
        bc_name = BC_SET.get(name, name)
        self.name, self.policy, self.omega, self.coll = name, policy, omega, coll
        vs = getattr(xa.velocity_set, lat_name)(policy, xa.ComputeBackend.MI355X)
        xa.init(velocity_set=vs, default_backend=xa.ComputeBackend.MI355X, default_precision_policy=policy)
        self.vs = vs
        self.grid = xa.grid.grid_factory(shape, compute_backend=xa.ComputeBackend.MI355X)
        self.bcs = make_bcs(bc_name, self.grid, xa)
        force_vector, force_field = None, None
        self.force, self.force_field_np = None, None
        shape3 = tuple(self.grid.field_shape)
        if name == "periodic_forced":
            force_vector = np.array([1e-4, -2e-4, 5e-5])
            self.force = force_vector
        if name == "periodic_field":
            rng = np.random.default_rng(seed + 7)
            self.force_field_np = 1e-4 * rng.standard_normal((vs.d,) + shape3)
            force_field = torch.from_numpy(self.force_field_np).to(policy.compute_precision.torch_dtype).cuda()
            # the oracle sees the same compute-precision values
            self.force_field_np = force_field.double().cpu().numpy()
        self.stepper = xa.operator.stepper.IncompressibleNavierStokesStepper(
            self.grid, self.bcs, collision_type=coll, force_vector=force_vector, force_field=force_field
        )
        self.f0, self.f1, self.bc_mask, self.missing = self.stepper.prepare_fields()
        self.o_coll = ORACLE_COLLISION[coll]
        self.cparam = float(self.stepper.collision.collision_param)
        self.fmt, self.comp = policy_codes(policy)
        self.lat = vs.lattice_id
        self.shape3 = shape3
        # oracle state
        self.o_mask, self.o_missing = oracle_masks(self.lat, vs.q, vs.d, shape3, self.bcs)
        f_init = pop_alloc(self.fmt, vs.q, shape3)
        if init == "tgv":
            g = [2 * np.pi / s * np.arange(s) for s in shape3]
            X, Y, Z = np.meshgrid(*g, indexing="ij", sparse=True)
            u = np.zeros((3,) + shape3)
            u[0] = 0.04 * np.sin(X) * np.cos(Y) * np.cos(Z)
            u[1] = -0.04 * np.cos(X) * np.sin(Y) * np.cos(Z)
            rho = np.ones(shape3)
            orc.equilibrium(self.lat, self.fmt, self.comp, rho, u, f_init)
            upload(f_init, self.f0)
            xa.grid.copy_population(self.f1, self.f0)
        elif init == "random":
            rho, u = random_init(vs, shape3, seed)
            orc.equilibrium(self.lat, self.fmt, self.comp, rho, u, f_init)
            upload(f_init, self.f0)
            xa.grid.copy_population(self.f1, self.f0)
        else:
            rho, u = np.ones(shape3), np.zeros((vs.d,) + shape3)
            orc.equilibrium(self.lat, self.fmt, self.comp, rho, u, f_init)
        self.init_rho, self.init_u = rho, u
        self.o_f0 = pop_copy(f_init)
        self.o_f1 = pop_copy(f_init)
        self.o_aux = self._oracle_aux(policy)
        self.o_bcs = [oracle_bc_desc(bc) for bc in self.bcs]

    def _oracle_aux(self, policy, f1=None):
        """Prescribed values: stashed in f_1[0] as the reference does, or, for the block format
        (whose int16 mantissas cannot hold them), returned as the oracle's side table."""
        fmt = policy_codes(policy)[0]
        f1 = self.o_f1 if f1 is None else f1
        side = np.zeros(self.o_mask.size)
        for bc in self.bcs:
            if bc.needs_aux_init:
                vals = oracle_aux(bc, self.o_mask, self.vs.d, policy)
                if fmt == orc.I16_BLOCK:
                    side[self.o_mask.reshape(-1) == bc.id] = vals[self.o_mask.reshape(-1) == bc.id]
                else:
                    orc.aux_init(fmt, f1, self.o_mask, bc.id, vals)
        return side if fmt == orc.I16_BLOCK else None

    def step_product(self, n, use_run=False):
        if use_run:
            self.f0, self.f1 = self.stepper.run(self.f0, self.f1, self.bc_mask, self.missing, self.omega, n)
            return
        for t in range(n):
            self.f0, self.f1 = self.stepper(self.f0, self.f1, self.bc_mask, self.missing, self.omega, t)
            self.f0, self.f1 = self.f1, self.f0

    def step_oracle(self, n):
        for _ in range(n):
            orc.step(self.lat, self.o_coll, self.fmt, self.comp, self.o_f0, self.o_f1, self.o_mask, self.o_missing,
                     self.o_bcs, self.omega, force=self.force, force_field=self.force_field_np, cparam=self.cparam,
                     aux=self.o_aux)
            self.o_f0, self.o_f1 = self.o_f1, self.o_f0

    def oracle_reference_f(self, steps, policy):
        """Decoded populations of an oracle-only run of the same case under another policy
        (e.g. FP64FP64: the noise floor against which FP32 results are judged)."""
        fmt, comp = policy_codes(policy)
        f0 = pop_alloc(fmt, self.vs.q, self.shape3)
        orc.equilibrium(self.lat, fmt, comp, self.init_rho, self.init_u, f0)
        f1 = pop_copy(f0)
        aux = self._oracle_aux(policy, f1)
        for _ in range(steps):
            orc.step(self.lat, self.o_coll, fmt, comp, f0, f1, self.o_mask, self.o_missing, self.o_bcs, self.omega,
                     force=self.force, force_field=self.force_field_np, cparam=self.cparam, aux=aux)
            f0, f1 = f1, f0
        return decode(f0, fmt, self.vs.w)

    def product_f(self):
        return decode(to_np(self.f0), self.fmt, self.vs.w)

    def oracle_f(self):
        return decode(self.o_f0, self.fmt, self.vs.w)
