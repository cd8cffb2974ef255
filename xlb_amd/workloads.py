"""Benchmark workloads built through the operator API, as the reference's example scripts build
them: the lid-driven cavity of examples/performance/mlups_3d.py:45-90 and
examples/cfd/lid_driven_cavity_2d.py, the wind tunnel of examples/cfd/flow_past_sphere_3d.py:41-125,
a periodic Taylor-Green vortex (examples/cfd/taylor_green_vortex.py), a differentiable 2-D channel
with a per-cell body force and a differentiable periodic decaying-turbulence box.

Every builder returns a ``Workload``: the stepper, the four fields of ``prepare_fields`` (with
f_0 replaced by the workload's initial state where it has one) and omega.
"""

import math
from dataclasses import dataclass, field

import numpy as np
import torch

import xlb_amd as xa


@dataclass
class Workload:
    stepper: object
    f_0: torch.Tensor
    f_1: torch.Tensor
    bc_mask: torch.Tensor
    missing_mask: torch.Tensor
    omega: float
    extra: dict = field(default_factory=dict)


def _setup(lattice, policy, shape):
    vs = getattr(xa.velocity_set, lattice)(policy, xa.ComputeBackend.MI355X)
    xa.init(velocity_set=vs, default_backend=xa.ComputeBackend.MI355X, default_precision_policy=policy)
    xa.operator.boundary_condition.boundary_condition_registry.reset()
    return vs, xa.grid.grid_factory(shape, compute_backend=xa.ComputeBackend.MI355X)


def _faces(box, names, d):
    walls = [sum((box[n][i] for n in names), []) for i in range(d)]
    return np.unique(np.array(walls), axis=-1).tolist()


def _eq_state(stepper, f, rho, u):
    """f <- feq(rho, u) in f's store format (the product's equilibrium kernel)."""
    cd = stepper.compute_dtype
    eq = xa.operator.equilibrium.QuadraticEquilibrium(stepper.velocity_set, stepper.precision_policy, stepper.compute_backend)
    eq(rho.to(cd).contiguous(), u.to(cd).contiguous(), f)
    return f


def periodic_box_3d(n, policy, omega=1.0, seed=0):
    """D3Q19 BGK periodic box without boundary conditions (the pure step kernel), started from a
    seeded random near-equilibrium flow."""
    vs, grid = _setup("D3Q19", policy, (n, n, n))
    st = xa.operator.stepper.IncompressibleNavierStokesStepper(grid, [], collision_type="BGK")
    f0, f1, bm, mm = st.prepare_fields()
    g = torch.Generator(device="cpu").manual_seed(seed)
    u = (0.02 * torch.randn((3, n, n, n), generator=g, dtype=torch.float32)).to(f0.device)
    rho = torch.ones((1, n, n, n), dtype=torch.float32, device=f0.device)
    _eq_state(st, f0, rho, u)
    del u, rho
    xa.grid.copy_population(f1, f0)
    return Workload(st, f0, f1, bm, mm, omega)


def lid_driven_cavity_3d(n, policy, lid_u=0.02, omega=1.0, walls="fullway", collision="BGK"):
    """mlups_3d.py:45-90: EquilibriumBC lid (top face without edges), bounce-back on the other
    five full faces, rest start.  walls = "fullway" (as mlups_3d.py) or "halfway"; collision =
    "BGK" (mlups_3d.py) or "SmagorinskyLESBGK" (the LES cavity of
    examples/out_of_core/lid_driven_cavity.py: lid 0.06, tau 0.501)."""
    BC = xa.operator.boundary_condition
    vs, grid = _setup("D3Q19", policy, (n, n, n))
    box = grid.bounding_box_indices()
    lid = grid.bounding_box_indices(remove_edges=True)["top"]
    wall_idx = _faces(box, ["bottom", "left", "right", "front", "back"], 3)
    wall = BC.FullwayBounceBackBC(indices=wall_idx) if walls == "fullway" else BC.HalfwayBounceBackBC(indices=wall_idx)
    bcs = [BC.EquilibriumBC(rho=1.0, u=(lid_u, 0.0, 0.0), indices=lid), wall]
    st = xa.operator.stepper.IncompressibleNavierStokesStepper(grid, bcs, collision_type=collision)
    f0, f1, bm, mm = st.prepare_fields()
    return Workload(st, f0, f1, bm, mm, omega)


def wind_tunnel_sphere(shape, policy, radius, center, u_in=0.04, reynolds=200.0, lattice="D3Q27", collision="KBC",
                       sphere_bc="halfway"):
    """flow_past_sphere_3d.py:41-125: sphere as HalfwayBounceBack cells (sphere_bc="halfway") or
    GradsApproximationBC cells ("grads"), RegularizedBC velocity inlet at x = 0,
    ExtrapolationOutflowBC at x = nx-1, FullwayBounceBack on the y/z faces; omega from the Reynolds
    number on the diameter."""
    BC = xa.operator.boundary_condition
    vs, grid = _setup(lattice, policy, shape)
    box = grid.bounding_box_indices()
    ne = grid.bounding_box_indices(remove_edges=True)
    walls = _faces(box, ["bottom", "top", "front", "back"], 3)
    g = np.indices(shape)
    r2 = sum((g[i] - center[i]) ** 2 for i in range(3))
    sphere = [tuple(int(v) for v in a) for a in np.where(r2 < radius**2)]
    bcs = [
        BC.FullwayBounceBackBC(indices=walls),
        BC.RegularizedBC("velocity", prescribed_value=(u_in, 0.0, 0.0), indices=ne["left"]),
        BC.ExtrapolationOutflowBC(indices=ne["right"]),
        BC.HalfwayBounceBackBC(indices=sphere) if sphere_bc == "halfway" else BC.GradsApproximationBC(indices=sphere),
    ]
    nu = u_in * 2 * radius / reynolds
    omega = 1.0 / (3.0 * nu + 0.5)
    st = xa.operator.stepper.IncompressibleNavierStokesStepper(grid, bcs, collision_type=collision)
    f0, f1, bm, mm = st.prepare_fields()
    return Workload(st, f0, f1, bm, mm, omega, extra=dict(sphere_bc=bcs[3]))


def taylor_green_velocity(shape, u0, device="cuda"):
    """u = u0 (sin x cos y cos z, -cos x sin y cos z, 0) on the grid scaled to [0, 2 pi)^3."""
    nx, ny, nz = shape
    x = (2 * math.pi / nx) * torch.arange(nx, dtype=torch.float64, device=device).reshape(-1, 1, 1)
    y = (2 * math.pi / ny) * torch.arange(ny, dtype=torch.float64, device=device).reshape(1, -1, 1)
    z = (2 * math.pi / nz) * torch.arange(nz, dtype=torch.float64, device=device).reshape(1, 1, -1)
    u = torch.zeros((3,) + tuple(shape), dtype=torch.float64, device=device)
    u[0] = u0 * torch.sin(x) * torch.cos(y) * torch.cos(z)
    u[1] = -u0 * torch.cos(x) * torch.sin(y) * torch.cos(z)
    return u


def taylor_green_3d(shape, policy, u0=0.04, nu=0.01):
    """Periodic D3Q19 BGK Taylor-Green vortex, rho = 1 equilibrium start."""
    vs, grid = _setup("D3Q19", policy, shape)
    st = xa.operator.stepper.IncompressibleNavierStokesStepper(grid, [], collision_type="BGK")
    f0, f1, bm, mm = st.prepare_fields()
    u = taylor_green_velocity(shape, u0, f0.device)
    rho = torch.ones((1,) + tuple(shape), dtype=torch.float64, device=f0.device)
    _eq_state(st, f0, rho, u)
    del u, rho
    xa.grid.copy_population(f1, f0)
    return Workload(st, f0, f1, bm, mm, 1.0 / (3.0 * nu + 0.5))


def differentiable_channel_2d(shape, policy, u_in=0.05, omega=1.6, force_std=1e-6, seed=0):
    """D2Q9 BGK channel: parabolic regularized velocity inlet on the left (normal-speed profile as
    flow_past_sphere_3d.py:85-97), extrapolation outflow on the right, halfway bounce-back on the
    full top and bottom faces (lid_driven_cavity_2d.py:55), and a learnable per-cell body-force
    field g ~ N(0, force_std^2) (seeded) applied by the exact-difference scheme."""
    BC = xa.operator.boundary_condition
    vs, grid = _setup("D2Q9", policy, shape)
    box = grid.bounding_box_indices()
    ne = grid.bounding_box_indices(remove_edges=True)
    H = shape[1] - 1.0

    def parabola(index):
        yy = np.asarray(index[1], dtype=np.float64)
        return 4.0 * u_in * yy * (H - yy) / (H * H)

    bcs = [
        BC.HalfwayBounceBackBC(indices=_faces(box, ["bottom", "top"], 2)),
        BC.RegularizedBC("velocity", profile=parabola, indices=ne["left"]),
        BC.ExtrapolationOutflowBC(indices=ne["right"]),
    ]
    rng = np.random.default_rng(seed)
    g = torch.from_numpy(force_std * rng.standard_normal((2,) + tuple(grid.field_shape)))
    g = g.to(policy.compute_precision.torch_dtype).cuda()
    st = xa.operator.stepper.IncompressibleNavierStokesStepper(grid, bcs, collision_type="BGK", force_field=g)
    f0, f1, bm, mm = st.prepare_fields()
    return Workload(st, f0, f1, bm, mm, omega, extra=dict(force_field=g))


def turbulence_velocity(n, rms=0.03, band=(2, 8), slope=-5.0 / 3.0, seed=0, device="cuda"):
    """Seeded Gaussian random velocity field on the periodic n^3 box: energy spectrum E(k) ~
    k^slope for band[0] <= |k| <= band[1] (zero elsewhere), projected divergence-free in Fourier
    space, scaled to the given rms per component."""
    gen = torch.Generator(device="cpu").manual_seed(seed)
    k1 = torch.fft.fftfreq(n, d=1.0 / n, dtype=torch.float64)
    kx, ky, kz = torch.meshgrid(k1, k1, k1, indexing="ij")
    kk = torch.sqrt(kx * kx + ky * ky + kz * kz)
    amp = torch.zeros_like(kk)
    sel = (kk >= band[0]) & (kk <= band[1])
    amp[sel] = torch.sqrt(kk[sel] ** slope / (4 * math.pi * kk[sel] ** 2))
    uh = torch.complex(torch.randn((3, n, n, n), generator=gen, dtype=torch.float64),
                       torch.randn((3, n, n, n), generator=gen, dtype=torch.float64)) * amp
    k = torch.stack([kx, ky, kz]).to(torch.complex128)
    k2 = torch.where(kk > 0, kk * kk, torch.ones_like(kk))
    uh = uh - k * (k * uh).sum(0, keepdim=True) / k2
    u = torch.fft.ifftn(uh, dim=(1, 2, 3)).real
    u = u * (rms / u.pow(2).mean().sqrt())
    return u.to(device)


def band_energy(u, band=(4, 8)):
    """Kinetic energy 0.5 sum |u_hat|^2 / n^3 of the modes with band[0] <= |k| <= band[1]
    (differentiable; u is (3, n, n, n))."""
    n = u.shape[-1]
    k1 = torch.fft.fftfreq(n, d=1.0 / n, dtype=torch.float64, device=u.device)
    kx, ky, kz = torch.meshgrid(k1, k1, k1, indexing="ij")
    kk = torch.sqrt(kx * kx + ky * ky + kz * kz)
    sel = ((kk >= band[0]) & (kk <= band[1])).to(u.dtype)
    uh = torch.fft.fftn(u, dim=(1, 2, 3), norm="forward")
    return 0.5 * ((uh.real ** 2 + uh.imag ** 2) * sel).sum()


def decaying_turbulence_3d(n, policy, omega=1.8):
    """Periodic D3Q19 BGK box for the differentiable decaying-turbulence rollout; the initial
    state is feq(1, u0) of a ``turbulence_velocity`` field, built by the caller so that it can
    carry gradients."""
    vs, grid = _setup("D3Q19", policy, (n, n, n))
    st = xa.operator.stepper.IncompressibleNavierStokesStepper(grid, [], collision_type="BGK")
    f0, f1, bm, mm = st.prepare_fields()
    return Workload(st, f0, f1, bm, mm, omega)


def icosphere(center, radius, subdivisions):
    """Triangle soup (20 * 4^k, 3, 3) of a subdivided icosahedron projected on a sphere, the kind of
    mesh the reference's wind tunnel loads from STL (windtunnel_3d.py:76-95)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    tris = np.array([[v[a], v[b], v[c]] for a, b, c in f], dtype=float)
    tris /= np.linalg.norm(tris, axis=-1, keepdims=True)
    for _ in range(subdivisions):
        a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]
        ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
        ab, bc, ca = (x / np.linalg.norm(x, axis=-1, keepdims=True) for x in (ab, bc, ca))
        tris = np.concatenate([np.stack(s, 1) for s in ((a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca))])
    return tris * radius + np.asarray(center, float)
