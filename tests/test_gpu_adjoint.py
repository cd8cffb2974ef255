"""GPU checks of the hand-written adjoint step (xlbh_adjoint_step) and the checkpointed rollout.

The reference differentiates its Warp kernels with wp.Tape (autodiff_stepper_subroutine.py);
its JAX-backend gradients cannot be produced here (jax is absent), so the adjoint is pinned by
the mathematics instead:
* central finite differences of the forward stepper in FP64 (the forward itself is pinned
  against the oracle by the stepper tests), entry by entry on tiny grids and along random and
  boundary-localised directions on larger ones, for every boundary kind, body force and lattice;
* FP32 adjoints against FP64 adjoints of the same trajectory;
* checkpoint spacing (every step, every k, sqrt(N), none) changes nothing but summation order;
  BF16 checkpoints of f - w stay within a stated bound.
Relative error bars for FP64 against central FD (eps = 1e-6): 1e-8 of the largest entry for
elementwise checks and 2e-8 for directional ones; the measured worst cases are 1.2e-9 and 5.5e-9,
the FD round-off level (the target for the method is 1e-4).
"""

import numpy as np
import pytest
import torch

import xlb_amd as xa
from xlb_amd.autodiff import equilibrium_torch, macroscopic_torch, rollout
from lbm_cases import make_bcs

pytestmark = pytest.mark.gpu

P = xa.PrecisionPolicy
FD_ELEMENT_TOL = 1e-8
FD_DIRECTION_TOL = 2e-8

ADJ_CASES = {
    # name: (lattice, shape, boundary set, omega, force)
    "tiny2d_channel": ("D2Q9", (8, 6), "channel2d_zouhe", 1.2, "field"),
    "tiny2d_regularized": ("D2Q9", (8, 6), "channel2d_regularized", 1.7, "field"),
    "tiny2d_zouhe_pressure": ("D2Q9", (8, 6), "channel2d_zouhe_pressure", 1.1, None),
    "tiny2d_cavity": ("D2Q9", (7, 6), "cavity2d", 1.5, "uniform"),
    "tiny3d_sphere": ("D3Q19", (8, 6, 6), "sphere3d", 1.6, "field"),
    "tiny3d_do_nothing": ("D3Q27", (5, 4, 4), "do_nothing", 1.3, "field"),
    "tiny3d_grads_sphere": ("D3Q19", (8, 6, 6), "sphere3d_grads", 1.6, "field"),
    "tiny3d_grads_cavity": ("D3Q19", (5, 5, 4), "cavity_grads", 1.4, None),
    "tiny2d_zouhe_interior": ("D2Q9", (8, 6), "channel2d_zouhe_interior", 1.2, "field"),
    "tiny3d_smagorinsky": ("D3Q19", (6, 5, 5), "cavity_halfway", 1.9, "field", "SmagorinskyLESBGK"),
    "tiny2d_smagorinsky": ("D2Q9", (8, 6), "channel2d_regularized", 1.95, "uniform", "SmagorinskyLESBGK"),
    "tiny2d_kbc": ("D2Q9", (8, 6), "channel2d_zouhe", 1.9, "field", "KBC"),
    # not with do-nothing cells: a cell that never streams loses its higher-order part dh in one KBC
    # step, and gamma's term (sp1 / sp2) dh is homogeneous of degree 0 in dh, not differentiable there
    "tiny3d_kbc": ("D3Q27", (6, 5, 5), "cavity_halfway", 1.7, "uniform", "KBC"),
    "periodic2d": ("D2Q9", (24, 20), "periodic", 1.3, "field"),
    "periodic3d_uniform": ("D3Q19", (10, 8, 6), "periodic", 1.6, "uniform"),
    "channel2d_zouhe": ("D2Q9", (32, 16), "channel2d_zouhe", 1.2, "field"),
    "channel2d_regularized": ("D2Q9", (32, 16), "channel2d_regularized", 1.7, "field"),
    "cavity_fullway": ("D3Q19", (12, 12, 12), "cavity_fullway", 1.0, "field"),
    "cavity_halfway": ("D3Q19", (12, 12, 12), "cavity_halfway", 1.6779, None),
    "sphere3d": ("D3Q19", (24, 16, 16), "sphere3d", 1.6, "field"),
    "sphere3d_q27": ("D3Q27", (20, 12, 12), "sphere3d", 1.6, "field"),
    "do_nothing": ("D3Q19", (12, 10, 10), "do_nothing", 1.0, "field"),
    "grads_sphere": ("D3Q19", (24, 16, 16), "sphere3d_grads", 1.6, "field"),
    "grads_sphere_q27": ("D3Q27", (20, 12, 12), "sphere3d_grads", 1.6, "field"),
    "smagorinsky_sphere": ("D3Q19", (24, 16, 16), "sphere3d", 1.9, "field", "SmagorinskyLESBGK"),
    "smagorinsky_q27": ("D3Q27", (16, 12, 12), "sphere3d_grads", 1.95, None, "SmagorinskyLESBGK"),
    "kbc_sphere_q27": ("D3Q27", (20, 12, 12), "sphere3d", 1.9, "field", "KBC"),
    "kbc_channel2d": ("D2Q9", (32, 16), "channel2d_regularized", 1.95, "field", "KBC"),
}

TINY = [k for k in ADJ_CASES if k.startswith("tiny")]
LARGER = [k for k in ADJ_CASES if not k.startswith("tiny")]


class Case:
    def __init__(self, name, policy, seed=0):
        lat, shape, bcname, omega, force = ADJ_CASES[name][:5]
        collision = ADJ_CASES[name][5] if len(ADJ_CASES[name]) > 5 else "BGK"
        vs = getattr(xa.velocity_set, lat)(policy, xa.ComputeBackend.MI355X)
        xa.init(velocity_set=vs, default_backend=xa.ComputeBackend.MI355X, default_precision_policy=policy)
        grid = xa.grid.grid_factory(shape, compute_backend=xa.ComputeBackend.MI355X)
        bcs = make_bcs(bcname, grid, xa)
        shape3 = tuple(grid.field_shape)
        rng = np.random.default_rng(seed)
        cdt = policy.compute_precision.torch_dtype
        self.ff = None
        fv = None
        if force == "field":
            self.ff = torch.from_numpy(2e-4 * rng.standard_normal((vs.d,) + shape3)).to(cdt).cuda()
        elif force == "uniform":
            fv = np.array([1e-4, -2e-4, 5e-5][: vs.d])
        self.st = xa.operator.stepper.IncompressibleNavierStokesStepper(
            grid, bcs, collision_type=collision, force_vector=fv, force_field=self.ff
        )
        _, _, self.bc_mask, self.missing = self.st.prepare_fields()
        # generic state: equilibrium of a random flow plus a non-equilibrium part
        rho = torch.from_numpy(1.0 + 0.01 * rng.standard_normal((1,) + shape3)).cuda()
        u = torch.from_numpy(0.03 * rng.standard_normal((vs.d,) + shape3)).cuda()
        w = torch.tensor(np.asarray(vs.w, dtype=np.float64), device="cuda").reshape((-1,) + (1,) * 3)
        f = equilibrium_torch(rho, u, vs) + 2e-3 * w * torch.from_numpy(rng.standard_normal((vs.q,) + shape3)).cuda()
        self.f0 = f.to(self.st.store_dtype).contiguous()
        self.W = torch.from_numpy(rng.standard_normal((vs.q,) + shape3)).cuda()
        self.vs, self.omega, self.policy = vs, omega, policy

    def forward(self, f0, steps, ff=None):
        if ff is not None:
            self.st.collision.force_field = ff
        a = f0.clone()
        b = torch.empty_like(a)
        for t in range(steps):
            self.st(a, b, self.bc_mask, self.missing, self.omega, t)
            a, b = b, a
        return a

    def loss(self, fT):
        fT = fT.double()
        _, u = macroscopic_torch(fT, self.vs)
        return (self.W * fT).sum() + 50.0 * (u * u).sum()

    def gradients(self, steps, every=None, ckpt="store", overlap="auto", fused="auto"):
        f0 = self.f0.clone().requires_grad_(True)
        ff = self.ff.clone().requires_grad_(True) if self.ff is not None else None
        fT = rollout(self.st, f0, self.bc_mask, self.missing, self.omega, steps, force_field=ff,
                     checkpoint_every=every, checkpoint_dtype=ckpt, overlap=overlap, fused=fused)
        self.loss(fT).backward()
        return f0.grad.double(), (ff.grad.double() if ff is not None else None)

    def loss_at(self, f0, steps, ff=None):
        return float(self.loss(self.forward(f0, steps, ff)))


def rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def central(fun, eps, richardson):
    """Central difference of fun(t) at t = 0; with ``richardson`` the two-step extrapolation
    (4 D(eps/2) - D(eps)) / 3, whose truncation error is O(eps^4).  KBC's entropic gamma is a ratio of
    two quantities quadratic in fneq, so plain central differences carry an O((eps/fneq)^2) error
    there."""
    d1 = (fun(eps) - fun(-eps)) / (2 * eps)
    if not richardson:
        return d1
    d2 = (fun(eps / 2) - fun(-eps / 2)) / eps
    return (4 * d2 - d1) / 3


@pytest.mark.parametrize("name", TINY)
def test_adjoint_elementwise_fd(name):
    """Every entry of dL/df0 and dL/dg against central FD (FP64, 3 steps, checkpoint every 2)."""
    c = Case(name, P.FP64FP64)
    steps, eps = 3, 1e-6
    rich = c.st.collision_type == "KBC"
    g_f, g_g = c.gradients(steps, every=2)
    fd = torch.zeros_like(g_f)
    flat0 = c.f0.reshape(-1)
    for i in range(flat0.numel()):
        def fun(t):
            f = c.f0.clone()
            f.view(-1)[i] += t
            return c.loss_at(f, steps)

        fd.view(-1)[i] = central(fun, eps, rich)
    scale = fd.abs().max().item()
    err = (g_f - fd).abs().max().item()
    print(f"{name}: f0 max|adj - fd| / scale = {err / scale:.2e}")
    assert err <= FD_ELEMENT_TOL * scale, f"{name}: max|adj - fd| = {err:.3e} (scale {scale:.3e})"
    if c.ff is not None:
        fdg = torch.zeros_like(g_g)
        for i in range(c.ff.numel()):
            def fung(t):
                g = c.ff.clone()
                g.view(-1)[i] += t
                return c.loss_at(c.f0, steps, g)

            fdg.view(-1)[i] = central(fung, eps, rich)
        c.st.collision.force_field = c.ff
        scale = fdg.abs().max().item()
        err = (g_g - fdg).abs().max().item()
        print(f"{name}: force max|adj - fd| / scale = {err / scale:.2e}")
        assert err <= FD_ELEMENT_TOL * scale, f"{name}: force grad max|adj - fd| = {err:.3e} (scale {scale:.3e})"


def _near_boundary(c):
    """1 at boundary cells and their face neighbours, else 0 (shape of bc_mask); None if no BC."""
    bc = (c.bc_mask != 0).double()
    if not bc.any():
        return None
    near = bc.clone()
    for ax in range(1, bc.dim()):
        if bc.shape[ax] > 1:
            near = torch.maximum(near, torch.roll(bc, 1, ax))
            near = torch.maximum(near, torch.roll(bc, -1, ax))
    return near


def _directions(shape, near, seed):
    """A global random direction and the same one confined to the cells of ``near``."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    v = torch.randn(shape, generator=g, dtype=torch.float64).cuda()
    dirs = [("global", v)]
    if near is not None:
        dirs.append(("boundary", v * near))
    return dirs


@pytest.mark.parametrize("name", LARGER)
def test_adjoint_directional_fd(name):
    """Directional derivatives along random and boundary-localised directions (FP64, 8 steps,
    sqrt-N checkpointing)."""
    c = Case(name, P.FP64FP64, seed=1)
    steps, eps = 8, 1e-6
    rich = c.st.collision_type == "KBC"
    g_f, g_g = c.gradients(steps)
    near = _near_boundary(c)
    for tag, v in _directions(c.f0.shape, near, 3):
        fd = central(lambda t: c.loss_at(c.f0 + t * v, steps), eps, rich)
        ad = float((g_f * v).sum())
        print(f"{name} f0/{tag}: rel {rel(ad, fd):.2e}")
        # KBC at regularized boundary cells: their non-equilibrium part is the second-order Hermite
        # projection, which leaves dh (the part outside the shear) small, so gamma's sp1 / sp2 is
        # steep there and FD loses digits (measured 1.7e-7); 1e-6 for KBC's boundary directions
        tol = 1e-6 if rich and tag == "boundary" else FD_DIRECTION_TOL
        assert rel(ad, fd) <= tol, f"{name} f0/{tag}: adjoint {ad:.10e} vs FD {fd:.10e}"
    if c.ff is not None:
        for tag, v in _directions(c.ff.shape, near, 5):
            fd = central(lambda t: c.loss_at(c.f0, steps, c.ff + t * v), eps, rich)
            c.st.collision.force_field = c.ff
            ad = float((g_g * v).sum())
            print(f"{name} force/{tag}: rel {rel(ad, fd):.2e}")
            assert rel(ad, fd) <= FD_DIRECTION_TOL, f"{name} force/{tag}: adjoint {ad:.10e} vs FD {fd:.10e}"


@pytest.mark.parametrize("name", ["channel2d_zouhe", "sphere3d", "periodic3d_uniform"])
def test_adjoint_fp32_matches_fp64(name):
    """FP32 compute/store adjoint of the same case against the FP64 one: relative L2 <= 1e-5."""
    g64, gg64 = Case(name, P.FP64FP64, seed=2).gradients(10)
    g32, gg32 = Case(name, P.FP32FP32, seed=2).gradients(10)
    r = ((g32 - g64).norm() / g64.norm()).item()
    assert r <= 1e-5, f"{name}: FP32 vs FP64 adjoint rel L2 {r:.3e}"
    if gg64 is not None:
        r = ((gg32 - gg64).norm() / gg64.norm()).item()
        assert r <= 1e-4, f"{name}: FP32 vs FP64 force gradient rel L2 {r:.3e}"


def test_checkpoint_spacing_invariance():
    """Gradients do not depend on the checkpoint spacing beyond summation order (atomic adds at
    boundary cells); BF16 checkpoints of f - w stay within 1e-4 relative L2."""
    steps = 13
    c = Case("channel2d_zouhe", P.FP64FP64, seed=4)
    ref_f, ref_g = c.gradients(steps, every=steps)
    for every, overlap, fused in ((1, True, True), (3, True, True), (None, True, False), (3, False, True),
                                  (None, False, False)):
        g_f, g_g = c.gradients(steps, every=every, overlap=overlap, fused=fused)
        assert ((g_f - ref_f).norm() / ref_f.norm()).item() <= 1e-12
        assert ((g_g - ref_g).norm() / ref_g.norm()).item() <= 1e-12
    c32 = Case("channel2d_zouhe", P.FP32FP32, seed=4)
    e_f, e_g = c32.gradients(steps, every=4)
    b_f, b_g = c32.gradients(steps, every=4, ckpt="bf16")
    b2_f, _ = c32.gradients(steps, every=4, ckpt="bf16", overlap=False)
    assert ((b2_f - b_f).norm() / b_f.norm()).item() <= 1e-6
    r = ((b_f - e_f).norm() / e_f.norm()).item()
    assert r <= 1e-4, f"BF16 checkpoints: rel L2 {r:.3e}"


def test_adjoint_rejects_bad_arguments():
    c = Case("tiny2d_cavity", P.FP64FP64)
    lam = torch.zeros_like(c.f0)
    with pytest.raises(ValueError):
        c.st.adjoint_step(c.f0, lam, lam, c.bc_mask, c.missing, c.omega)
    # KBC is not defined for D3Q19 (kbc.py:183-185): the C ABI rejects its adjoint too
    from xlb_amd import _lib
    c3 = Case("periodic3d_uniform", P.FP64FP64)
    lam3 = torch.zeros_like(c3.f0)
    nx, ny, nz, plane, masks, force_mode, fu_ptr, keep, ff = c3.st._launch_args(c3.f0, None, c3.bc_mask, c3.missing)
    rc = _lib.lib().xlbh_adjoint_step(c3.st.lattice_id, _lib.KBC, c3.st.precision_policy.store_format,
                                      c3.st.compute_id, _lib.ptr(c3.f0), _lib.ptr(lam3), _lib.ptr(torch.zeros_like(lam3)),
                                      *masks, None, 0, force_mode, fu_ptr, _lib.ptr(ff), None, nx, ny, nz,
                                      plane, float(c3.omega), 0.0, _lib.stream_of(c3.f0))
    assert rc != 0 and b"D2Q9 and D3Q27" in _lib.lib().xlbh_last_error()
    # a mask with a solid (255) cell: refused by both entry points, with and without a workspace,
    # at an unaligned mask start (the scan's byte loop) and at an aligned one (its 16-byte loop)
    c = Case("sphere3d", P.FP64FP64)
    st = c.st
    nx, ny, nz, plane, masks, force_mode, fu_ptr, keep, ff = st._launch_args(c.f0, None, c.bc_mask, c.missing)
    lam_out, lam_in = c.W.clone(), torch.empty_like(c.W)
    ncell = nx * ny * nz
    for offset in (1, 0):
        mask = torch.zeros(ncell + 16, dtype=torch.uint8, device="cuda")
        view = mask[offset:offset + ncell]
        view.copy_(c.bc_mask.reshape(-1))
        view[ncell - 3] = 255
        args = (st.lattice_id, st._collision_id(), st.precision_policy.store_format, st.compute_id)
        rc = _lib.lib().xlbh_adjoint_step(*args, _lib.ptr(c.f0), _lib.ptr(lam_out), _lib.ptr(lam_in), _lib.ptr(view),
                                          *masks[1:], None, 0, force_mode, fu_ptr, _lib.ptr(ff), None, nx, ny, nz,
                                          plane, float(c.omega), st._collision_param(), _lib.stream_of(c.f0))
        assert rc != 0 and b"solid" in _lib.lib().xlbh_last_error()
        ws = st.adjoint_workspace(c.bc_mask)
        for workspace, flags in ((None, 0), (ws, _lib.ADJ_BEGIN | _lib.ADJ_END)):
            rc = _lib.lib().xlbh_adjoint_run(*args, _lib.ptr(c.f0), None, _lib.ptr(lam_out), _lib.ptr(lam_in),
                                             _lib.ptr(view), *masks[1:], None, 0, force_mode, fu_ptr, _lib.ptr(ff),
                                             None, nx, ny, nz, plane, float(c.omega), st._collision_param(), 0, 1,
                                             _lib.ptr(workspace), flags, _lib.stream_of(c.f0))
            assert rc != 0 and b"solid" in _lib.lib().xlbh_last_error()
        view[ncell - 3] = 0
        rc = _lib.lib().xlbh_adjoint_step(*args, _lib.ptr(c.f0), _lib.ptr(lam_out), _lib.ptr(lam_in), _lib.ptr(view),
                                          *masks[1:], None, 0, force_mode, fu_ptr, _lib.ptr(ff), None, nx, ny, nz,
                                          plane, float(c.omega), st._collision_param(), _lib.stream_of(c.f0))
        assert rc == 0, _lib.lib().xlbh_last_error()


@pytest.mark.parametrize("name", ["tiny2d_channel", "tiny2d_regularized", "sphere3d", "do_nothing", "grads_sphere"])
def test_boundary_pass_list_equals_grid_scan(name):
    """The boundary pass over the compacted cell list (what the stepper passes) and over the whole
    grid (bc_cells = NULL through the C ABI) give the same adjoint."""
    import ctypes

    from xlb_amd import _lib

    c = Case(name, P.FP64FP64, seed=5)
    st = c.st
    lam_out = c.W.clone()
    via_list = torch.empty_like(lam_out)
    st.adjoint_step(c.f0, lam_out, via_list, c.bc_mask, c.missing, c.omega)
    via_scan = torch.empty_like(lam_out)
    nx, ny, nz, plane, masks, force_mode, fu_ptr, keep, ff = st._launch_args(c.f0, None, c.bc_mask, c.missing)
    _lib.check(_lib.lib().xlbh_adjoint_step(
        st.lattice_id, st._collision_id(), st.precision_policy.store_format, st.compute_id, _lib.ptr(c.f0),
        _lib.ptr(lam_out), _lib.ptr(via_scan), *masks, None, 0, force_mode, fu_ptr, _lib.ptr(ff), None, nx, ny, nz,
        plane, float(c.omega), st._collision_param(), _lib.stream_of(c.f0)))
    torch.cuda.synchronize()
    assert st._bc_cells is not None and st._bc_cells.numel() > 0
    err = (via_list - via_scan).abs().max().item()
    assert err <= 1e-12 * via_scan.abs().max().item()


@pytest.mark.parametrize("name", ["channel2d_zouhe", "channel2d_regularized", "tiny2d_zouhe_pressure", "sphere3d",
                                  "sphere3d_q27", "do_nothing", "cavity_fullway", "grads_sphere", "grads_sphere_q27",
                                  "tiny3d_grads_cavity", "tiny2d_zouhe_interior", "smagorinsky_sphere",
                                  "smagorinsky_q27", "kbc_sphere_q27", "kbc_channel2d"])
def test_fused_and_split_reverse_agree(name):
    """The fused one-kernel reverse step (boundary contributions carried in the workspace, Zou-He /
    regularized Jacobian by the hand-written one-pass adjoint) against the two-kernel reverse step
    (atomics, Zou-He Jacobian by forward-mode dual numbers; Grad's by the same one-pass adjoint),
    over a checkpointed rollout in FP64."""
    c = Case(name, P.FP64FP64, seed=6)
    f_fused, g_fused = c.gradients(9, every=4, fused=True)
    f_split, g_split = c.gradients(9, every=4, fused=False)
    assert ((f_fused - f_split).norm() / f_split.norm()).item() <= 1e-12
    if g_split is not None:
        assert ((g_fused - g_split).norm() / g_split.norm()).item() <= 1e-12


def test_rollout_keeps_its_own_force_field():
    """Each rollout's backward pass recomputes its states with the force field of its own forward
    pass: two rollouts with different fields run before either backward, and the caller's tensor
    changed in place before the backward, give each the gradients of a rollout run alone, and the
    force gradient matches central FD along a random direction."""
    c = Case("channel2d_zouhe", P.FP64FP64, seed=6)
    steps = 6
    fields = (c.ff.clone(), -2.0 * c.ff.clone())

    def start(ff):
        f0 = c.f0.clone().requires_grad_(True)
        g = ff.clone().requires_grad_(True)
        fT = rollout(c.st, f0, c.bc_mask, c.missing, c.omega, steps, force_field=g, checkpoint_every=2)
        return f0, g, c.loss(fT)

    alone = []
    for ff in fields:
        f0, g, loss = start(ff)
        loss.backward()
        alone.append((f0.grad.clone(), g.grad.clone()))
    runs = [start(ff) for ff in fields]
    with torch.no_grad():
        for _, g, _ in runs:
            g.mul_(5.0)  # the caller changes its field tensors after the forward passes
    for (f0, g, loss), (ref_f, ref_g) in zip(runs, alone):
        loss.backward()
        assert ((f0.grad - ref_f).norm() / ref_f.norm()).item() <= 1e-12
        assert ((g.grad - ref_g).norm() / ref_g.norm()).item() <= 1e-12
    v = torch.randn(c.ff.shape, generator=torch.Generator().manual_seed(7), dtype=torch.float64).cuda()
    for ff, (_, ref_g) in zip(fields, alone):
        fd = central(lambda t: c.loss_at(c.f0, steps, ff + t * v), 1e-6, False)
        ad = float((ref_g * v).sum())
        assert rel(ad, fd) <= FD_DIRECTION_TOL, f"adjoint {ad:.10e} vs FD {fd:.10e}"


def test_fp64_q27_reverse_is_bit_stable():
    """The FP64 D3Q27 reverse kernels use every VGPR and spill to scratch.  Their fused path carries
    the boundary contributions in a workspace (no atomics), so its gradients are bit-identical
    between two identical rollouts and with or without the next segment recorded on a second
    stream; a build whose code for these kernels went wrong showed run-to-run differences here."""
    c = Case("smagorinsky_q27", P.FP64FP64, seed=6)
    a, _ = c.gradients(9, every=4, fused=True, overlap=True)
    b, _ = c.gradients(9, every=4, fused=True, overlap=True)
    n, _ = c.gradients(9, every=4, fused=True, overlap=False)
    assert torch.equal(a, b)
    assert torch.equal(a, n)
