"""The two differentiable workloads at full size, gradients checked against central finite
differences of FP64 forward runs.

* 2-D channel (D2Q9 BGK 512x256, parabolic regularized inlet u = 0.05, extrapolation outflow, halfway
  walls, per-cell body force g ~ N(0, 1e-12)); loss = mean squared deviation of u at x = 480 from
  a seeded target profile after 800 steps; adjoint with a checkpoint every 50 steps; dL/dg checked
  on 16 random (cell, component) entries.  The FD step is 1e-7 on the FP64 trajectory; the
  configured step 1e-3, a thousand times the force itself, is checked beside it with bars that
  allow for its O(h^2) truncation.
* 3-D decaying turbulence (D3Q19 BGK 128^3 periodic, seeded k^-5/3 field in 2 <= |k| <= 8, rms
  0.03, divergence-free); loss = kinetic energy in 4 <= |k| <= 8 after 2000 steps; gradient with
  respect to the initial velocity through the equilibrium initialisation, sqrt(N) checkpoints;
  FP32 compute with BF16 checkpoints, its split, and FP64; checked along 32 random
  divergence-free Fourier modes of the band.

Bars: FP64 adjoint vs FD <= 1e-4 relative (measured ~1e-8); FP32 adjoints vs FD stated per test.
"""

import math

import numpy as np
import pytest
import torch

import xlb_amd as xa
from xlb_amd.autodiff import equilibrium_torch, macroscopic_torch, rollout
from xlb_amd.workloads import band_energy, decaying_turbulence_3d, differentiable_channel_2d, turbulence_velocity

pytestmark = pytest.mark.gpu

P = xa.PrecisionPolicy


def _channel_loss(vs, target):
    def loss(fT):
        _, u = macroscopic_torch(fT[:, 480:481].double(), vs)
        return ((u[:, 0, :, 0] - target) ** 2).mean()
    return loss


def _channel_target():
    rng = np.random.default_rng(1)
    yy = np.arange(256, dtype=np.float64)
    t = np.zeros((2, 256))
    t[0] = 4 * 0.05 * yy * (255 - yy) / 255**2 * (1 + 0.05 * rng.standard_normal(256))
    return torch.from_numpy(t).cuda()


def _channel_grad(policy, steps, every):
    wl = differentiable_channel_2d((512, 256), policy, seed=0)
    loss = _channel_loss(wl.stepper.velocity_set, _channel_target())
    g = wl.extra["force_field"].clone().requires_grad_(True)
    fT = rollout(wl.stepper, wl.f_0, wl.bc_mask, wl.missing_mask, wl.omega, steps, force_field=g, checkpoint_every=every)
    loss(fT).backward()
    return wl, loss, g.grad.double()


def test_channel_force_gradient_fd():
    steps, every, eps = 800, 50, 1e-7
    wl, loss, grad64 = _channel_grad(P.FP64FP64, steps, every)
    st = wl.stepper
    g0 = wl.extra["force_field"]
    rng = np.random.default_rng(7)
    picks = [(int(rng.integers(2)), int(rng.integers(1, 511)), int(rng.integers(1, 255))) for _ in range(16)]

    def L(g):
        st.collision.force_field = g
        a, _ = st.run(wl.f_0.clone(), wl.f_1.clone(), wl.bc_mask, wl.missing_mask, wl.omega, steps)
        return float(loss(a))

    def central(h):
        out = []
        for d, x, y in picks:
            gp, gm = g0.clone(), g0.clone()
            gp[d, x, y, 0] += h
            gm[d, x, y, 0] -= h
            out.append((L(gp) - L(gm)) / (2 * h))
        return np.array(out)

    fd = central(eps)
    ad64 = np.array([float(grad64[d, x, y, 0]) for d, x, y in picks])
    err64 = np.abs(ad64 - fd) / np.abs(fd)
    print(f"channel FP64: max per-entry rel err {err64.max():.2e} over 16 cells; |fd| in [{np.abs(fd).min():.2e}, {np.abs(fd).max():.2e}]")
    assert err64.max() <= 1e-4
    # The configured FD step, 1e-3 (a thousand times the force): the O(h^2) truncation of the
    # central difference shows (measured rel L2 1.3e-5, worst entry 6.4e-4); bars twice that.
    fd3 = central(1e-3)
    l2_3 = np.linalg.norm(ad64 - fd3) / np.linalg.norm(fd3)
    worst3 = (np.abs(ad64 - fd3) / np.abs(fd3)).max()
    print(f"channel FP64, FD step 1e-3: rel L2 {l2_3:.2e}, max per-entry rel err {worst3:.2e}")
    assert l2_3 <= 3e-5 and worst3 <= 1.3e-3
    _, _, grad32 = _channel_grad(P.FP32FP32, steps, every)
    ad32 = np.array([float(grad32[d, x, y, 0]) for d, x, y in picks])
    r32 = np.linalg.norm(ad32 - fd) / np.linalg.norm(fd)
    full = float((grad32 - grad64).norm() / grad64.norm())
    print(f"channel FP32: rel L2 vs FD over the 16 cells {r32:.2e}; whole-field rel L2 vs FP64 adjoint {full:.2e}")
    assert r32 <= 1e-3
    assert full <= 1e-3


def _band_modes(n, count, seed, band=(4, 8)):
    """Random divergence-free real Fourier modes with band[0] <= |k| <= band[1], max amplitude 1."""
    rng = np.random.default_rng(seed)
    grid = torch.arange(n, dtype=torch.float64, device="cuda")
    x, y, z = torch.meshgrid(grid, grid, grid, indexing="ij")
    modes = []
    while len(modes) < count:
        k = rng.integers(-band[1], band[1] + 1, size=3)
        kk = float(np.linalg.norm(k))
        if not (band[0] <= kk <= band[1]):
            continue
        a = rng.standard_normal(3)
        a -= k * (a @ k) / (k @ k)
        a /= np.linalg.norm(a)
        phase = 2 * math.pi * (k[0] * x + k[1] * y + k[2] * z) / n + rng.uniform(0, 2 * math.pi)
        modes.append(torch.stack([float(a[i]) * torch.cos(phase) for i in range(3)]))
    return modes


def _turbulence_grad(policy, steps, ckpt):
    n = 128
    wl = decaying_turbulence_3d(n, policy, omega=1.95)
    vs = wl.stepper.velocity_set
    u0 = turbulence_velocity(n, seed=3).requires_grad_(True)
    rho = torch.ones((1, n, n, n), dtype=torch.float64, device="cuda")
    f0 = equilibrium_torch(rho, u0, vs).to(wl.stepper.store_dtype)
    fT = rollout(wl.stepper, f0, wl.bc_mask, wl.missing_mask, wl.omega, steps, checkpoint_dtype=ckpt)
    _, u = macroscopic_torch(fT.double(), vs)
    band_energy(u).backward()
    return wl, u0.detach(), u0.grad


def test_turbulence_velocity_gradient_fd():
    steps, eps, n = 2000, 1e-7, 128
    wl, u0, grad64 = _turbulence_grad(P.FP64FP64, steps, "store")
    st, vs = wl.stepper, wl.stepper.velocity_set
    rho = torch.ones((1, n, n, n), dtype=torch.float64, device="cuda")

    def L(u):
        f = equilibrium_torch(rho, u, vs).contiguous()
        a, _ = st.run(f, torch.empty_like(f), wl.bc_mask, wl.missing_mask, wl.omega, steps)
        _, uT = macroscopic_torch(a, vs)
        return float(band_energy(uT))

    modes = _band_modes(n, 32, seed=11)
    fd = np.array([(L(u0 + eps * m) - L(u0 - eps * m)) / (2 * eps) for m in modes])
    ad64 = np.array([float((grad64 * m).sum()) for m in modes])
    err64 = np.abs(ad64 - fd) / np.abs(fd)
    print(f"turbulence FP64: max per-mode rel err {err64.max():.2e} over 32 modes; |fd| in [{np.abs(fd).min():.2e}, {np.abs(fd).max():.2e}]")
    assert err64.max() <= 1e-4
    del wl, st
    torch.cuda.empty_cache()
    # (policy, checkpoints, bar).  C5 as configured is FP32 arithmetic with BF16 checkpoints: the
    # first row (the deviation form of the FP32 step and the checkpoint codec).  The others split
    # the error: plain FP32 arithmetic, the older f - w checkpoints, FP32 checkpoints, and FP64
    # arithmetic on FP32-stored states.  Bars are twice the measurements (in the comments).
    rows = ((P.FP32FP32_DEVIATION, "bf16", 2.1e-5),  # 1.02e-5
            (P.FP32FP32, "bf16", 3.2e-4),  # 1.59e-4
            (P.FP32FP32, "bf16_shifted", 5e-4),  # 2.44e-4
            (P.FP32FP32, "store", 3.2e-4),  # 1.56e-4
            (P.FP64FP32, "store", 8e-6))  # 3.8e-6
    for policy, ckpt, bar in rows:
        _, _, g = _turbulence_grad(policy, steps, ckpt)
        ad = np.array([float((g * m).sum()) for m in modes])
        r = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
        full = float((g - grad64).norm() / grad64.norm())
        print(f"turbulence {policy.name}, checkpoints {ckpt}: rel L2 vs FD over 32 modes {r:.2e}; whole-field rel L2 "
              f"vs FP64 adjoint {full:.2e}")
        assert r <= bar
        del g
        torch.cuda.empty_cache()
