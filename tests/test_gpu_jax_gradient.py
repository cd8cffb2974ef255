"""The HIP adjoint against reference-semantics gradients: torch autograd through
oracle/jax_semantics.py, the FP64 restatement of XLB's JAX step (jax itself is absent, so
`jax.grad` over xlb/operator/stepper/nse_stepper.py:147-192 cannot run; the comparison with XLB's
own JAX gradient stays parity unpinned).

The case is the differentiable channel (C4) reduced to 64 x 32 and 100 steps: halfway walls on
the full top and bottom faces, a regularized parabolic velocity inlet and an extrapolation outlet
on the faces without edges, a per-cell body force, BGK, FP64.  On this geometry the JAX and Warp
semantics coincide (tests/test_jax_semantics_cpu.py), so the two gradients, dL/df_0 and dL/dg of
L = sum W * f_T, must agree to FP64 round-off accumulated over 100 steps.
"""

import numpy as np
import pytest
import torch

import xlb_amd as xa
from oracle import jax_semantics as js
from test_jax_semantics_cpu import channel, initial_state
from xlb_amd.autodiff import rollout
from xlb_amd.workloads import differentiable_channel_2d

pytestmark = pytest.mark.gpu


def test_hip_adjoint_matches_autograd_of_the_jax_restatement():
    shape, steps, omega = (64, 32), 100, 1.6
    L = js.Lattice(0)
    f, force = initial_state(L, shape, 3)
    W = torch.from_numpy(np.random.default_rng(9).standard_normal((L.q,) + shape))
    # reference semantics: autograd through the restatement (CPU, FP64)
    bcs, _ = channel(shape)
    bc_mask, missing = js.jax_masks(L, shape, bcs)
    f_ref = f.clone().requires_grad_(True)
    g_ref = force.clone().requires_grad_(True)
    (W * js.rollout(L, f_ref, bc_mask, missing, bcs, omega, steps, g_ref)).sum().backward()
    # the product: the hand-written adjoint through the checkpointed rollout (GPU, FP64)
    wl = differentiable_channel_2d(shape, xa.PrecisionPolicy.FP64FP64, omega=omega)
    q, nx, ny = L.q, shape[0], shape[1]
    f0 = f.reshape(q, nx, ny, 1).cuda().requires_grad_(True)
    g = force.reshape(2, nx, ny, 1).cuda().requires_grad_(True)
    fT = rollout(wl.stepper, f0, wl.bc_mask, wl.missing_mask, omega, steps, force_field=g, checkpoint_every=10)
    (W.reshape(q, nx, ny, 1).cuda() * fT).sum().backward()
    gf = f0.grad.reshape(q, nx, ny).cpu()
    gg = g.grad.reshape(2, nx, ny).cpu()
    ef = float((gf - f_ref.grad).norm() / f_ref.grad.norm())
    eg = float((gg - g_ref.grad).norm() / g_ref.grad.norm())
    print(f"HIP adjoint vs autograd of the JAX-semantics restatement (64x32 channel, 100 steps, FP64): "
          f"dL/df0 rel L2 {ef:.2e}, dL/dg rel L2 {eg:.2e}")
    assert ef <= 1e-10 and eg <= 1e-10
