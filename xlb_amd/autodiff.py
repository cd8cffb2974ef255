"""Differentiable rollouts of the fused step (reverse mode by the hand-written adjoint kernel).

``rollout(stepper, f0, bc_mask, missing_mask, omega, num_steps, force_field=None, ...)`` returns the
populations after ``num_steps`` steps as a tensor that participates in torch autograd: gradients
flow to ``f0`` and to the per-cell body-force field.  Memory is bounded by checkpointing: every
``checkpoint_every`` steps (default ceil(sqrt(num_steps))) the state is kept, optionally with its
populations in bfloat16 (the library's checkpoint codec: the non-equilibrium part in bfloat16,
rho and u in float32); the backward pass recomputes each segment from its checkpoint and runs
the adjoint step (xlbh_adjoint_step) through it in reverse.

The reference differentiates with Warp's tape over k-step segments with stored checkpoints
(examples/out_of_core/autodiff_lbm.py:64-127, subroutine/autodiff_stepper_subroutine.py:287-555);
here the adjoint is a hand-written kernel and the per-cell force gradient is accumulated in FP64
over cells and time.

Also provides differentiable torch forms of the equilibrium and of the macroscopic moments for
building initial states and losses.
"""

import math

import numpy as np
import torch

from xlb_amd import _lib
from xlb_amd.operator.collision import ForcedCollision


def equilibrium_torch(rho, u, velocity_set):
    """feq(rho, u) as differentiable torch ops (quadratic_equilibrium.py:18-25); rho (1,...),
    u (d,...) -> (q,...).  Sums over directions are explicit elementwise ops (no BLAS call, which
    misbehaved on ROCm for fields of several GB)."""
    c = np.asarray(velocity_set.c, dtype=np.float64)
    w = np.asarray(velocity_set.w, dtype=np.float64)
    usqr = 1.5 * (u * u).sum(0)
    out = []
    for l in range(velocity_set.q):
        cu = None
        for d in range(velocity_set.d):
            if c[d, l] != 0:
                term = float(c[d, l]) * u[d]
                cu = term if cu is None else cu + term
        if cu is None:
            out.append(rho[0] * w[l] * (1.0 - usqr))
        else:
            cu = 3.0 * cu
            out.append(rho[0] * w[l] * (1.0 + cu * (1.0 + 0.5 * cu) - usqr))
    return torch.stack(out)


def macroscopic_torch(f, velocity_set):
    """rho (1, ...), u (d, ...) of populations f (q, ...) as differentiable torch ops."""
    c = np.asarray(velocity_set.c, dtype=np.float64)
    rho = f.sum(0, keepdim=True)
    mom = []
    for d in range(velocity_set.d):
        pos = [l for l in range(velocity_set.q) if c[d, l] > 0]
        neg = [l for l in range(velocity_set.q) if c[d, l] < 0]
        mom.append(f[pos].sum(0) - f[neg].sum(0))
    return rho, torch.stack(mom) / rho


class _Checkpoint:
    """State at a segment start.  ``mode``:

    * "store": a copy in the store dtype;
    * "bf16": the checkpoint codec of the library (xlbh_checkpoint_encode): the non-equilibrium
      part f_l - feq_l(rho, u) as bfloat16 plus rho - 1 and u as float32 ((2 Q + 4 (d + 1)) bytes per
      cell; 0.71 of an FP32 field for D3Q19);
    * "bf16_shifted": f - w_l as bfloat16 (half of an FP32 field; the rounding then also perturbs
      rho and u, by about 2^-9 of |f - w_l|)."""

    def __init__(self, f, mode, w, stepper):
        self.mode = mode
        if mode == "bf16":
            q, ext = f.shape[0], tuple(f.shape[1:])
            self.data = torch.empty((q,) + ext, dtype=torch.bfloat16, device=f.device)
            self.moments = torch.empty((stepper.velocity_set.d + 1,) + ext, dtype=torch.float32, device=f.device)
            _lib.check(_lib.lib().xlbh_checkpoint_encode(
                stepper.lattice_id, stepper.precision_policy.store_format, _lib.ptr(f), _lib.ptr(self.data),
                _lib.ptr(self.moments), *ext, f.stride(0), _lib.stream_of(f)), "checkpoint encode")
        elif mode == "bf16_shifted":
            self.data = (f.float() - w).to(torch.bfloat16)
        else:
            self.data = f.clone()

    def state(self, scratch, w, stepper):
        """A store-dtype tensor holding the state (``scratch`` is filled for 16-bit checkpoints)."""
        if self.mode == "bf16":
            _lib.check(_lib.lib().xlbh_checkpoint_decode(
                stepper.lattice_id, stepper.precision_policy.store_format, _lib.ptr(self.data), _lib.ptr(self.moments),
                _lib.ptr(scratch), *scratch.shape[1:], scratch.stride(0), _lib.stream_of(scratch)), "checkpoint decode")
            _lib.mark_written(scratch)
            return scratch
        if self.mode == "bf16_shifted":
            scratch.copy_(self.data.float() + w)
            return scratch
        return self.data


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f0, force_field, stepper, bc_mask, missing_mask, omega, num_steps, every, ckpt_mode, overlap,
                fused):
        col = stepper.collision
        field = None
        if force_field is not None:
            if not isinstance(col, ForcedCollision) or col.force_field is None:
                raise ValueError("a differentiable force field needs a stepper built with force_field=...")
            # A private copy: the backward pass recomputes this rollout's states with this field even
            # if the caller changes its tensor, or runs another rollout on the stepper, in between.
            field = force_field.detach().to(stepper.compute_dtype, copy=True).contiguous()
            col.force_field = field
        w = torch.as_tensor(np.asarray(stepper.velocity_set.w), dtype=torch.float32, device=f0.device).reshape(-1, 1, 1, 1)
        a = f0.detach().to(stepper.store_dtype).clone().contiguous()
        b = torch.empty_like(a)
        ckpts = {}
        for s in range(0, num_steps, every):
            ckpts[s] = _Checkpoint(a, ckpt_mode, w, stepper)
            a, b = stepper.run(a, b, bc_mask, missing_mask, omega, min(every, num_steps - s), timestep=s)
        ctx.stepper = stepper
        ctx.masks = (bc_mask, missing_mask)
        ctx.omega = omega
        ctx.num_steps = num_steps
        ctx.every = every
        ctx.ckpts = ckpts
        ctx.w = w
        ctx.has_force = force_field is not None
        ctx.force_field = field
        ctx.force_shape = None if force_field is None else tuple(force_field.shape)
        ctx.force_dtype = None if force_field is None else force_field.dtype
        ctx.f0_dtype = f0.dtype
        ctx.overlap = bool(overlap)
        ctx.fused = bool(fused)
        return a

    @staticmethod
    def backward(ctx, grad_out):
        st = ctx.stepper
        bc_mask, missing_mask = ctx.masks
        lam_a = grad_out.detach().to(st.compute_dtype).contiguous().clone()
        lam_b = torch.empty_like(lam_a)
        grad_force = None
        if ctx.has_force:
            grad_force = torch.zeros(ctx.force_shape, dtype=torch.float64, device=lam_a.device)
        n, k = ctx.num_steps, ctx.every
        shape = tuple(lam_a.shape)
        starts = sorted(ctx.ckpts.keys(), reverse=True)
        bf16 = any(c.mode != "store" for c in ctx.ckpts.values())
        nbuf = 2 if ctx.overlap else 1
        states = [torch.empty((max(1, min(k, n) - 1),) + shape, dtype=st.store_dtype, device=lam_a.device)
                  for _ in range(nbuf)]
        scratch = [torch.empty(shape, dtype=st.store_dtype, device=lam_a.device) if bf16 else None for _ in range(nbuf)]
        main = torch.cuda.current_stream(lam_a.device)
        side = torch.cuda.Stream(device=lam_a.device) if ctx.overlap else None
        workspace = st.adjoint_workspace(bc_mask) if ctx.fused else None
        col = st.collision
        caller_field = getattr(col, "force_field", None)
        if ctx.has_force:
            col.force_field = ctx.force_field

        def record(idx):
            """Restore the checkpoint of segment idx and record its states into buffer idx % nbuf."""
            s = starts[idx]
            b = idx % nbuf
            f_s = ctx.ckpts[s].state(scratch[b], ctx.w, st)
            st.record(f_s, states[b], bc_mask, missing_mask, ctx.omega, min(k, n - s) - 1, timestep=s)
            return f_s

        try:
            f_cur = record(0)
            for idx, s in enumerate(starts):
                length = min(k, n - s)
                f_next = None
                if idx + 1 < len(starts) and side is not None:
                    # the buffers of segment idx + 1 were last read by the sweep enqueued on main
                    side.wait_stream(main)
                    with torch.cuda.stream(side):
                        f_next = record(idx + 1)
                out = st.adjoint_run(f_cur, states[idx % nbuf], lam_a, lam_b, bc_mask, missing_mask, ctx.omega,
                                     length, timestep=s, grad_force=grad_force, recompute=False, workspace=workspace,
                                     begin=(idx == 0), end=(idx == len(starts) - 1))
                if out is lam_b:
                    lam_a, lam_b = lam_b, lam_a
                if idx + 1 < len(starts):
                    if side is not None:
                        main.wait_stream(side)
                    else:
                        f_next = record(idx + 1)
                f_cur = f_next
        finally:
            if ctx.has_force:
                col.force_field = caller_field
        if side is not None:
            for t in states + [x for x in scratch if x is not None] + ([ctx.force_field] if ctx.has_force else []):
                t.record_stream(side)
        g_f0 = lam_a.to(ctx.f0_dtype)
        g_force = grad_force.to(ctx.force_dtype) if grad_force is not None else None
        return g_f0, g_force, None, None, None, None, None, None, None, None, None


# Grids up to this many cells leave the GPU partly idle in one step kernel: by default the backward
# pass then records the next segment on a second stream while the current segment's adjoint runs,
# and uses the fused one-kernel reverse step.
OVERLAP_MAX_CELLS = 1 << 22


def rollout(stepper, f0, bc_mask, missing_mask, omega, num_steps, force_field=None, checkpoint_every=None,
            checkpoint_dtype="store", overlap="auto", fused="auto"):
    """Differentiable ``num_steps``-step rollout; see the module docstring.

    checkpoint_dtype: "store" keeps checkpoints in the stepper's store dtype; "bf16" keeps the
    non-equilibrium part f - feq(rho, u) as bfloat16 with rho and u in float32 (0.71 of an FP32
    field for D3Q19); "bf16_shifted" keeps f - w as bfloat16 (half the memory, coarser).  The
    recomputed trajectory starts from the decoded state.
    overlap: record segment i-1 on a second stream while the reverse sweep of segment i runs
    (True / False / "auto" = grids of at most OVERLAP_MAX_CELLS cells); it doubles the recorded
    state workspace.
    fused: the one-kernel reverse step (True / False / "auto" = every lattice but D3Q27, whose fused
    kernel spills registers); it needs a workspace of three population-sized fields in the compute
    dtype, and measured 5-10 % faster than the two-kernel step on 256^3-class D3Q19 grids and 2x
    faster on launch-bound 2-D grids."""
    if num_steps < 1:
        raise ValueError("num_steps must be >= 1")
    every = checkpoint_every or max(1, int(math.ceil(math.sqrt(num_steps))))
    if checkpoint_dtype not in ("store", "bf16", "bf16_shifted"):
        raise ValueError("checkpoint_dtype must be 'store', 'bf16' or 'bf16_shifted'")
    small = f0[0].numel() <= OVERLAP_MAX_CELLS
    if overlap == "auto":
        overlap = small
    if fused == "auto":
        fused = stepper.velocity_set.q != 27
    return _Rollout.apply(f0, force_field, stepper, bc_mask, missing_mask, omega, int(num_steps), int(every),
                          checkpoint_dtype, bool(overlap), bool(fused))
