"""Torch restatement of the reference's JAX backend for the fused step, differentiable by autograd.

TEST INFRASTRUCTURE ONLY: the source of reference-semantics gradients for the adjoint tests.
Nothing in the product package imports this module.  jax is not installed here, so XLB's own
`jax.grad` cannot run; this module restates, with whole-array torch operations in FP64, what the
JAX implementations compute, so torch autograd differentiates the same composition of functions
that `jax.grad` would:

* IncompressibleNavierStokesStepper.jax_implementation_pull (xlb/operator/stepper/nse_stepper.py:147-192):
  stream -> streaming-step BCs in list order -> macroscopic -> equilibrium -> collision ->
  for each BC: update_bc_auxilary_data, then the collision-step BCs -> result;
* Stream (stream.py:18-51): jnp.roll of each direction by c_l;
* macroscopic / quadratic equilibrium / BGK (macroscopic.py, zero_moment.py, first_moment.py,
  quadratic_equilibrium.py:18-25, bgk.py:19-24), exact-difference forcing
  (exact_difference_force.py:45-69, forced_collision.py:34-40), with a per-cell force field in place
  of the reference's uniform vector (the differentiable channel learns one);
* the JAX boundary conditions: HalfwayBounceBackBC (bc_halfway_bounce_back.py:50-60),
  FullwayBounceBackBC (bc_fullway_bounce_back.py:44-50), EquilibriumBC (bc_equilibrium.py:58-66),
  DoNothingBC (bc_do_nothing.py:44-48), ZouHeBC and RegularizedBC with the JAX normals (sum over
  every missing main direction, bc_zouhe.py:137-141) and the full prescribed velocity vector
  (bc_zouhe.py:183-264, bc_regularized.py:66-123), ExtrapolationOutflowBC with its normal taken once
  from the index set (bc_extrapolation_outflow.py:63-129);
* IndicesBoundaryMasker.jax_implementation (indices_boundary_masker.py:42-101): the missing mask
  by streaming a padded all-True border (one pad cell per side, one device).

Only what a single-device FP64 rollout needs is restated (no push scheme, no KBC here).
"""

from collections import Counter

import numpy as np
import torch

from oracle import oracle as orc


class Lattice:
    """c (d, q), w (q,), opp, qi (q, d(d+1)/2) of the oracle's pinned tables, as torch FP64."""

    def __init__(self, lat):
        t = orc.lattice(lat)
        self.d = 2 if lat == 0 else 3
        self.c = np.asarray(t["c"][: self.d], dtype=np.int64)
        self.q = self.c.shape[1]
        self.w = torch.tensor(t["w"], dtype=torch.float64)
        self.opp = list(int(o) for o in t["opp"])
        self.qi = torch.tensor(t["qi"][:, : self.d * (self.d + 1) // 2], dtype=torch.float64)
        self.main = [l for l in range(self.q) if int(np.abs(self.c[:, l]).sum()) == 1]
        cc = []
        for a in range(self.d):
            for b in range(a, self.d):
                cc.append(self.c[a] * self.c[b])
        self.cc = torch.tensor(np.array(cc), dtype=torch.float64)  # (nt, q)
        self.ct = torch.tensor(self.c, dtype=torch.float64)

    def bw(self):
        """w broadcast over a (q, ...) field of dimension d."""
        return self.w.reshape((-1,) + (1,) * self.d)


def stream(L, f):
    """Pull streaming: f_l(x) <- f_l(x - c_l), periodic (stream.py:18-51)."""
    axes = tuple(range(1, L.d + 1))
    return torch.stack([torch.roll(f[l], shifts=tuple(int(v) for v in L.c[:, l]), dims=tuple(a - 1 for a in axes))
                        for l in range(L.q)])


def macroscopic(L, f):
    rho = f.sum(0, keepdim=True)
    u = torch.tensordot(L.ct, f, dims=([1], [0])) / rho
    return rho, u


def equilibrium(L, rho, u):
    cu = 3.0 * torch.tensordot(L.ct, u, dims=([0], [0]))
    usqr = 1.5 * (u * u).sum(0, keepdim=True)
    return rho * L.bw() * (1.0 + cu * (1.0 + 0.5 * cu) - usqr)


def bgk(f, feq, omega):
    return f - omega * (f - feq)


def exact_difference(L, fpc, feq, rho, u, force):
    return fpc + equilibrium(L, rho, u + force) - feq


class Bc:
    """One boundary condition: kind, id, index set and, for Zou-He / regularized, the prescribed
    velocity vector field (d, ...) or density."""

    STREAMING, COLLISION = 0, 1

    def __init__(self, kind, bc_id, indices, velocity=None, rho=None, u=None, needs_padding=False):
        self.kind, self.id = kind, bc_id
        self.indices = np.asarray(indices)
        self.velocity, self.rho, self.u = velocity, rho, u
        self.needs_padding = needs_padding
        self.step = Bc.COLLISION if kind == orc.FULLWAY_BB else Bc.STREAMING
        if kind == orc.EXTRAPOLATION_OUTFLOW:
            self.normal = self._extrapolation_normal(self.indices)

    @staticmethod
    def _extrapolation_normal(indices):
        """bc_extrapolation_outflow.py:63-77: the most common value of each coordinate; the axis
        with the largest count is the normal, negative if that value is 0."""
        common = [Counter(int(v) for v in coord).most_common(1)[0] for coord in indices]
        counts = np.array([k for _, k in common])
        values = np.array([v for v, _ in common])
        normal = counts // counts.max()
        if values[np.argmax(counts)] == 0:
            normal = -normal
        return normal


def jax_masks(L, shape, bcs):
    """bc_mask (1, ...) uint8 and missing mask (q, ...) bool by the JAX masker
    (indices_boundary_masker.py:42-101, one device: one pad cell per side)."""
    d = L.d
    pad = (1,) * d
    bmap = np.zeros(tuple(s + 2 for s in shape), dtype=np.uint8)
    grid_mask = np.ones((L.q,) + tuple(s + 2 for s in shape), dtype=bool)
    inner = tuple(slice(1, -1) for _ in range(d))
    grid_mask[(slice(None),) + inner] = False
    for bc in bcs:
        idx = bc.indices[:d] + np.array(pad)[:, None]
        bmap[tuple(idx)] = bc.id
        interior = np.all((bc.indices[:d] > 0) & (bc.indices[:d] < np.array(shape)[:, None] - 1), axis=0)
        if interior.any() and bc.needs_padding:
            grid_mask[(slice(None),) + tuple(idx)] = True
            push = (idx[:, :, None] + L.c[:, None, :]).reshape(d, -1)
            bmap[tuple(push)] = bc.id
    streamed = stream(L, torch.from_numpy(grid_mask.astype(np.float64))).numpy() > 0.5
    missing = streamed[(slice(None),) + inner]
    return torch.from_numpy(bmap[inner][None].copy()), torch.from_numpy(missing.copy())


def _normals(L, missing):
    """-sum of c over the missing main directions (bc_zouhe.py:137-141)."""
    m = missing[L.main].to(torch.float64)
    return -torch.tensordot(L.ct[:, L.main], m, dims=([1], [0]))


def _known_middle(L, missing):
    known = missing[L.opp]
    middle = ~(missing | known)
    return known.to(torch.float64), middle.to(torch.float64)


def _zouhe(L, bc, f_post, missing, regularize):
    normals = _normals(L, missing)
    known, middle = _known_middle(L, missing)
    fsum = (f_post * middle).sum(0, keepdim=True) + 2.0 * (f_post * known).sum(0, keepdim=True)
    if bc.velocity is not None:
        vel = bc.velocity
        unormal = (normals * vel).sum(0, keepdim=True)
        rho = fsum / (1.0 + unormal)
    else:
        rho = torch.full_like(fsum, float(bc.rho))
        vel = (-1.0 + fsum / rho) * normals
    feq = equilibrium(L, rho, vel)
    fbd = torch.where(missing, f_post[L.opp] + feq - feq[L.opp], f_post)
    if regularize:
        fneq = fbd - feq
        pi = torch.tensordot(L.cc, fneq, dims=([1], [0]))
        fbd = feq + 4.5 * L.bw() * torch.tensordot(L.qi, pi, dims=([1], [0]))
    return fbd


def apply_bc(L, bc, f_pre, f_post, bc_mask, missing):
    """The JAX implementation of one BC: f_pre / f_post as the stepper passes them."""
    boundary = (bc_mask == bc.id).expand(f_post.shape)
    if bc.kind == orc.HALFWAY_BB or bc.kind == orc.EXTRAPOLATION_OUTFLOW:
        return torch.where(missing & boundary, f_pre[L.opp], f_post)
    if bc.kind == orc.FULLWAY_BB:
        return torch.where(boundary, f_pre[L.opp], f_post)
    if bc.kind == orc.DO_NOTHING:
        return torch.where(boundary, f_pre, f_post)
    if bc.kind == orc.EQUILIBRIUM:
        rho = torch.tensor([float(bc.rho)], dtype=torch.float64).reshape((1,) + (1,) * L.d)
        u = torch.tensor(list(bc.u)[: L.d], dtype=torch.float64).reshape((L.d,) + (1,) * L.d)
        return torch.where(boundary, equilibrium(L, rho, u).expand(f_post.shape), f_post)
    if bc.kind in (orc.ZOUHE_VELOCITY, orc.ZOUHE_PRESSURE, orc.REGULARIZED_VELOCITY, orc.REGULARIZED_PRESSURE):
        fbd = _zouhe(L, bc, f_post, missing, bc.kind in (orc.REGULARIZED_VELOCITY, orc.REGULARIZED_PRESSURE))
        return torch.where(boundary, fbd, f_post)
    raise ValueError(f"BC kind {bc.kind} is not restated")


def extrapolation_aux(L, bc, f_pre, f_post, bc_mask, missing):
    """ExtrapolationOutflowBC.update_bc_auxilary_data, JAX form (bc_extrapolation_outflow.py:94-118)."""
    cs = 1.0 / np.sqrt(3.0)
    dims = tuple(range(1, L.d + 1))
    boundary = (bc_mask == bc.id).expand(f_post.shape)
    n = tuple(int(v) for v in bc.normal)
    neighbour = torch.roll(boundary, shifts=tuple(-v for v in n), dims=dims)
    fpop = torch.where(boundary, f_pre, f_post)
    fpop_nb = torch.roll(torch.where(neighbour, f_pre, f_post), shifts=n, dims=dims)
    extrapolated = cs * fpop_nb + (1.0 - cs) * fpop
    known = missing[L.opp]
    return torch.where(boundary & known, extrapolated[L.opp], f_post)


def step(L, f_0, bc_mask, missing, bcs, omega, force=None):
    """One pull step (nse_stepper.py:147-192); returns f_1."""
    f_ps = stream(L, f_0)
    for bc in bcs:
        if bc.step == Bc.STREAMING:
            f_ps = apply_bc(L, bc, f_0, f_ps, bc_mask, missing)
    rho, u = macroscopic(L, f_ps)
    feq = equilibrium(L, rho, u)
    f_pc = bgk(f_ps, feq, omega)
    if force is not None:
        f_pc = exact_difference(L, f_pc, feq, rho, u, force)
    for bc in bcs:
        if bc.kind == orc.EXTRAPOLATION_OUTFLOW:
            f_pc = extrapolation_aux(L, bc, f_ps, f_pc, bc_mask, missing)
        if bc.step == Bc.COLLISION:
            f_pc = apply_bc(L, bc, f_ps, f_pc, bc_mask, missing)
    return f_pc


def rollout(L, f_0, bc_mask, missing, bcs, omega, steps, force=None):
    f = f_0
    for _ in range(steps):
        f = step(L, f, bc_mask, missing, bcs, omega, force)
    return f
