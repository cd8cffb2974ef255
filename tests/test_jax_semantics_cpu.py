"""Pins oracle/jax_semantics.py, the restatement of the reference's JAX step used as the source of
reference-semantics gradients for the adjoint.

* On the differentiable channel's geometry (halfway-bounce-back walls on the full top and bottom
  faces, a regularized parabolic velocity inlet and an extrapolation outlet on the faces without
  edges) the JAX and Warp semantics coincide: every boundary cell lies on the domain edge (so both
  maskers give the same missing mask), no boundary has edge or corner cells with two normals, and
  the inflow is normal (so the JAX full velocity vector equals the Warp normal speed times n).
  There the restatement must reproduce the C++ oracle (Warp semantics, pinned by
  test_oracle_cpu.py) to FP64 round-off, including a per-cell body force.
* Where the semantics differ (Zou-He at a face's corner cells: the JAX normal sums both missing
  main directions, the Warp normal takes the first) the two must differ, which shows the
  restatement follows the JAX rule rather than copying the Warp one.
* The masker restatement reproduces the JAX rule that every domain-edge cell gets the missing
  directions that point out of the domain, whether or not a boundary condition is there.
* Autograd through the restatement matches central differences (FP64).
"""

import numpy as np
import pytest
import torch

from oracle import jax_semantics as js
from oracle import oracle as orc


def _faces(shape, no_edge):
    nx, ny = shape
    g = np.indices(shape)
    sel = {"left": g[0] == 0, "right": g[0] == nx - 1, "bottom": g[1] == 0, "top": g[1] == ny - 1}
    if no_edge:
        sel["left"] &= (g[1] > 0) & (g[1] < ny - 1)
        sel["right"] &= (g[1] > 0) & (g[1] < ny - 1)
    return {k: np.array(np.where(v)) for k, v in sel.items()}


def channel(shape, inlet_kind=orc.REGULARIZED_VELOCITY, inlet_edges=False, u_in=0.05):
    """(JAX-semantics BC list, parabolic inlet speed per cell) of the differentiable channel."""
    nx, ny = shape
    full, ne = _faces(shape, False), _faces(shape, True)
    walls = np.concatenate([full["bottom"], full["top"]], axis=1)
    inlet = full["left"] if inlet_edges else ne["left"]
    H = ny - 1.0
    y = np.arange(ny, dtype=np.float64)
    speed = np.broadcast_to(4.0 * u_in * y * (H - y) / (H * H), shape).copy()
    vel = torch.zeros((2,) + shape, dtype=torch.float64)
    vel[0] = torch.from_numpy(speed)
    bcs = [js.Bc(orc.HALFWAY_BB, 1, walls, needs_padding=True),
           js.Bc(inlet_kind, 2, inlet, velocity=vel, needs_padding=True),
           js.Bc(orc.EXTRAPOLATION_OUTFLOW, 3, ne["right"])]
    return bcs, speed


def initial_state(L, shape, seed):
    rng = np.random.default_rng(seed)
    rho = torch.from_numpy(1.0 + 0.01 * rng.standard_normal((1,) + shape))
    u = torch.from_numpy(0.03 * rng.standard_normal((2,) + shape))
    f = js.equilibrium(L, rho, u) + 2e-3 * L.bw() * torch.from_numpy(rng.standard_normal((L.q,) + shape))
    force = torch.from_numpy(1e-4 * rng.standard_normal((2,) + shape))
    return f, force


def warp_run(L, shape, bcs, speed, f, force, omega, steps):
    """The same case through the C++ oracle (Warp semantics, Warp masker)."""
    q = L.q
    shape3 = shape + (1,)
    bc_mask = np.zeros((1,) + shape3, np.uint8)
    missing = np.zeros((q,) + shape3, np.uint8)
    for bc in bcs:
        idx = np.vstack([bc.indices, np.zeros(bc.indices.shape[1], int)]).astype(np.int32)
        interior = np.all((bc.indices > 0) & (bc.indices < np.array(shape)[:, None] - 1), axis=0) & bc.needs_padding
        orc.mask_indices(0, idx, interior.astype(np.uint8), bc.id, bc_mask, missing)
    f0 = f.numpy().reshape((q,) + shape3).copy()
    f1 = f0.copy()
    orc.aux_init(orc.F64, f1, bc_mask, 2, speed.reshape(-1))
    descs = [dict(id=bc.id, kind=bc.kind) for bc in bcs]
    ff = force.numpy().reshape((2,) + shape3)
    for _ in range(steps):
        orc.step(0, 0, orc.F64, 1, f0, f1, bc_mask, missing, descs, omega, force_field=ff)
        f0, f1 = f1, f0
    return f0.reshape((q,) + shape), bc_mask, missing


@pytest.mark.parametrize("inlet_kind", [orc.REGULARIZED_VELOCITY, orc.ZOUHE_VELOCITY])
def test_jax_restatement_matches_warp_oracle_on_the_channel(inlet_kind):
    L = js.Lattice(0)
    shape = (40, 20)
    bcs, speed = channel(shape, inlet_kind)
    f, force = initial_state(L, shape, 1)
    bc_mask, missing = js.jax_masks(L, shape, bcs)
    out = js.rollout(L, f, bc_mask, missing, bcs, 1.6, 60, force)
    ref, w_mask, w_missing = warp_run(L, shape, bcs, speed, f, force, 1.6, 60)
    np.testing.assert_array_equal(bc_mask.numpy(), w_mask[..., 0])
    np.testing.assert_array_equal(missing.numpy(), w_missing[..., 0].astype(bool))
    err = np.abs(out.numpy() - ref).max()
    assert err < 1e-14, err


def test_jax_and_warp_semantics_differ_at_inlet_corners():
    """A Zou-He inlet over the whole left face (corners included): at the corners the JAX normal
    (-1, +-1) differs from the Warp one (first missing main direction), so the trajectories differ
    there after one step; elsewhere on the first step they agree."""
    L = js.Lattice(0)
    shape = (16, 12)
    bcs, speed = channel(shape, orc.ZOUHE_VELOCITY, inlet_edges=True)
    bcs = [bcs[1], bcs[2]]  # no walls: the inlet owns the corner cells
    f, force = initial_state(L, shape, 2)
    bc_mask, missing = js.jax_masks(L, shape, bcs)
    out = js.rollout(L, f, bc_mask, missing, bcs, 1.3, 1, force).numpy()
    ref, _, _ = warp_run(L, shape, bcs, speed, f, force, 1.3, 1)
    diff = np.abs(out - ref).max(axis=0)
    corners = np.zeros(shape, bool)
    corners[0, 0] = corners[0, -1] = True
    assert diff[corners].min() > 1e-6
    assert diff[~corners].max() < 1e-15


def test_jax_masker_marks_every_domain_edge():
    """indices_boundary_masker.py:42-101: the padded border is solid for the streaming of the grid
    mask, so every edge cell has its outward-pulling directions missing, boundary or not; an
    interior padded BC marks the neighbours' missing directions and its push cells' ids."""
    L = js.Lattice(0)
    shape = (9, 7)
    body = js.Bc(orc.HALFWAY_BB, 4, np.array([[4, 4], [3, 4]]), needs_padding=True)
    bc_mask, missing = js.jax_masks(L, shape, [body])
    m = missing.numpy()
    for l in range(L.q):
        cx, cy = L.c[:, l]
        for x in range(shape[0]):
            for y in range(shape[1]):
                src = (x - cx, y - cy)
                out_of_domain = not (0 <= src[0] < shape[0] and 0 <= src[1] < shape[1])
                in_body = src in ((4, 3), (4, 4))
                assert m[l, x, y] == (out_of_domain or in_body), (l, x, y)
    ids = bc_mask.numpy()[0]
    expect = np.zeros(shape, np.uint8)
    for p in ((4, 3), (4, 4)):
        for l in range(L.q):
            expect[p[0] + L.c[0, l], p[1] + L.c[1, l]] = 4
    np.testing.assert_array_equal(ids, expect)


def test_autograd_of_the_restatement_matches_central_differences():
    L = js.Lattice(0)
    shape = (12, 8)
    bcs, _ = channel(shape)
    f, force = initial_state(L, shape, 3)
    bc_mask, missing = js.jax_masks(L, shape, bcs)
    W = torch.from_numpy(np.random.default_rng(4).standard_normal((L.q,) + shape))

    def loss(f0, g):
        return (W * js.rollout(L, f0, bc_mask, missing, bcs, 1.4, 5, g)).sum()

    f0 = f.clone().requires_grad_(True)
    g = force.clone().requires_grad_(True)
    loss(f0, g).backward()
    rng = np.random.default_rng(5)
    for _ in range(3):
        v = torch.from_numpy(rng.standard_normal(f.shape))
        vg = torch.from_numpy(rng.standard_normal(force.shape))
        eps = 1e-6
        with torch.no_grad():
            fd = (loss(f + eps * v, force + eps * vg) - loss(f - eps * v, force - eps * vg)) / (2 * eps)
        ad = (f0.grad * v).sum() + (g.grad * vg).sum()
        assert abs(float(ad - fd)) <= 1e-8 * abs(float(fd)), (float(ad), float(fd))
