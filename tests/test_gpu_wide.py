"""Planes above 4 GiB: the 64-bit-offset instantiation of the fused step.

A field whose planes exceed 4 GiB (more than 2^30 cells in FP32, more than 2^31 in 16-bit) is
stepped with 64-bit cell offsets.  No CPU oracle finishes at these sizes, so the check rests on
translation invariance: with periodic wrap in x and x-invariant boundaries, a field that repeats
with period 2 along x steps to a field that repeats with period 2, and each period equals the
step of the 2-plane grid, which the 32-bit-offset kernel runs and the oracle tests cover.  The two
instantiations share their arithmetic, so the comparison is bit-exact.

Each case holds two population fields of about 82 GB (163 GB of HBM) and frees them afterwards.
"""

import numpy as np
import pytest
import torch

import xlb_amd as xa
from xlb_amd import _lib
from xlb_amd.helper.initializers import initialize_eq

pytestmark = pytest.mark.gpu

P = xa.PrecisionPolicy
NY, NZ = 512, 1024


def _face(xs, ys, zs):
    g = np.meshgrid(np.asarray(xs), np.asarray(ys), np.asarray(zs), indexing="ij")
    return [list(map(int, a.reshape(-1))) for a in g]


def _small_case(policy, with_bcs, seed):
    """The 2 x NY x NZ grid: random equilibrium populations and, optionally, walls and a lid on
    the y and z faces (every boundary is x-invariant)."""
    vs = xa.velocity_set.D3Q19(policy, xa.ComputeBackend.MI355X)
    xa.init(velocity_set=vs, default_backend=xa.ComputeBackend.MI355X, default_precision_policy=policy)
    grid = xa.grid.grid_factory((2, NY, NZ), compute_backend=xa.ComputeBackend.MI355X)
    bcs = []
    if with_bcs:
        BC = xa.operator.boundary_condition
        inner_y = range(1, NY - 1)
        bcs = [
            BC.HalfwayBounceBackBC(indices=_face([0, 1], [0], range(NZ))),
            BC.FullwayBounceBackBC(indices=_face([0, 1], [NY - 1], range(NZ))),
            BC.EquilibriumBC(rho=1.0, u=(0.03, 0.01, 0.0), indices=_face([0, 1], inner_y, [NZ - 1])),
            BC.ExtrapolationOutflowBC(indices=_face([0, 1], inner_y, [0])),
        ]
    stepper = xa.operator.stepper.IncompressibleNavierStokesStepper(grid, boundary_conditions=bcs)
    f0, f1, bc_mask, missing = stepper.prepare_fields()
    gen = torch.Generator(device="cuda").manual_seed(seed)
    cdt = stepper.compute_dtype
    rho = 1.0 + 0.01 * torch.randn((1, 2, NY, NZ), generator=gen, device="cuda", dtype=cdt)
    u = 0.03 * torch.randn((3, 2, NY, NZ), generator=gen, device="cuda", dtype=cdt)
    f0 = initialize_eq(f0, grid, vs, policy, xa.ComputeBackend.MI355X, rho=rho, u=u)
    stepper._launch_args(f0, f1, bc_mask, missing)  # derive the packed bits and the BC table
    return stepper, f0, f1, bc_mask


def _tile_x(small, reps):
    """(c, 2, NY, NZ) -> (c, 2 * reps, NY, NZ), contiguous, without a second full-size temporary."""
    c = small.shape[0]
    big = torch.empty((c, 2 * reps, NY, NZ), dtype=small.dtype, device=small.device)
    big.view(c, reps, 2, NY, NZ).copy_(small.view(c, 1, 2, NY, NZ).expand(c, reps, 2, NY, NZ))
    return big


def _args(stepper, bc_mask, bits, with_bcs):
    if not with_bcs:
        return (None, None, None, None)
    return (_lib.ptr(bc_mask), _lib.ptr(bits), None, _lib.ptr(stepper._table))


def _run(stepper, f_a, f_b, masks, nx, omega, steps):
    fu, keep = _lib.doubles((0.0, 0.0, 0.0))
    _lib.check(
        _lib.lib().xlbh_run(
            stepper.lattice_id, _lib.BGK, stepper.precision_policy.store_format, stepper.compute_id,
            _lib.ptr(f_a), _lib.ptr(f_b), *masks, _lib.FORCE_NONE, fu, None, nx, NY, NZ, f_a.stride(0), float(omega), 0.0, 0,
            int(steps), _lib.stream_of(f_a),
        ),
        "xlbh_run",
    )
    return f_a if steps % 2 == 0 else f_b


@pytest.mark.parametrize(
    "policy,reps,with_bcs",
    [
        (P.FP32FP32, 1025, False),  # 2050 x 512 x 1024 FP32: 1.07 G cells, 4.30 GB planes
        (P.FP32FP32, 1025, True),
        (P.FP32FP16_SHIFTED, 2049, False),  # 4098 x 512 x 1024: 2.15 G cells (indices above 2^31)
    ],
)
def test_step_with_planes_above_4gib_matches_the_tiled_small_grid(policy, reps, with_bcs):
    stepper, s0, s1, s_mask = _small_case(policy, with_bcs, seed=11)
    nx = 2 * reps
    itemsize = s0.element_size()
    assert nx * NY * NZ * itemsize > 2**32, "the case must need 64-bit offsets"
    need = 2 * 19 * nx * NY * NZ * itemsize + (5 * nx * NY * NZ if with_bcs else 0)
    torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info()
    if need > free - (8 << 30):
        pytest.skip(f"needs {need / 1e9:.0f} GB of free HBM, {free / 1e9:.0f} GB free")
    print(f"wide step: policy={policy.name} grid=({nx}, {NY}, {NZ}) bcs={with_bcs}", flush=True)
    omega = 1.7
    steps = 3
    s_start = s0.clone()
    s_out = _run(stepper, s0, s1, _args(stepper, s_mask, stepper._bits, with_bcs), 2, omega, steps)
    torch.cuda.synchronize()
    assert not torch.equal(s_out, s_start), "the small grid's step changed nothing"
    b0 = b1 = b_out = b_mask = b_bits = None
    try:
        if with_bcs:
            b_mask = _tile_x(s_mask, reps)
            b_bits = _tile_x(stepper._bits.view(1, 2, NY, NZ), reps)
        b0 = _tile_x(s_start, reps)
        b1 = torch.empty_like(b0)
        b_out = _run(stepper, b0, b1, _args(stepper, b_mask, b_bits, with_bcs), nx, omega, steps)
        torch.cuda.synchronize()
        want = s_out.view(19, 1, 2, NY, NZ)
        for l in range(19):
            bad = int((b_out[l].view(reps, 2, NY, NZ) != want[l]).sum().item())
            assert bad == 0, f"plane {l}: {bad} values differ from the small grid's step"
    finally:
        b0 = b1 = b_out = b_mask = b_bits = None
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
