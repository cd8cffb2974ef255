"""The x-slab step (xlbh_step_slab through IncompressibleNavierStokesStepper.step_slab): a domain
split into x-slabs, each with one ghost plane per side, stepped slab by slab with the crossing
populations exchanged between steps, equals the whole-domain run bit for bit.

Between steps each slab's lower ghost plane receives only the populations with c_x = +1 of the
lower neighbour's last plane, and its upper ghost plane only those with c_x = -1 of the upper
neighbour's first plane (the reference's sharded exchange moves the same sets,
distribute.py:23-44); everything else in the ghost planes is NaN from the start, so a read of any
other ghost value would show in the comparison.  The exchange wraps around (the last slab's upper
neighbour is the first slab), as the whole-domain pull wraps in x.  The device copies here stand
in for the send/recv between GPUs of a sharded run.
"""

import numpy as np
import pytest
import torch

import xlb_amd as xa
from xlb_amd.workloads import differentiable_channel_2d, lid_driven_cavity_3d, periodic_box_3d, wind_tunnel_sphere

pytestmark = pytest.mark.gpu

P = xa.PrecisionPolicy

BITS = {torch.float64: torch.int64, torch.float32: torch.int32, torch.float16: torch.int16, torch.bfloat16: torch.int16}


def _planes(t, a, b):
    """Planes a - 1 .. b of t along x (dimension 1), wrapped, as a dense tensor."""
    idx = torch.arange(a - 1, b + 1, device=t.device) % t.shape[1]
    return t[:, idx].contiguous()


def _slab_run(wl, bounds, steps):
    """Runs `steps` steps of the workload as x-slabs [a, b) and returns the gathered state."""
    st = wl.stepper
    cx = np.asarray(st.velocity_set.c)[0]
    up = torch.from_numpy(np.nonzero(cx == 1)[0]).cuda()  # read from the lower ghost plane
    dn = torch.from_numpy(np.nonzero(cx == -1)[0]).cuda()  # read from the upper ghost plane
    force = getattr(st.collision, "force_field", None)
    slabs = []
    for a, b in bounds:
        f_in = _planes(wl.f_0, a, b)
        keep_lo, keep_hi = f_in[up, 0].clone(), f_in[dn, -1].clone()
        f_in[:, 0] = float("nan")
        f_in[:, -1] = float("nan")
        f_in[up, 0] = keep_lo
        f_in[dn, -1] = keep_hi
        slabs.append(dict(fi=f_in, fo=torch.full_like(f_in, float("nan")), bm=_planes(wl.bc_mask, a, b),
                          mm=_planes(wl.missing_mask, a, b), g=None if force is None else _planes(force, a, b)))
    for t in range(steps):
        for s in slabs:
            if force is not None:
                st.collision.force_field = s["g"]
            st.step_slab(s["fi"], s["fo"], s["bm"], s["mm"], wl.omega, t, 1, s["fo"].shape[1] - 1)
        for i, s in enumerate(slabs):
            lo, hi = slabs[i - 1], slabs[(i + 1) % len(slabs)]
            s["fo"][up, 0] = lo["fo"][up, -2]
            s["fo"][dn, -1] = hi["fo"][dn, 1]
        for s in slabs:
            s["fi"], s["fo"] = s["fo"], s["fi"]
    if force is not None:
        st.collision.force_field = force
    return torch.cat([s["fi"][:, 1:-1] for s in slabs], dim=1)


CASES = {
    # (workload builder, slab bounds along x, steps)
    "periodic_d3q19_fp32": (lambda: periodic_box_3d(24, P.FP32FP32, omega=1.7), [(0, 7), (7, 16), (16, 24)], 30),
    "periodic_d3q19_fp64": (lambda: periodic_box_3d(16, P.FP64FP64, omega=1.7), [(0, 5), (5, 16)], 20),
    "cavity_d3q19_fp32": (lambda: lid_driven_cavity_3d(20, P.FP32FP32, lid_u=0.05), [(0, 9), (9, 20)], 40),
    "cavity_d3q19_fp16_shifted": (lambda: lid_driven_cavity_3d(20, P.FP32FP16_SHIFTED, lid_u=0.05),
                                  [(0, 4), (4, 13), (13, 20)], 40),
    "cavity_d3q19_smagorinsky": (lambda: lid_driven_cavity_3d(18, P.FP32FP32, lid_u=0.06, omega=1.6,
                                                              collision="SmagorinskyLESBGK"), [(0, 10), (10, 18)], 30),
    # the sphere (x 10..18) straddles the boundary at x = 12; regularized inlet, extrapolation outlet
    "windtunnel_d3q27_kbc": (lambda: wind_tunnel_sphere((40, 20, 20), P.FP32FP32, 4, (14, 10, 10)),
                             [(0, 12), (12, 25), (25, 40)], 30),
    # D2Q9 with a per-cell force field (sliced per slab) and a parabolic regularized inlet
    "channel_d2q9_force": (lambda: differentiable_channel_2d((48, 24), P.FP32FP32, force_std=1e-5),
                           [(0, 20), (20, 48)], 40),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_slab_steps_match_whole_domain(case):
    build, bounds, steps = CASES[case]
    wl = build()
    assert bounds[0][0] == 0 and bounds[-1][1] == wl.f_0.shape[1]
    assert all(b - a >= 3 for a, b in bounds)
    f0, f1 = wl.f_0.clone(), wl.f_1.clone()
    slab = _slab_run(wl, bounds, steps)
    full, _ = wl.stepper.run(f0, f1, wl.bc_mask, wl.missing_mask, wl.omega, steps)
    full = full.contiguous()
    assert bool(torch.isfinite(slab.float()).all()), "a ghost value other than the crossing populations was read"
    bits = BITS[full.dtype]
    same = slab.view(bits) == full.view(bits)
    print(f"{case}: {len(bounds)} slabs {bounds}, {steps} steps: {int((~same).sum())} of {same.numel()} values differ")
    assert bool(same.all())


def test_edge_planes_then_interior():
    """Updating a slab's two edge planes first and its interior afterwards (the order that lets
    the edge planes' exchange overlap the interior's update) gives the same state as one call."""
    wl = lid_driven_cavity_3d(20, P.FP32FP32, lid_u=0.05)
    st = wl.stepper
    a, b = 0, 20
    fi = _planes(wl.f_0, a, b)
    bm, mm = _planes(wl.bc_mask, a, b), _planes(wl.missing_mask, a, b)
    one = torch.zeros_like(fi)
    split = torch.zeros_like(fi)
    n = fi.shape[1]
    st.step_slab(fi, one, bm, mm, wl.omega, 0, 1, n - 1)
    st.step_slab(fi, split, bm, mm, wl.omega, 0, 1, 2)
    st.step_slab(fi, split, bm, mm, wl.omega, 0, n - 2, n - 1)
    st.step_slab(fi, split, bm, mm, wl.omega, 0, 2, n - 2)
    assert torch.equal(one.view(torch.int32), split.view(torch.int32))
    # the ghost planes stay unwritten
    assert bool((one[:, 0] == 0).all()) and bool((one[:, -1] == 0).all())
