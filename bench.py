"""Benchmark of the fused LBM step on one MI355X.

Headline workload (the reference's own benchmark, examples/performance/mlups_3d.py:45-90, at
edge 256): D3Q19 BGK lid-driven cavity 256^3, rho = 1 / u = 0 equilibrium start, lid
EquilibriumBC(rho=1, u=(0.02, 0, 0)) on the top face without edges, FullwayBounceBackBC on the
other five full faces, omega = 1.0.  One step = one fused stream-collide-BC pass over the grid;
the timed loop runs K steps with the buffer swap below the C ABI (xlbh_run), all inputs resident
in HBM.  value = MLUPS = 256^3 K / t / 1e6 with the FP32/FP32 policy.

Also reported on the same line:
* store16 -- the same workload with 16-bit population storage (plain FP16, and f - w_l shifted
  FP16 / BF16), FP32 arithmetic;
* precision -- periodic D3Q19 BGK Taylor-Green vortex 512x256x256, u0 = 0.04, nu = 0.01, 2000
  steps: relative L2 velocity error and kinetic-energy decay of every storage format against an
  FP64/FP64 run of the same kernel (itself matched to the FP64 CPU oracle to 1e-12 by the tests);
* adjoint -- backward (checkpoint recompute + adjoint kernels) over forward time per step for the
  differentiable D2Q9 channel 512x256 (800 steps, checkpoint every 50), the 256^3 cavity and the
  128^3 decaying turbulence (2000 steps, BF16 checkpoints every 45);
* kbc_sphere -- D3Q27 KBC wind tunnel 256x128x128, sphere r = 12, 5000 steps with the
  momentum-exchange drag every 100 steps inside the timed loop, and the drag coefficient at the
  end; grads_sphere -- the same with a GradsApproximationBC sphere;
* mesh_masker -- surface voxelisation of an 81,920-triangle sphere mesh into the same grid;
* les_cavity -- the LES lid-driven cavity of examples/out_of_core/lid_driven_cavity.py (D3Q19
  SmagorinskyLESBGK, Cs = 0.17, 256^3, lid u = 0.06, tau = 0.501, FP32), with the Q-criterion
  of its velocity field (examples' RenderQCriterionSubroutine) timed on the final state.
* slab -- one GPU's share of an x-sharded run: the 256^3 periodic D3Q19 BGK FP32 data as a slab
  of 256 planes plus two ghost planes, stepped with xlbh_step_slab at dense and at the library's
  recommended plane stride, beside the whole-domain step; edge planes, interior and ghost fill.
roofline.achieved = algorithmic bytes per launch (2 Q S_store per cell) / average kernel duration
from HIP events on the launch stream; roofline.traffic = FETCH_SIZE + WRITE_SIZE per launch from
rocprofv3 (profiles/traffic_step_d3q19_cavity256.json, written by --summarize-pmc).
cpu_baseline = the C++/OpenMP oracle (the parity reference of the tests, a restatement of the
reference algorithm) timed on a bounded number of steps of the headline workload on the host cores.

Usage:
  python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 256] [--no-cpu] [--quick]
  python bench.py --profile [--steps K]        # headline kernel + a 1 GiB copy, for rocprofv3
  python bench.py --summarize-pmc FETCH.csv WRITE.csv   # -> profiles/traffic_step_d3q19_cavity256.json
"""

import argparse
import csv
import json
import re
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS & % HBM roofline, D3Q19 BGK 256³ (FP32 vs FP16-store); adjoint/fwd time"
HBM_PEAK_GBS = 8000.0
COPY_CALIBRATION_BYTES = 1 << 30
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "traffic_step_d3q19_cavity{n}.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def store_bytes(torch, t):
    return torch.empty((), dtype=t.dtype).element_size()


def time_run(torch, wl, steps, warmup):
    """Warm up, then time ``steps`` steps of ``stepper.run`` with HIP events on the launch stream
    and the host clock; returns (wall_s, event_s, f_0, f_1)."""
    st = wl.stepper
    f0, f1 = st.run(wl.f_0, wl.f_1, wl.bc_mask, wl.missing_mask, wl.omega, warmup)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    f0, f1 = st.run(f0, f1, wl.bc_mask, wl.missing_mask, wl.omega, steps, timestep=warmup)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, ev0.elapsed_time(ev1) / 1e3, f0, f1


def headline_leg(xa, torch, policy, n, steps, warmup, python_loop=True):
    from xlb_amd.workloads import lid_driven_cavity_3d

    wl = lid_driven_cavity_3d(n, policy)
    wall, ev_s, f0, f1 = time_run(torch, wl, steps, warmup)
    cells = n**3
    res = dict(policy=policy.name, mlups=cells * steps / wall / 1e6, mlups_events=cells * steps / ev_s / 1e6,
               ms_per_step=wall / steps * 1e3)
    if python_loop:
        # the reference's loop through the public __call__ (mlups_3d.py:78-80)
        py_steps = min(steps, 200)
        st = wl.stepper
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for i in range(py_steps):
            f0, f1 = st(f0, f1, wl.bc_mask, wl.missing_mask, wl.omega, i)
            f0, f1 = f1, f0
        torch.cuda.synchronize()
        res["mlups_python_loop"] = cells * py_steps / (time.perf_counter() - t1) / 1e6
    res["finite"] = bool(torch.isfinite(f0.float()).all().item())
    bpl = 2 * wl.stepper.velocity_set.q * store_bytes(torch, f0)
    kernel_s = ev_s / steps
    res.update(bytes_per_lu=bpl, kernel_ms=kernel_s * 1e3, achieved_gbs=bpl * cells / kernel_s / 1e9)
    del wl, f0, f1
    torch.cuda.empty_cache()
    return res


def decoded_velocity(torch, f, policy, vs, chunk=32):
    """Velocity (d, ...) in float64 of stored populations, by explicit sums over the directions in
    x-chunks (no BLAS call: torch's tensordot on fields of this size, or on strided views into
    them, returned wrong moments on ROCm)."""
    import xlb_amd as xa

    c = np.asarray(vs.c, dtype=np.float64)
    w = np.asarray(vs.w, dtype=np.float64)
    u = torch.empty((vs.d,) + tuple(f.shape[1:]), dtype=torch.float64, device=f.device)
    dec = xa.grid.decode_populations(f, vs.w, block=True) if policy.block else None
    for i in range(0, f.shape[1], chunk):
        rho = torch.zeros_like(u[0, i:i + chunk])
        mom = torch.zeros_like(u[:, i:i + chunk])
        for l in range(vs.q):
            fl = dec[l, i:i + chunk] if dec is not None else f[l, i:i + chunk].double()
            if policy.shifted:
                fl = fl + w[l]
            rho += fl
            for d in range(vs.d):
                if c[d, l] != 0:
                    mom[d] += c[d, l] * fl
        u[:, i:i + chunk] = mom / rho
    return u


def precision_leg(xa, torch, shape=(512, 256, 256), steps=2000):
    """Taylor-Green vortex: every storage format against the FP64/FP64 run."""
    from xlb_amd.workloads import taylor_green_3d

    P = xa.PrecisionPolicy
    out, u64, ke0 = {}, None, None
    cells = int(np.prod(shape))
    for pol in (P.FP64FP64, P.FP32FP32, P.FP32FP32_DEVIATION, P.FP32I16_BLOCK, P.FP32FP16_SHIFTED,
                P.FP32BF16_SHIFTED, P.FP32FP16, P.FP32BF16):
        wl = taylor_green_3d(shape, pol)
        vs = wl.stepper.velocity_set
        if ke0 is None:
            ke0 = float(0.5 * decoded_velocity(torch, wl.f_0, pol, vs).pow(2).sum(0).mean())
        wall, ev_s, f0, f1 = time_run(torch, wl, steps, 0)
        u = decoded_velocity(torch, f0, pol, vs)
        ke = float(0.5 * u.pow(2).sum(0).mean())
        rec = dict(mlups=cells * steps / ev_s / 1e6, ke_ratio=ke / ke0)
        if u64 is None:
            u64 = u
        else:
            rec["rel_l2_u_vs_fp64"] = float((u - u64).norm() / u64.norm())
            rec["max_abs_u_vs_fp64"] = float((u - u64).abs().max())
        out[pol.name] = rec
        log(f"bench: precision {pol.name}: " + ", ".join(f"{k}={v:.3e}" for k, v in rec.items()))
        del wl, f0, f1, u
        torch.cuda.empty_cache()
    return dict(workload=f"periodic D3Q19 BGK Taylor-Green vortex {shape[0]}x{shape[1]}x{shape[2]}, u0=0.04, "
                         f"nu=0.01, {steps} steps; errors against the FP64/FP64 run", formats=out)


def precision_viscosity_sweep(xa, torch, shape=(512, 256, 256), steps=2000, nus=(0.003, 0.03, 0.1)):
    """The C3 Taylor-Green vortex at other viscosities: relative L2 velocity error of the shifted
    16-bit formats against FP64, to show how the storage error depends on the relaxation time."""
    from xlb_amd.workloads import taylor_green_3d

    P = xa.PrecisionPolicy
    res = {}
    for nu in nus:
        row = {"tau": 3 * nu + 0.5}
        u64 = None
        for pol in (P.FP64FP64, P.FP32I16_BLOCK, P.FP32FP16_SHIFTED, P.FP32BF16_SHIFTED):
            wl = taylor_green_3d(shape, pol, nu=nu)
            vs = wl.stepper.velocity_set
            _, _, f0, f1 = time_run(torch, wl, steps, 0)
            u = decoded_velocity(torch, f0, pol, vs)
            if u64 is None:
                u64 = u
                row["ke_rms_u"] = float(u.pow(2).sum(0).mean().sqrt())
            else:
                row[pol.name] = float((u - u64).norm() / u64.norm())
            del wl, f0, f1, u
            torch.cuda.empty_cache()
        res[f"nu={nu}"] = row
        log(f"bench: precision sweep nu={nu}: " + ", ".join(f"{k}={v:.3e}" for k, v in row.items()))
        del u64
        torch.cuda.empty_cache()
    return dict(workload=f"periodic D3Q19 BGK Taylor-Green vortex {shape[0]}x{shape[1]}x{shape[2]}, u0=0.04, "
                         f"{steps} steps, viscosities {list(nus)}; shifted 16-bit formats against FP64", rows=res)


def headline_precision_leg(xa, torch, n=256, steps=1000):
    """Velocity error of every storage format on the headline cavity after ``steps`` steps,
    against the FP64/FP64 run of the same workload."""
    from xlb_amd.workloads import lid_driven_cavity_3d

    P = xa.PrecisionPolicy
    out, u64 = {}, None
    for pol in (P.FP64FP64, P.FP32FP32, P.FP32FP32_DEVIATION, P.FP32I16_BLOCK, P.FP32FP16_SHIFTED,
                P.FP32BF16_SHIFTED, P.FP32FP16, P.FP32BF16):
        wl = lid_driven_cavity_3d(n, pol)
        f0, _ = wl.stepper.run(wl.f_0, wl.f_1, wl.bc_mask, wl.missing_mask, wl.omega, steps)
        u = decoded_velocity(torch, f0, pol, wl.stepper.velocity_set)
        if u64 is None:
            u64 = u
            out[pol.name] = dict(max_abs_u=float(u.abs().max()))
        else:
            out[pol.name] = dict(rel_l2_u_vs_fp64=float((u - u64).norm() / u64.norm()),
                                 max_abs_u_vs_fp64=float((u - u64).abs().max()))
        log(f"bench: cavity precision {pol.name}: " + ", ".join(f"{k}={v:.3e}" for k, v in out[pol.name].items()))
        del wl, f0, u
        torch.cuda.empty_cache()
    return dict(workload=f"headline cavity D3Q19 BGK {n}^3 (lid u=0.02, omega=1.0), {steps} steps; errors against "
                         "the FP64/FP64 run", formats=out)


def _sync_time(torch, fn):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0, r


def adjoint_leg(xa, torch, quick=False):
    """Backward / forward time per step of the checkpointed adjoint rollout, best of three runs
    each.  The forward reference is the plain forward (``stepper.run``, the step loop below the
    ABI); the backward is the whole reverse pass (checkpoint recompute + adjoint kernels)."""
    from xlb_amd.autodiff import macroscopic_torch, rollout
    from xlb_amd.workloads import differentiable_channel_2d, lid_driven_cavity_3d

    def measure(wl, steps, every, loss, grad_wrt_force):
        st = wl.stepper
        a0, b0 = wl.f_0.clone(), wl.f_1.clone()
        st.run(a0, b0, wl.bc_mask, wl.missing_mask, wl.omega, steps)  # warm-up
        t_run = float("inf")
        for _ in range(3):
            a0.copy_(wl.f_0)
            t, _ = _sync_time(torch, lambda: st.run(a0, b0, wl.bc_mask, wl.missing_mask, wl.omega, steps))
            t_run = min(t_run, t)
        def inputs():
            if grad_wrt_force:
                x = wl.extra["force_field"].clone().requires_grad_(True)
                return x, wl.f_0, dict(force_field=x)
            x = wl.f_0.clone().requires_grad_(True)
            return x, x, {}

        # untimed warm-up of the reverse pass (first-launch loading and workspace allocation)
        x, f0, args = inputs()
        loss(rollout(st, f0, wl.bc_mask, wl.missing_mask, wl.omega, min(steps, 2 * every), checkpoint_every=every,
                     **args)).backward()
        t_roll = t_bwd = float("inf")
        for _ in range(3):
            x, f0, args = inputs()
            t, fT = _sync_time(torch, lambda: rollout(st, f0, wl.bc_mask, wl.missing_mask, wl.omega, steps,
                                                       checkpoint_every=every, **args))
            t_roll = min(t_roll, t)
            L = loss(fT)
            torch.cuda.synchronize()
            t, _ = _sync_time(torch, lambda: L.backward())
            t_bwd = min(t_bwd, t)
            del fT, L
        return dict(fwd_ms_per_step=t_run / steps * 1e3, rollout_fwd_ms_per_step=t_roll / steps * 1e3,
                    bwd_ms_per_step=t_bwd / steps * 1e3, ratio_bwd_over_fwd=t_bwd / t_run,
                    ratio_fwd_plus_bwd_over_fwd=(t_roll + t_bwd) / t_run, grad_norm=float(x.grad.norm()))

    res = {}
    # differentiable channel: loss = mean squared deviation of u at x = 480 from a seeded target
    steps, every = (200, 50) if quick else (800, 50)
    wl = differentiable_channel_2d((512, 256), xa.PrecisionPolicy.FP32FP32, seed=0)
    vs = wl.stepper.velocity_set
    rng = np.random.default_rng(1)
    yy = np.arange(256, dtype=np.float64)
    target = np.zeros((2, 256))
    target[0] = 4 * 0.05 * yy * (255 - yy) / 255**2 * (1 + 0.05 * rng.standard_normal(256))
    target = torch.from_numpy(target).cuda()

    def channel_loss(fT):
        _, u = macroscopic_torch(fT[:, 480:481].double(), vs)
        return ((u[:, 0, :, 0] - target) ** 2).mean()

    r = measure(wl, steps, every, channel_loss, True)
    r["workload"] = (f"D2Q9 BGK channel 512x256, regularized parabolic inlet u=0.05, extrapolation outlet, halfway "
                     f"walls, per-cell force field, {steps} steps, checkpoint every {every}, FP32; gradient w.r.t. the "
                     "force field")
    res["channel2d"] = r
    log("bench: adjoint channel2d: " + ", ".join(f"{k}={v:.4g}" for k, v in r.items() if k != "workload"))
    del wl
    torch.cuda.empty_cache()

    # headline cavity: loss on one z-slice, sqrt-N checkpoints
    n, steps = (128, 32) if quick else (256, 64)
    wl = lid_driven_cavity_3d(n, xa.PrecisionPolicy.FP32FP32)
    wl.f_0, wl.f_1 = wl.stepper.run(wl.f_0, wl.f_1, wl.bc_mask, wl.missing_mask, wl.omega, 20)
    every = int(np.ceil(np.sqrt(steps)))
    r = measure(wl, steps, every, lambda fT: (fT[:, :, :, n // 2].double() ** 2).sum(), False)
    r["workload"] = (f"D3Q19 BGK cavity {n}^3 (headline setup), {steps} steps, checkpoint every {every}, FP32; "
                     "gradient w.r.t. the initial populations")
    res["cavity3d"] = r
    log("bench: adjoint cavity3d: " + ", ".join(f"{k}={v:.4g}" for k, v in r.items() if k != "workload"))
    del wl
    torch.cuda.empty_cache()

    res["turbulence3d"] = turbulence_adjoint(xa, torch, quick)
    return res


def turbulence_adjoint(xa, torch, quick=False):
    """C5: gradient of the band-limited kinetic energy after 2000 steps of 128^3 decaying turbulence
    with respect to the initial velocity (through the equilibrium initialisation), FP32 with BF16
    checkpoints every sqrt(N) steps.  Backward / forward time per step, best of three."""
    from xlb_amd.autodiff import equilibrium_torch, macroscopic_torch, rollout
    from xlb_amd.workloads import band_energy, decaying_turbulence_3d, turbulence_velocity

    n, steps = (64, 400) if quick else (128, 2000)
    wl = decaying_turbulence_3d(n, xa.PrecisionPolicy.FP32FP32, omega=1.95)
    st, vs = wl.stepper, wl.stepper.velocity_set
    u_init = turbulence_velocity(n, seed=3)
    rho = torch.ones((1, n, n, n), dtype=torch.float64, device="cuda")

    def start(u):
        return equilibrium_torch(rho, u, vs).to(st.store_dtype)

    def loss(fT):
        _, u = macroscopic_torch(fT.double(), vs)
        return band_energy(u)

    every = int(np.ceil(np.sqrt(steps)))
    f0 = start(u_init).contiguous()
    f1 = torch.empty_like(f0)
    st.run(f0.clone(), f1, wl.bc_mask, wl.missing_mask, wl.omega, steps)  # warm-up
    t_run = float("inf")
    for _ in range(3):
        a = f0.clone()
        t, _ = _sync_time(torch, lambda: st.run(a, f1, wl.bc_mask, wl.missing_mask, wl.omega, steps))
        t_run = min(t_run, t)
    u0 = u_init.clone().requires_grad_(True)
    loss(rollout(st, start(u0), wl.bc_mask, wl.missing_mask, wl.omega, 2 * every, checkpoint_every=every,
                 checkpoint_dtype="bf16")).backward()  # untimed warm-up of the reverse pass
    t_roll = t_bwd = float("inf")
    for _ in range(3):
        u0 = u_init.clone().requires_grad_(True)
        t, fT = _sync_time(torch, lambda: rollout(st, start(u0), wl.bc_mask, wl.missing_mask, wl.omega, steps,
                                                   checkpoint_every=every, checkpoint_dtype="bf16"))
        t_roll = min(t_roll, t)
        L = loss(fT)
        torch.cuda.synchronize()
        t, _ = _sync_time(torch, lambda: L.backward())
        t_bwd = min(t_bwd, t)
        del fT, L
    r = dict(fwd_ms_per_step=t_run / steps * 1e3, rollout_fwd_ms_per_step=t_roll / steps * 1e3,
             bwd_ms_per_step=t_bwd / steps * 1e3, ratio_bwd_over_fwd=t_bwd / t_run,
             ratio_fwd_plus_bwd_over_fwd=(t_roll + t_bwd) / t_run, grad_norm=float(u0.grad.norm()),
             workload=(f"D3Q19 BGK decaying turbulence {n}^3 periodic, k^-5/3 band 2..8, rms 0.03, {steps} steps, "
                       f"BF16 checkpoints every {every}, FP32; gradient of the 4..8 band energy w.r.t. the initial "
                       "velocity"))
    log("bench: adjoint turbulence3d: " + ", ".join(f"{k}={v:.4g}" for k, v in r.items() if k != "workload"))
    del wl, f0, f1
    torch.cuda.empty_cache()
    return r


def kbc_sphere_leg(xa, torch, steps=5000, quick=False, sphere_bc="halfway"):
    """D3Q27 KBC wind tunnel with the drag every 100 steps inside the timed loop; the sphere is a
    halfway bounce-back or a Grad's-approximation wall."""
    from xlb_amd.workloads import wind_tunnel_sphere

    shape = (128, 64, 64) if quick else (256, 128, 128)
    radius, center = (6, (32, 32, 32)) if quick else (12, (64, 64, 64))
    if quick:
        steps = min(steps, 500)
    wl = wind_tunnel_sphere(shape, xa.PrecisionPolicy.FP32FP32, radius, center, sphere_bc=sphere_bc)
    st = wl.stepper
    mt = xa.operator.force.MomentumTransfer(wl.extra["sphere_bc"])
    f0, f1 = st.run(wl.f_0, wl.f_1, wl.bc_mask, wl.missing_mask, wl.omega, 100)
    torch.cuda.synchronize()
    drags = []
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(100, steps - done)
        f0, f1 = st.run(f0, f1, wl.bc_mask, wl.missing_mask, wl.omega, k, timestep=100 + done)
        drags.append(mt.force_tensor(f0, f1, wl.bc_mask, wl.missing_mask))
        done += k
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = int(np.prod(shape))
    bpl = 2 * 27 * 4
    mlups = cells * steps / wall / 1e6
    wall_name = "halfway bounce-back" if sphere_bc == "halfway" else "GradsApproximationBC"
    res = dict(workload=f"D3Q27 KBC wind tunnel {shape[0]}x{shape[1]}x{shape[2]}, {wall_name} sphere r={radius}, "
                        f"Re=200, {steps} steps with momentum-exchange drag every 100, FP32",
               mlups=mlups, achieved_gbs=mlups * 1e6 * bpl / 1e9, frac=mlups * 1e6 * bpl / 1e9 / HBM_PEAK_GBS,
               drag_last=[float(v) for v in drags[-1].cpu()], finite=bool(torch.isfinite(f0).all().item()))
    res["cd_last"] = 2.0 * res["drag_last"][0] / (0.04**2 * np.pi * radius**2)
    log(f"bench: kbc sphere ({sphere_bc}) {mlups:.0f} MLUPS ({res['achieved_gbs']:.0f} GB/s), drag {res['drag_last']}")
    del wl, st, f0, f1
    torch.cuda.empty_cache()
    return res


def les_cavity_leg(xa, torch, n, steps, warmup):
    """SmagorinskyLESBGK lid-driven cavity (out_of_core/lid_driven_cavity.py:38-47 defaults) and
    the Q-criterion post-processing of its final velocity field."""
    from xlb_amd.workloads import lid_driven_cavity_3d

    wl = lid_driven_cavity_3d(n, xa.PrecisionPolicy.FP32FP32, lid_u=0.06, omega=1.0 / 0.501,
                              collision="SmagorinskyLESBGK")
    wall, ev_s, f0, f1 = time_run(torch, wl, steps, warmup)
    cells = n**3
    res = dict(workload=f"D3Q19 SmagorinskyLESBGK (Cs 0.17) lid-driven cavity {n}^3, lid u = 0.06, tau = 0.501, "
                        f"FP32, {steps} steps", mlups=cells * steps / wall / 1e6, kernel_ms=ev_s / steps * 1e3,
               achieved_gbs=152 * cells / (ev_s / steps) / 1e9)
    res["frac"] = res["achieved_gbs"] / HBM_PEAK_GBS
    res["finite"] = bool(torch.isfinite(f0).all().item())
    g = wl.stepper.grid
    rho, u = g.create_field(1), g.create_field(3)
    xa.operator.macroscopic.Macroscopic()(f0, rho, u)
    norm_mu, qf = g.create_field(1), g.create_field(1)
    qc = xa.operator.postprocess.QCriterion()
    qc(u, wl.bc_mask, norm_mu, qf)
    reps = 20
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
        qc(u, wl.bc_mask, norm_mu, qf)
    ev1.record(stream)
    torch.cuda.synchronize()
    q_s = ev0.elapsed_time(ev1) / 1e3 / reps
    # algorithmic bytes per cell: 3 velocity components + bc_mask in, |curl u| and Q out
    res.update(q_criterion_ms=q_s * 1e3, q_criterion_gbs=(3 * 4 + 1 + 2 * 4) * cells / q_s / 1e9,
               max_u=float(u.abs().max().item()), max_q=float(qf.max().item()))
    log(f"bench: les cavity {res['mlups']:.0f} MLUPS ({res['achieved_gbs']:.0f} GB/s), "
        f"q-criterion {res['q_criterion_ms']:.3f} ms ({res['q_criterion_gbs']:.0f} GB/s)")
    del wl, f0, f1, rho, u, norm_mu, qf
    torch.cuda.empty_cache()
    return res


def periodic_leg(xa, torch, n, steps, warmup):
    """The pure step kernel: D3Q19 BGK FP32 periodic box without boundary conditions."""
    from xlb_amd.workloads import periodic_box_3d

    wl = periodic_box_3d(n, xa.PrecisionPolicy.FP32FP32)
    wall, ev_s, f0, f1 = time_run(torch, wl, steps, warmup)
    cells = n**3
    res = dict(workload=f"D3Q19 BGK FP32 periodic {n}^3, no boundary conditions", mlups=cells * steps / wall / 1e6,
               kernel_ms=ev_s / steps * 1e3, achieved_gbs=152 * cells / (ev_s / steps) / 1e9)
    res["frac"] = res["achieved_gbs"] / HBM_PEAK_GBS
    del wl, f0, f1
    torch.cuda.empty_cache()
    return res


def slab_leg(xa, torch, n=256, reps=50):
    """One GPU's share of an x-sharded run: an x-slab of n planes of n x n cells plus a ghost
    plane on each side, stepped with xlbh_step_slab (D3Q19 BGK FP32, periodic box data), beside
    the whole-domain n^3 step.  The slab's populations are timed with planes of the slab's own
    size (dense) and with the library's recommended plane stride (xlbh_plane_stride, which pads a
    plane just above a power of two); then, at the recommended stride, the two edge planes and
    the interior as separate calls, and the ghost fill: the c_x = +1 / -1 populations of one
    plane each (the data a sharded run exchanges per step and face)."""
    from xlb_amd.workloads import periodic_box_3d

    wl = periodic_box_3d(n, xa.PrecisionPolicy.FP32FP32, omega=1.7)
    st = wl.stepper
    stream = torch.cuda.current_stream()

    def timed(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e3

    res = dict(workload=f"D3Q19 BGK FP32, x-slab of {n} planes + 2 ghost planes ({n + 2}x{n}x{n}) vs the {n}^3 "
                        "whole-domain step (periodic box data)")
    res["whole_domain_us"] = timed(lambda: st(wl.f_0, wl.f_1, wl.bc_mask, wl.missing_mask, wl.omega, 0))
    idx = torch.arange(-1, n + 1, device="cuda") % n
    f_src = wl.f_0[:, idx].contiguous()
    bm, mm = wl.bc_mask[:, idx].contiguous(), wl.missing_mask[:, idx].contiguous()
    del wl
    torch.cuda.empty_cache()
    plane_cells = n * n
    from xlb_amd import _lib

    recommended = int(_lib.lib().xlbh_plane_stride(_lib.D3Q19, _lib.F32, n + 2, n, n))
    for label, stride in (("dense", (n + 2) * plane_cells), ("recommended", recommended)):
        bufs = [torch.empty(19 * stride, dtype=torch.float32, device="cuda") for _ in range(2)]
        a, b = [u.as_strided(tuple(f_src.shape), (stride, plane_cells, n, 1)) for u in bufs]
        a.copy_(f_src)
        b.copy_(f_src)
        us = timed(lambda: st.step_slab(a, b, bm, mm, 1.7, 0, 1, n + 1))
        res[f"slab_{label}_us"] = us
        res[f"slab_{label}_plane_mib"] = stride * 4 / 2**20
        res[f"slab_{label}_mlups"] = n**3 / us
        if label == "recommended":
            res["edge_planes_us"] = timed(lambda: (st.step_slab(a, b, bm, mm, 1.7, 0, 1, 2),
                                                   st.step_slab(a, b, bm, mm, 1.7, 0, n, n + 1)))
            res["interior_us"] = timed(lambda: st.step_slab(a, b, bm, mm, 1.7, 0, 2, n))
            cx = np.asarray(st.velocity_set.c)[0]
            up = [int(l) for l in np.nonzero(cx == 1)[0]]
            dn = [int(l) for l in np.nonzero(cx == -1)[0]]

            def fill():
                for l in up:
                    b[l, 0].copy_(a[l, n])
                for l in dn:
                    b[l, n + 1].copy_(a[l, 1])

            res["ghost_fill_us"] = timed(fill)
            res["ghost_fill_bytes"] = (len(up) + len(dn)) * plane_cells * 4
        del a, b, bufs
        torch.cuda.empty_cache()
    log(f"bench: slab share {res['slab_recommended_us']:.1f} us (planes {res['slab_recommended_plane_mib']:.1f} MiB), "
        f"{res['slab_dense_us']:.1f} us (dense), whole domain {res['whole_domain_us']:.1f} us")
    return res


def copy_leg(torch, nbytes=2 * 19 * 4 * 256**3 // 2, reps=20):
    """Device-to-device copy rate (torch copy_ of one 1.27 GB buffer into another): the practical
    streaming ceiling the step kernel is compared with."""
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        b.copy_(a)
    ev1.record()
    torch.cuda.synchronize()
    s = ev0.elapsed_time(ev1) / 1e3 / reps
    del a, b
    torch.cuda.empty_cache()
    return dict(bytes=2 * nbytes, gbs=2 * nbytes / s / 1e9)


def pull_pattern_leg(xa, torch, policy, n=256, reps=50):
    """The step kernel's memory pattern alone: the Stream operator (the same Q pulled loads, Q
    non-temporal stores, addressing, launch shape and occupancy as the fused step, no arithmetic)
    on the headline cavity's fields.  Its rate is the ceiling the step's access pattern allows."""
    from xlb_amd.workloads import lid_driven_cavity_3d

    wl = lid_driven_cavity_3d(n, policy)
    stream = xa.operator.stream.Stream(wl.stepper.velocity_set, policy, wl.stepper.compute_backend)
    a, b = wl.f_0, wl.f_1
    for _ in range(5):
        stream(a, b)
        a, b = b, a
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        stream(a, b)
        a, b = b, a
    ev1.record()
    torch.cuda.synchronize()
    s = ev0.elapsed_time(ev1) / 1e3 / reps
    bpl = 2 * wl.stepper.velocity_set.q * store_bytes(torch, a)
    del wl, a, b
    torch.cuda.empty_cache()
    return dict(policy=policy.name, kernel_ms=s * 1e3, gbs=bpl * n**3 / s / 1e9, mlups=n**3 / s / 1e6)


def c1_leg(xa, torch, with_cpu=True):
    """C1: 64^3 lid-driven cavity, halfway walls, lid u = (0.05, 0, 0), Re = 100, 1000 steps, FP32:
    GPU rate, and the oracle's rate on the host cores for the same 1000 steps."""
    from xlb_amd.workloads import lid_driven_cavity_3d

    n, steps = 64, 1000
    omega = 1.0 / (3 * 0.05 * n / 100 + 0.5)
    wl = lid_driven_cavity_3d(n, xa.PrecisionPolicy.FP32FP32, lid_u=0.05, omega=omega, walls="halfway")
    wall, ev_s, f0, f1 = time_run(torch, wl, steps, 20)
    res = dict(workload="lid-driven cavity D3Q19 BGK 64^3, halfway walls, lid u=0.05, Re=100, 1000 steps, FP32",
               gpu_mlups=n**3 * steps / wall / 1e6)
    if with_cpu:
        from oracle import oracle as orc

        threads = len(os.sched_getaffinity(0))
        orc.set_threads(threads)
        mask = wl.bc_mask.cpu().numpy()
        miss = wl.missing_mask.cpu().numpy().astype(np.uint8)
        bcs = [dict(id=int(bc.id), kind=int(bc.kind), rho=float(getattr(bc, "rho", 0.0)),
                    u=tuple(float(v) for v in getattr(bc, "u", (0.0, 0.0, 0.0)))) for bc in wl.stepper.boundary_conditions]
        a = np.zeros((19, n, n, n), np.float32)
        orc.equilibrium(1, orc.F32, 0, np.ones((n, n, n)), np.zeros((3, n, n, n)), a)
        b = a.copy()
        t0 = time.perf_counter()
        for _ in range(steps):
            orc.step(1, 0, orc.F32, 0, a, b, mask, miss, bcs, omega)
            a, b = b, a
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = dict(value=n**3 * steps / dt / 1e6, unit="MLUPS", cores=threads, kind="port",
                                   sample="all 1000 steps of C1, C++/OpenMP oracle, FP32")
        res["gpu_over_cpu"] = res["gpu_mlups"] / res["cpu_baseline"]["value"]
    del wl, f0, f1
    torch.cuda.empty_cache()
    return res


def mesh_leg(xa, torch, shape=(256, 128, 128), subdivisions=6, reps=5):
    """MeshBoundaryMasker: surface voxelisation of an icosphere (20 * 4^k triangles, radius 30) into
    the C2 wind-tunnel grid, best of `reps` (HIP events)."""
    from xlb_amd.workloads import _setup, icosphere

    vs, grid = _setup("D3Q19", xa.PrecisionPolicy.FP32FP32, shape)
    tris = icosphere((shape[0] / 4, shape[1] / 2, shape[2] / 2), 30.0, subdivisions)
    masker = xa.operator.boundary_masker.MeshBoundaryMasker(vs, xa.PrecisionPolicy.FP32FP32, xa.ComputeBackend.MI355X)
    best = float("inf")
    for _ in range(reps):
        bc = xa.operator.boundary_condition.HalfwayBounceBackBC(mesh_vertices=tris.reshape(-1, 3))
        bc_mask = torch.zeros((1,) + shape, dtype=torch.uint8, device="cuda")
        missing = torch.zeros((vs.q,) + shape, dtype=torch.bool, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        masker(bc, bc_mask, missing)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    cells = int(np.prod(shape))
    res = dict(workload=f"icosphere r=30, {tris.shape[0]} triangles, voxelised into {shape[0]}x{shape[1]}x{shape[2]} "
                        "(D3Q19 masks), host call to masks ready",
               ms=best * 1e3, solid_voxels=int((bc_mask == 255).sum().item()),
               boundary_voxels=int((bc_mask == bc.id).sum().item()), mcells_per_s=cells / best / 1e6)
    log(f"bench: mesh masker {res['ms']:.2f} ms for {tris.shape[0]} triangles")
    del bc_mask, missing
    torch.cuda.empty_cache()
    return res


def cpu_leg(n, budget_s=12.0, max_steps=40):
    """The C++/OpenMP oracle on the headline workload, FP32: a bounded sample of steps."""
    from oracle import oracle as orc

    threads = len(os.sched_getaffinity(0))
    orc.set_threads(threads)
    shp = (n, n, n)
    q = 19
    g = np.indices(shp)
    top = (g[2] == n - 1) & (g[0] > 0) & (g[0] < n - 1) & (g[1] > 0) & (g[1] < n - 1)
    walls = (g[2] == 0) | (g[0] == 0) | (g[0] == n - 1) | (g[1] == 0) | (g[1] == n - 1)
    bc_mask = np.zeros((1,) + shp, np.uint8)
    miss = np.zeros((q,) + shp, np.uint8)
    for bid, sel in ((1, top), (2, walls)):
        idx = np.array(np.where(sel), dtype=np.int32)
        orc.mask_indices(1, idx, np.zeros(idx.shape[1], np.uint8), bid, bc_mask, miss)
    del g, top, walls
    f0 = np.zeros((q,) + shp, np.float32)
    orc.equilibrium(1, orc.F32, 0, np.ones(shp), np.zeros((3,) + shp), f0)
    f1 = f0.copy()
    bcs = [dict(id=1, kind=orc.EQUILIBRIUM, rho=1.0, u=(0.02, 0.0, 0.0)), dict(id=2, kind=orc.FULLWAY_BB)]
    for _ in range(2):
        orc.step(1, 0, orc.F32, 0, f0, f1, bc_mask, miss, bcs, 1.0)
        f0, f1 = f1, f0
    steps, t0 = 0, time.perf_counter()
    while steps < max_steps and (time.perf_counter() - t0) < budget_s:
        orc.step(1, 0, orc.F32, 0, f0, f1, bc_mask, miss, bcs, 1.0)
        f0, f1 = f1, f0
        steps += 1
    dt = time.perf_counter() - t0
    return dict(value=n**3 * steps / dt / 1e6, unit="MLUPS", cores=threads, kind="port",
                sample=f"{steps} steps of the {n}^3 D3Q19 BGK cavity (FP32) after 2 warm-up steps, C++/OpenMP oracle")


def load_traffic(n):
    path = TRAFFIC_JSON.format(n=n)
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh).get("bytes_per_launch")
    return None


def profile_pass(xa, torch, n, steps, policies=("FP32FP32",)):
    """What the rocprofv3 runs profile: one 1 GiB device copy (the counter calibration), then
    ``steps`` headline steps for each storage policy named."""
    a = torch.ones(COPY_CALIBRATION_BYTES // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    for name in policies:
        r = headline_leg(xa, torch, xa.PrecisionPolicy[name], n, steps, 5, python_loop=False)
        log(f"bench: profile pass {name} {r['mlups']:.0f} MLUPS, kernel {r['kernel_ms']:.4f} ms")


def _read_pmc(path, counter):
    """[(dispatch id, kernel name, value)] of one counter from a rocprofv3 counter_collection.csv,
    in dispatch order."""
    rows = []
    with open(path, newline="") as fh:
        for row in csv.DictReader(fh):
            if row.get("Counter_Name") == counter:
                rows.append((int(row["Dispatch_Id"]), row["Kernel_Name"], float(row["Counter_Value"])))
    return sorted(rows)


def summarize_pmc(fetch_csv, write_csv, n):
    """FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of each step kernel profiled, each scaled by what
    the same counter reads for the 1 GiB calibration copy (the first three copy dispatches of the
    profiled process, issued by ``profile_pass`` before anything else is copied).  The FP32 step
    kernel's figures go to profiles/traffic_step_d3q19_cavity{n}.json (the headline line's
    ``traffic``); a 16-bit kernel's to profiles/traffic_step16_<store>_d3q19_cavity{n}.json."""
    fetch, write = _read_pmc(fetch_csv, "FETCH_SIZE"), _read_pmc(write_csv, "WRITE_SIZE")

    def calib(rows):
        vals = [v for _, k, v in rows if "copy" in k.lower() and "xlbh" not in k][:3]
        if len(vals) != 3:
            raise SystemExit("calibration copy dispatches not found")
        return float(np.median(vals))

    kernels = sorted({k for _, k, _ in fetch if "k_step" in k})
    if not kernels:
        raise SystemExit("no step kernel in the counter file")
    copy_fetch, copy_write = calib(fetch), calib(write)
    kib = COPY_CALIBRATION_BYTES / 1024.0
    f_scale, w_scale = kib / copy_fetch, kib / copy_write
    store_names = {0: "f32", 1: "f64", 2: "f16", 3: "bf16", 4: "f16_shifted", 5: "bf16_shifted", 6: "i16_block"}
    results = {}
    for kernel in kernels:
        store = int(re.search(r"k_step\w*<\d+, \d+, (\d+),", kernel).group(1))
        sbytes = {0: 4, 1: 8}.get(store, 2)
        step_fetch = float(np.median([v for _, k, v in fetch if k == kernel]))
        step_write = float(np.median([v for _, k, v in write if k == kernel]))
        traffic = (step_fetch * f_scale + step_write * w_scale) * 1024.0
        algorithmic = 2 * 19 * sbytes * n**3
        out = dict(kernel=kernel, copy_bytes=COPY_CALIBRATION_BYTES,
                   copy_fetch_kib=copy_fetch, copy_write_kib=copy_write, fetch_scale=f_scale, write_scale=w_scale,
                   step_fetch_kib=step_fetch, step_write_kib=step_write,
                   fetch_bytes=step_fetch * f_scale * 1024.0, write_bytes=step_write * w_scale * 1024.0,
                   bytes_per_launch=traffic, algorithmic_bytes_per_launch=algorithmic,
                   traffic_over_algorithmic=traffic / algorithmic)
        path = TRAFFIC_JSON.format(n=n) if store == 0 else os.path.join(
            os.path.dirname(TRAFFIC_JSON), f"traffic_step16_{store_names[store]}_d3q19_cavity{n}.json")
        with open(path, "w") as fh:
            json.dump(out, fh, indent=1)
        results[kernel] = out
    print(json.dumps(results, indent=1))
    return results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-store16", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the precision, adjoint and KBC legs")
    ap.add_argument("--quick", action="store_true", help="smaller extra legs")
    ap.add_argument("--profile", action="store_true")
    ap.add_argument("--profile-policies", default="FP32FP32",
                    help="comma-separated precision policies the --profile pass steps (default FP32FP32)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"))
    ap.add_argument("--only", default=None,
                    help="comma-separated extra legs to run alone (kbc_sphere, grads_sphere, les_cavity, mesh_masker, "
                         "adjoint, periodic, precision_viscosity_sweep, slab), printing their JSON; for profiling")
    args = ap.parse_args()
    if args.summarize_pmc:
        summarize_pmc(args.summarize_pmc[0], args.summarize_pmc[1], args.size)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the package has no multi-GPU path (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import torch
    import xlb_amd as xa

    n = args.size
    if args.profile:
        profile_pass(xa, torch, n, args.steps, args.profile_policies.split(","))
        return
    if args.only:
        legs = {
            "kbc_sphere": lambda: kbc_sphere_leg(xa, torch, quick=args.quick),
            "grads_sphere": lambda: kbc_sphere_leg(xa, torch, quick=args.quick, sphere_bc="grads"),
            "les_cavity": lambda: les_cavity_leg(xa, torch, 128 if args.quick else n, args.steps, args.warmup),
            "mesh_masker": lambda: mesh_leg(xa, torch, subdivisions=4 if args.quick else 6),
            "adjoint": lambda: adjoint_leg(xa, torch, quick=args.quick),
            "periodic": lambda: periodic_leg(xa, torch, n, args.steps, args.warmup),
            "precision_viscosity_sweep": lambda: precision_viscosity_sweep(xa, torch),
            "slab": lambda: slab_leg(xa, torch, n),
        }
        names = args.only.split(",")
        unknown = [k for k in names if k not in legs]
        if unknown:
            sys.exit(f"unknown legs {unknown}; choose from {sorted(legs)}")
        print(json.dumps({k: legs[k]() for k in names}))
        return
    log(f"bench: fp32 leg, {n}^3, warmup {args.warmup}, steps {args.steps}")
    fp32 = headline_leg(xa, torch, xa.PrecisionPolicy.FP32FP32, n, args.steps, args.warmup)
    log(f"bench: fp32 {fp32['mlups']:.0f} MLUPS ({fp32['achieved_gbs']:.0f} GB/s)")
    store16 = {}
    if not args.no_store16:
        # In a kernel trace of this workload the first ~50 steps after the host's synchronisation
        # run up to 17 % slower for the block-format kernel and up to 7 % for shifted FP16 (the FP32
        # kernel's do not), then settle; a 20-step window would time that transient only.  The
        # 16-bit rates, and the FP32 rate they are compared with, are therefore measured over a
        # longer window of the same workload: 50 warm-up + 200 timed steps each.
        w16, s16 = max(args.warmup, 50), max(args.steps, 200)
        ref = fp32 if (s16, w16) == (args.steps, args.warmup) else headline_leg(
            xa, torch, xa.PrecisionPolicy.FP32FP32, n, s16, w16, python_loop=False)
        store16["window"] = dict(warmup=w16, steps=s16, fp32_mlups=ref["mlups"], fp32_kernel_ms=ref["kernel_ms"])
        for pol in (xa.PrecisionPolicy.FP32I16_BLOCK, xa.PrecisionPolicy.FP32FP16, xa.PrecisionPolicy.FP32FP16_SHIFTED,
                    xa.PrecisionPolicy.FP32BF16_SHIFTED):
            r = headline_leg(xa, torch, pol, n, s16, w16, python_loop=False)
            log(f"bench: {pol.name} {r['mlups']:.0f} MLUPS ({r['achieved_gbs']:.0f} GB/s), "
                f"{r['mlups'] / ref['mlups']:.2f} x FP32 over the same window")
            store16[pol.name] = dict(mlups=r["mlups"], achieved_gbs=r["achieved_gbs"], kernel_ms=r["kernel_ms"],
                                     speedup_vs_fp32=r["mlups"] / ref["mlups"], frac=r["achieved_gbs"] / HBM_PEAK_GBS)
    copy = copy_leg(torch)
    log(f"bench: device copy {copy['gbs']:.0f} GB/s")
    pull = pull_pattern_leg(xa, torch, xa.PrecisionPolicy.FP32FP32, n)
    log(f"bench: bare pull pattern (Stream) {pull['gbs']:.0f} GB/s ({pull['kernel_ms']:.4f} ms)")
    if store16:
        # the same pattern on 2-byte populations: the ceiling of the 16-bit step kernels
        pull16 = pull_pattern_leg(xa, torch, xa.PrecisionPolicy.FP32FP16_SHIFTED, n)
        log(f"bench: bare pull pattern, 2-byte populations {pull16['gbs']:.0f} GB/s ({pull16['kernel_ms']:.4f} ms)")
        store16["pull_pattern_gbs"] = pull16["gbs"]
        for name, rec in store16.items():
            if isinstance(rec, dict) and "achieved_gbs" in rec:
                rec["frac_of_pull_pattern"] = rec["achieved_gbs"] / pull16["gbs"]
    extras = {}
    if not args.no_extras:
        extras["periodic"] = periodic_leg(xa, torch, n, args.steps, args.warmup)
        log(f"bench: periodic {extras['periodic']['mlups']:.0f} MLUPS ({extras['periodic']['achieved_gbs']:.0f} GB/s)")
        extras["c1"] = c1_leg(xa, torch, with_cpu=not args.no_cpu)
        log(f"bench: c1 " + ", ".join(f"{k}={v}" for k, v in extras["c1"].items() if k != "workload"))
        pshape = (128, 64, 64) if args.quick else (512, 256, 256)
        extras["precision"] = precision_leg(xa, torch, pshape, 200 if args.quick else 2000)
        extras["precision_headline"] = headline_precision_leg(xa, torch, 128 if args.quick else n, 1000)
        extras["precision_viscosity_sweep"] = precision_viscosity_sweep(xa, torch, pshape, 200 if args.quick else 2000)
        extras["adjoint"] = adjoint_leg(xa, torch, quick=args.quick)
        extras["kbc_sphere"] = kbc_sphere_leg(xa, torch, quick=args.quick)
        extras["grads_sphere"] = kbc_sphere_leg(xa, torch, quick=args.quick, sphere_bc="grads")
        extras["mesh_masker"] = mesh_leg(xa, torch, subdivisions=4 if args.quick else 6)
        extras["les_cavity"] = les_cavity_leg(xa, torch, 128 if args.quick else n, args.steps, args.warmup)
        extras["slab"] = slab_leg(xa, torch, 128 if args.quick else n)
    cpu = None
    if not args.no_cpu:
        log("bench: cpu baseline (oracle) ...")
        cpu = cpu_leg(n)
        log(f"bench: cpu {cpu['value']:.1f} MLUPS on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": fp32["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fp32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"lid-driven cavity D3Q19 BGK {n}^3 (mlups_3d.py setup: EquilibriumBC lid u=(0.02,0,0), "
                        "FullwayBounceBackBC walls, omega=1.0, rest start)",
            "grid": [n, n, n],
            "lattice": "D3Q19",
            "collision": "BGK",
            "precision_policy": "FP32FP32",
            "parallelism": "single-GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": fp32["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": fp32["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": load_traffic(n),
            "kernel": "xlbh::k_step<D3Q19, BGK, F32, float>",
            "algorithmic_bytes_per_cell": fp32["bytes_per_lu"],
            "kernel_ms": fp32["kernel_ms"],
            "device_copy_gbs": copy["gbs"],
            "frac_of_device_copy": fp32["achieved_gbs"] / copy["gbs"],
            "pull_pattern_gbs": pull["gbs"],
            "frac_of_pull_pattern": fp32["achieved_gbs"] / pull["gbs"],
        },
        "cpu_baseline": cpu,
        "mlups_events": fp32["mlups_events"],
        "mlups_python_loop": fp32["mlups_python_loop"],
        "store16": store16,
        "finite": fp32["finite"],
    }
    out.update(extras)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
