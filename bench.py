"""Benchmark of the fused LBM step on one MI355X.

Workload (the reference's own benchmark, examples/performance/mlups_3d.py:45-90, at edge 256):
D3Q19 BGK lid-driven cavity 256^3, rho = 1 / u = 0 equilibrium start, lid EquilibriumBC(rho=1,
u=(0.02, 0, 0)) on the top face without edges, FullwayBounceBackBC on the other five full faces,
omega = 1.0.  One step = one fused stream-collide-BC pass over the grid; the timed loop runs
K steps with the buffer swap below the C ABI (xlbh_run), all inputs resident in HBM.

value = MLUPS = 256^3 * K / t / 1e6 with the FP32/FP32 policy.  The same workload is also timed
with 16-bit population storage (plain FP16, and f - w_l shifted FP16 / BF16) and reported under
"store16".  roofline.achieved = algorithmic bytes per launch (2 Q S_store per cell) / average
kernel duration from HIP events on the launch stream.  cpu_baseline = the C++/OpenMP oracle
(the parity checker, a restatement of the reference algorithm) timed on a bounded number of
steps of the same workload on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 256] [--no-cpu]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS & % HBM roofline, D3Q19 BGK 256³ (FP32 vs FP16-store); adjoint/fwd time"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cavity_bcs(xa, grid):
    box = grid.bounding_box_indices()
    box_no_edge = grid.bounding_box_indices(remove_edges=True)
    lid = box_no_edge["top"]
    walls = [box["bottom"][i] + box["left"][i] + box["right"][i] + box["front"][i] + box["back"][i] for i in range(3)]
    walls = np.unique(np.array(walls), axis=-1).tolist()
    BC = xa.operator.boundary_condition
    return [BC.EquilibriumBC(rho=1.0, u=(0.02, 0.0, 0.0), indices=lid), BC.FullwayBounceBackBC(indices=walls)]


def gpu_leg(xa, torch, policy, n, steps, warmup):
    vs = xa.velocity_set.D3Q19(policy, xa.ComputeBackend.MI355X)
    xa.init(velocity_set=vs, default_backend=xa.ComputeBackend.MI355X, default_precision_policy=policy)
    xa.operator.boundary_condition.boundary_condition_registry.reset()
    grid = xa.grid.grid_factory((n, n, n))
    stepper = xa.operator.stepper.IncompressibleNavierStokesStepper(grid, cavity_bcs(xa, grid), collision_type="BGK")
    f0, f1, bc_mask, missing = stepper.prepare_fields()
    omega = 1.0
    f0, f1 = stepper.run(f0, f1, bc_mask, missing, omega, warmup)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    f0, f1 = stepper.run(f0, f1, bc_mask, missing, omega, steps)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = ev0.elapsed_time(ev1)
    cells = n**3
    # Python-loop variant through the public __call__ (the reference's loop, mlups_3d.py:78-80)
    py_steps = min(steps, 200)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(py_steps):
        f0, f1 = stepper(f0, f1, bc_mask, missing, omega, i)
        f0, f1 = f1, f0
    torch.cuda.synchronize()
    py_wall = time.perf_counter() - t1
    finite = bool(torch.isfinite(f0.float()).all().item())
    store_bytes = torch.empty((), dtype=f0.dtype).element_size()
    bytes_per_lu = 2 * vs.q * store_bytes
    kernel_s = ev_ms / 1e3 / steps
    res = dict(
        policy=policy.name,
        mlups=cells * steps / wall / 1e6,
        mlups_events=cells * steps / (ev_ms / 1e3) / 1e6,
        mlups_python_loop=cells * py_steps / py_wall / 1e6,
        ms_per_step=wall / steps * 1e3,
        kernel_ms=kernel_s * 1e3,
        bytes_per_lu=bytes_per_lu,
        achieved_gbs=bytes_per_lu * cells / kernel_s / 1e9,
        finite=finite,
    )
    del f0, f1, bc_mask, missing, stepper
    torch.cuda.empty_cache()
    return res


def cpu_leg(n, budget_s=12.0, max_steps=40):
    """The C++/OpenMP oracle on the same workload, FP32: a bounded sample of steps."""
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
    path = os.path.join(ROOT, "profiles", f"traffic_step_d3q19_cavity{n}.json")
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh).get("bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-store16", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import torch
    import xlb_amd as xa

    n = args.size
    log(f"bench: fp32 leg, {n}^3, warmup {args.warmup}, steps {args.steps}")
    fp32 = gpu_leg(xa, torch, xa.PrecisionPolicy.FP32FP32, n, args.steps, args.warmup)
    log(f"bench: fp32 {fp32['mlups']:.0f} MLUPS ({fp32['achieved_gbs']:.0f} GB/s)")
    store16 = {}
    if not args.no_store16:
        for pol in (xa.PrecisionPolicy.FP32FP16, xa.PrecisionPolicy.FP32FP16_SHIFTED, xa.PrecisionPolicy.FP32BF16_SHIFTED):
            r = gpu_leg(xa, torch, pol, n, args.steps, args.warmup)
            log(f"bench: {pol.name} {r['mlups']:.0f} MLUPS ({r['achieved_gbs']:.0f} GB/s)")
            store16[pol.name] = dict(mlups=r["mlups"], achieved_gbs=r["achieved_gbs"], kernel_ms=r["kernel_ms"],
                                     speedup_vs_fp32=r["mlups"] / fp32["mlups"], frac=r["achieved_gbs"] / HBM_PEAK_GBS)
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
        },
        "cpu_baseline": cpu,
        "mlups_events": fp32["mlups_events"],
        "mlups_python_loop": fp32["mlups_python_loop"],
        "store16": store16,
        "finite": fp32["finite"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
