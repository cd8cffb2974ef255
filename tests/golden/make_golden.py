"""Writes the golden fixtures of the fused step: small seeded configurations covering every
boundary kind, every collision and the body force, run for 5 FP64 steps by the CPU oracle
(oracle/xlb_oracle.cpp, the restatement of nse_stepper.py:245-392).

Each fixture holds the inputs (initial populations, bc_mask, missing mask, per-cell prescribed
values, the BC table, omega, force) and the expected populations after 5 steps.  The masks are
built through the oracle's restatement of indices_boundary_masker.py:103-224 from the boundary
indices of the product's BC classes.

Usage: python tests/golden/make_golden.py [name ...]   (writes tests/golden/step_<name>.npz, all
fixtures when no name is given)
"""

import json
import os
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

STEPS = 5

GOLDEN = {
    # name: (lattice, shape, boundary set, collision, omega, force)
    "d3q19_cavity_halfway": ("D3Q19", (8, 8, 8), "cavity_halfway", "BGK", 1.6, None),
    "d3q19_cavity_fullway": ("D3Q19", (8, 7, 6), "cavity_fullway", "BGK", 1.0, None),
    "d3q27_sphere_kbc": ("D3Q27", (10, 8, 8), "sphere3d", "KBC", 1.9, None),
    "d2q9_channel_zouhe": ("D2Q9", (16, 8), "channel2d_zouhe", "BGK", 1.2, None),
    "d2q9_channel_regularized_kbc": ("D2Q9", (16, 8), "channel2d_regularized", "KBC", 1.7, None),
    "d2q9_channel_zouhe_pressure": ("D2Q9", (16, 8), "channel2d_zouhe_pressure", "BGK", 1.1, None),
    "d3q19_do_nothing_forced": ("D3Q19", (8, 6, 6), "do_nothing", "BGK", 1.3, "uniform"),
    "d2q9_periodic_force_field": ("D2Q9", (12, 10), "periodic", "BGK", 1.3, "field"),
    "d3q27_sphere_smagorinsky": ("D3Q27", (10, 8, 8), "sphere3d", "SmagorinskyLESBGK", 1.95, None),
    "d3q19_sphere_grads": ("D3Q19", (12, 10, 10), "sphere3d_grads", "BGK", 1.7, None),
    "d3q19_cavity_grads": ("D3Q19", (8, 8, 8), "cavity_grads", "BGK", 1.4, None),
}

COLLISION_ID = {"BGK": 0, "KBC": 1, "SmagorinskyLESBGK": 2}
SMAGORINSKY_COEF = 0.17  # SmagorinskyLESBGK's default (smagorinsky_les_bgk.py:24)


def build(name):
    """Inputs of fixture `name` as a dict of numpy arrays (no GPU needed)."""
    import xlb_amd as xa
    from lbm_cases import make_bcs, oracle_aux, oracle_bc_desc, oracle_masks
    from oracle import oracle as orc

    lat_name, shape, bcset, coll, omega, force = GOLDEN[name]
    policy = xa.PrecisionPolicy.FP64FP64
    vs = getattr(xa.velocity_set, lat_name)(policy, xa.ComputeBackend.MI355X)
    xa.init(velocity_set=vs, default_backend=xa.ComputeBackend.MI355X, default_precision_policy=policy)
    xa.operator.boundary_condition.boundary_condition_registry.reset()
    grid = xa.grid.grid_factory(shape, compute_backend=xa.ComputeBackend.MI355X)
    bcs = make_bcs(bcset, grid, xa)
    shape3 = tuple(grid.field_shape)
    lat = vs.lattice_id
    bc_mask, missing = oracle_masks(lat, vs.q, vs.d, shape3, bcs)
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    rho = 1.0 + 0.01 * rng.standard_normal(shape3)
    u = 0.03 * rng.standard_normal((vs.d,) + shape3)
    f0 = np.zeros((vs.q,) + shape3)
    orc.equilibrium(lat, orc.F64, 1, rho, u, f0)
    f0 += 1e-3 * np.asarray(vs.w).reshape((-1, 1, 1, 1)) * rng.standard_normal(f0.shape)
    aux = np.zeros(int(np.prod(shape3)))
    for bc in bcs:
        if bc.needs_aux_init:
            vals = oracle_aux(bc, bc_mask, vs.d, policy)
            sel = bc_mask.reshape(-1) == bc.id
            aux[sel] = vals[sel]
    out = dict(
        lattice=np.array(lat), collision=np.array(COLLISION_ID[coll]), omega=np.array(omega),
        f_init=f0, bc_mask=bc_mask, missing=missing, aux=aux,
        bcs=np.array(json.dumps([oracle_bc_desc(bc) for bc in bcs])),
    )
    if coll == "SmagorinskyLESBGK":
        out["cparam"] = np.array(SMAGORINSKY_COEF)
    if force == "uniform":
        out["force"] = np.array([1e-4, -2e-4, 5e-5])[: vs.d]
    if force == "field":
        out["force_field"] = 1e-4 * rng.standard_normal((vs.d,) + shape3)
    return out


def run_oracle(fx, steps=STEPS):
    """The oracle's FP64 populations after `steps` steps from the fixture's inputs."""
    from oracle import oracle as orc

    f0 = fx["f_init"].copy()
    f1 = f0.copy()
    bc_mask = fx["bc_mask"]
    aux = fx["aux"]
    bcs = json.loads(str(fx["bcs"]))
    for d in bcs:
        if d["kind"] in (orc.ZOUHE_VELOCITY, orc.ZOUHE_PRESSURE, orc.REGULARIZED_VELOCITY, orc.REGULARIZED_PRESSURE):
            orc.aux_init(orc.F64, f1, bc_mask, d["id"], aux)
    force = fx["force"] if "force" in fx else None
    field = fx["force_field"] if "force_field" in fx else None
    cparam = float(fx["cparam"]) if "cparam" in fx else 0.0
    for _ in range(steps):
        orc.step(int(fx["lattice"]), int(fx["collision"]), orc.F64, 1, f0, f1, bc_mask, fx["missing"], bcs,
                 float(fx["omega"]), force=force, force_field=field, cparam=cparam)
        f0, f1 = f1, f0
    return f0


def path(name):
    return os.path.join(HERE, f"step_{name}.npz")


def main(names):
    for name in names or GOLDEN:
        fx = build(name)
        fx["f_final"] = run_oracle(fx)
        np.savez_compressed(path(name), **fx)
        print(name, os.path.getsize(path(name)), "bytes")


if __name__ == "__main__":
    main(sys.argv[1:])
