"""The product against the committed golden fixtures (tests/golden/step_*.npz): the masks and
prescribed values it builds through the operator API are bit-identical to the fixture's, and 5
FP64 fused steps from the fixture's initial state reproduce the stored populations to 1e-12."""

import json
import os
import sys

import numpy as np
import pytest
import torch

import xlb_amd as xa

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden as mg  # noqa: E402
from lbm_cases import make_bcs  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(mg.GOLDEN))
def test_product_matches_golden(name):
    fx = dict(np.load(mg.path(name), allow_pickle=False))
    lat_name, shape, bcset, coll, omega, force = mg.GOLDEN[name]
    policy = xa.PrecisionPolicy.FP64FP64
    vs = getattr(xa.velocity_set, lat_name)(policy, xa.ComputeBackend.MI355X)
    xa.init(velocity_set=vs, default_backend=xa.ComputeBackend.MI355X, default_precision_policy=policy)
    xa.operator.boundary_condition.boundary_condition_registry.reset()
    grid = xa.grid.grid_factory(shape, compute_backend=xa.ComputeBackend.MI355X)
    bcs = make_bcs(bcset, grid, xa)
    kw = {}
    if "force" in fx:
        kw["force_vector"] = fx["force"]
    if "force_field" in fx:
        kw["force_field"] = torch.from_numpy(fx["force_field"]).cuda()
    st = xa.operator.stepper.IncompressibleNavierStokesStepper(grid, bcs, collision_type=coll, **kw)
    f0, f1, bc_mask, missing = st.prepare_fields()
    np.testing.assert_array_equal(bc_mask.cpu().numpy(), fx["bc_mask"])
    np.testing.assert_array_equal(missing.cpu().numpy().astype(np.uint8), fx["missing"])
    if st._aux is not None:
        cells = fx["bc_mask"].reshape(-1) != 0
        ids = {d["id"] for d in json.loads(str(fx["bcs"])) if d["kind"] >= 5 and d["kind"] <= 8}
        sel = np.isin(fx["bc_mask"].reshape(-1), list(ids)) & cells
        np.testing.assert_array_equal(st._aux.cpu().numpy()[sel], fx["aux"][sel])
    f0.copy_(torch.from_numpy(fx["f_init"]))
    f1.copy_(f0)
    f0, f1 = st.run(f0, f1, bc_mask, missing, float(fx["omega"]), mg.STEPS)
    err = np.abs(f0.cpu().numpy() - fx["f_final"]).max()
    assert err <= 1e-12, f"{name}: max|df| = {err:.3e}"
