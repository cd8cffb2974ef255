"""Vorticity and QCriterion (xlbh_vorticity / xlbh_q_criterion) against the post-processing oracle
on random velocity fields with scattered boundary cells: FP64 to 1e-13, FP32 to 1e-5 relative to
the field's scale; cells the reference skips keep their previous values.  Without a mask every
interior cell is evaluated (the oracle then runs with an all-zero mask)."""

import numpy as np
import pytest
import torch

import xlb_amd as xa
from oracle import postprocess as pp

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("masked", [True, False])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-13), (torch.float32, 1e-5)])
def test_postprocess_matches_oracle(dtype, tol, masked):
    rng = np.random.default_rng(1)
    shape = (20, 17, 23)
    u_np = 0.05 * rng.standard_normal((3,) + shape)
    bc_np = (rng.random((1,) + shape) < 0.03).astype(np.uint8) * 7
    if not masked:
        bc_np[...] = 0
    u = torch.from_numpy(u_np).to(dtype).cuda()
    bc = torch.from_numpy(bc_np).cuda()
    u_np = u.cpu().numpy()
    sentinel = -3.0
    vort = torch.full((3,) + shape, sentinel, dtype=dtype, device="cuda")
    mag = torch.full((1,) + shape, sentinel, dtype=dtype, device="cuda")
    mu = torch.full((1,) + shape, sentinel, dtype=dtype, device="cuda")
    q = torch.full((1,) + shape, sentinel, dtype=dtype, device="cuda")
    xa.operator.postprocess.Vorticity()(u, bc if masked else None, vort, mag)
    xa.operator.postprocess.QCriterion()(u, bc if masked else None, mu, q)
    o = {k: np.full(t.shape, sentinel, dtype=u_np.dtype) for k, t in (("vort", vort), ("mag", mag), ("mu", mu), ("q", q))}
    pp.vorticity(u_np, bc_np, o["vort"], o["mag"])
    pp.q_criterion(u_np, bc_np, o["mu"], o["q"])
    for k, t in (("vort", vort), ("mag", mag), ("mu", mu), ("q", q)):
        got = t.cpu().numpy()
        ref = o[k]
        skipped = ref == sentinel
        np.testing.assert_array_equal(got == sentinel, skipped, err_msg=k)
        scale = np.abs(ref[~skipped]).max()
        err = np.abs(got - ref)[~skipped].max()
        assert err <= tol * scale, f"{k}: {err:.3e} > {tol} x {scale:.3e}"
