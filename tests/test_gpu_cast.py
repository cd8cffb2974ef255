"""PrecisionCaster (precision_caster.py:15-82) through xlbh_cast against NumPy / torch's own
conversions: one rounding per value (round to nearest even), FP64 -> 16-bit through FP32."""

import numpy as np
import pytest
import torch

import xlb_amd as xa

pytestmark = pytest.mark.gpu

Pr = xa.Precision
PAIRS = [(a, b) for a in (Pr.FP64, Pr.FP32, Pr.FP16, Pr.BF16) for b in (Pr.FP64, Pr.FP32, Pr.FP16, Pr.BF16)]


def expected(x64, src, dst):
    """The value the cast must produce, computed with torch on the CPU."""
    t = torch.from_numpy(x64).to(src.torch_dtype)
    if src == Pr.FP64 and dst in (Pr.FP16, Pr.BF16):
        t = t.to(torch.float32)
    return t.to(dst.torch_dtype)


@pytest.mark.parametrize("src,dst", PAIRS)
def test_cast_matches_torch_conversions(src, dst):
    vs = xa.velocity_set.D3Q19(xa.PrecisionPolicy.FP32FP32, xa.ComputeBackend.MI355X)
    xa.init(velocity_set=vs, default_backend=xa.ComputeBackend.MI355X, default_precision_policy=xa.PrecisionPolicy.FP32FP32)
    rng = np.random.default_rng(3)
    shape = (19, 9, 8, 7)
    x = np.concatenate([rng.standard_normal(int(np.prod(shape)) // 2) * 0.1,
                        rng.uniform(-3, 3, int(np.prod(shape)) - int(np.prod(shape)) // 2)]).reshape(shape)
    x.reshape(-1)[:4] = [0.0, -0.0, 1.0 / 3.0, 65504.0]
    src_t = torch.from_numpy(x).to(src.torch_dtype).cuda()
    out = torch.empty(shape, dtype=dst.torch_dtype, device="cuda")
    xa.operator.precision_caster.PrecisionCaster(src, dst)(src_t, out)
    want = expected(x, src, dst)
    got = out.cpu()
    if dst in (Pr.FP16, Pr.BF16):
        assert torch.equal(got.view(torch.int16), want.view(torch.int16))
    else:
        assert torch.equal(got, want)


def test_cast_rejects_wrong_dtype():
    a = torch.zeros((2, 3, 3, 3), dtype=torch.float32, device="cuda")
    b = torch.zeros((2, 3, 3, 3), dtype=torch.float16, device="cuda")
    with pytest.raises(Exception, match="PrecisionCaster expects"):
        xa.operator.precision_caster.PrecisionCaster(Pr.FP64, Pr.FP16)(a, b)
