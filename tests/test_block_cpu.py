"""The oracle's restatement of the 16-bit block floating-point store (format 6: int16 mantissas of
f_l / w_l - 1, one exponent per row segment), pinned against what the format promises, on the CPU:

* encoding: every segment's exponent is the tightest one (its largest mantissa uses the top bit,
  16383 <= max |m| <= 32767), chosen from the largest |f_l / w_l - 1| computed independently in
  float64, and every decoded value lies within half a mantissa step of the FP64 value;
* a step in the block format equals the FP32 step of the same decoded state to within half a
  mantissa step of the output segment;
* a segment holding a non-finite value decodes to +inf and leaves the other segments alone.
"""

import numpy as np
import pytest

from lbm_cases import block_exps, block_geometry, decode, pop_alloc
from oracle import oracle as orc

LATS = {0: 9, 1: 19, 2: 27}
SHAPES = [(1, (3, 4, 300)), (1, (2, 5, 40)), (2, (2, 3, 130)), (0, (6, 70, 1)), (0, (3, 300, 1))]


def _state(lat, shape3, seed, du=0.03):
    rng = np.random.default_rng(seed)
    d = 2 if lat == 0 else 3
    rho = 1 + 0.02 * rng.standard_normal(shape3)
    u = du * rng.standard_normal((d,) + shape3)
    return rho, u


def _segments(a):
    """(per-cell segment index, number of segments) of a field's grid."""
    q, nx, ny, nz = a.shape
    bz, segs, rows = block_geometry((nx, ny, nz))
    if nz > 1:
        idx = (np.arange(nx)[:, None, None] * ny + np.arange(ny)[None, :, None]) * segs + np.arange(nz)[None, None] // bz
    else:
        idx = np.arange(nx)[:, None, None] * segs + (np.arange(ny) // bz)[None, :, None]
    return idx, segs * rows


@pytest.mark.parametrize("lat,shape3", SHAPES)
def test_block_equilibrium_encoding(lat, shape3):
    q = LATS[lat]
    w = orc.lattice(lat)["w"]
    rho, u = _state(lat, shape3, seed=lat + shape3[2])
    f64 = np.zeros((q,) + shape3)
    orc.equilibrium(lat, orc.F64, 1, rho, u, f64)
    a = pop_alloc(orc.I16_BLOCK, q, shape3)
    orc.equilibrium(lat, orc.I16_BLOCK, 0, rho, u, a)
    k = block_exps(a)
    idx, nseg = _segments(a)
    # the exponent: 14 - floor(log2 M), M the segment's largest |f / w - 1| (from the FP64 values:
    # within a float rounding of the float ones, so a power-of-two boundary may move it by one)
    x = np.abs(f64 / np.asarray(w).reshape(-1, 1, 1, 1) - 1).max(axis=0)
    M = np.zeros(nseg)
    np.maximum.at(M, idx.ravel(), x.ravel())
    want = 14 - np.floor(np.log2(M))
    near = np.abs(np.log2(M) - np.round(np.log2(M))) < 1e-6
    assert np.all((k == want) | (near & (np.abs(k - want) == 1)))
    # tight: the largest mantissa of each segment uses the top bit
    mmax = np.zeros(nseg)
    np.maximum.at(mmax, np.broadcast_to(idx, a.shape).ravel(), np.abs(a.astype(np.int64)).ravel())
    assert np.all((mmax >= 16383) & (mmax <= 32767))
    # every value within half a mantissa step (w 2^-k / 2) plus the FP32 arithmetic of feq
    step = np.asarray(w).reshape(-1, 1, 1, 1) * np.exp2(-k[idx])[None]
    err = np.abs(decode(a, orc.I16_BLOCK, w) - f64)
    assert np.all(err <= 0.5 * step + 4e-7 * np.abs(f64)), (err / step).max()


@pytest.mark.parametrize("lat,coll,shape3", [(1, 0, (3, 4, 300)), (1, 0, (4, 5, 64)), (2, 1, (2, 3, 130)),
                                             (0, 0, (6, 70, 1)), (0, 1, (3, 300, 1))])
def test_block_step_matches_fp32_step_within_half_a_step(lat, coll, shape3):
    q = LATS[lat]
    w = orc.lattice(lat)["w"]
    rho, u = _state(lat, shape3, seed=7)
    a0 = pop_alloc(orc.I16_BLOCK, q, shape3)
    orc.equilibrium(lat, orc.I16_BLOCK, 0, rho, u, a0)
    a1 = pop_alloc(orc.I16_BLOCK, q, shape3)
    orc.step(lat, coll, orc.I16_BLOCK, 0, a0, a1, None, None, [], 1.6)
    # the same decoded state through the FP32 step
    f0 = decode(a0, orc.I16_BLOCK, w).astype(np.float32)
    f1 = np.zeros_like(f0)
    orc.step(lat, coll, orc.F32, 0, f0, f1, None, None, [], 1.6)
    k = block_exps(a1)
    idx, _ = _segments(a1)
    step = np.asarray(w).reshape(-1, 1, 1, 1) * np.exp2(-k[idx])[None]
    err = np.abs(decode(a1, orc.I16_BLOCK, w) - f1.astype(np.float64))
    assert np.all(err <= 0.5 * step * (1 + 1e-6) + 2e-7 * np.abs(f1)), (err / step).max()


def test_block_nonfinite_segment_decodes_to_inf():
    lat, q, shape3 = 1, 19, (2, 3, 130)
    w = orc.lattice(lat)["w"]
    rho, u = _state(lat, shape3, seed=3)
    rho[1, 2, 129] = np.inf
    a = pop_alloc(orc.I16_BLOCK, q, shape3)
    orc.equilibrium(lat, orc.I16_BLOCK, 0, rho, u, a)
    k = block_exps(a)
    idx, _ = _segments(a)
    bad = idx[1, 2, 129]
    assert k[bad] == -127 and np.all(np.delete(k, bad) > -64 - 1)
    v = decode(a, orc.I16_BLOCK, w)  # float64: beyond the float range, i.e. +inf once rounded to float
    sel = (idx == bad)
    assert np.all(v[:, sel] > np.finfo(np.float32).max) and np.all(np.abs(v[:, ~sel]) < 1)


@pytest.mark.parametrize("lat,shape3", [(1, (3, 4, 300)), (0, (6, 70, 1))])
def test_package_decode_matches_the_format(lat, shape3):
    """xlb_amd.grid.decode_populations (and block_exponents) read a block field the way the format
    defines it: the same values as the test-side decoder, on a field the oracle wrote."""
    import torch

    from lbm_cases import _block_buffer

    import xlb_amd as xa

    q = LATS[lat]
    w = orc.lattice(lat)["w"]
    rho, u = _state(lat, shape3, seed=11)
    a = pop_alloc(orc.I16_BLOCK, q, shape3)
    orc.equilibrium(lat, orc.I16_BLOCK, 0, rho, u, a)
    t = torch.from_numpy(_block_buffer(a))[: a.size].view(a.shape)
    np.testing.assert_array_equal(xa.grid.block_exponents(t).numpy(), block_exps(a))
    got = xa.grid.decode_populations(t, w, block=True).numpy()
    np.testing.assert_allclose(got, decode(a, orc.I16_BLOCK, w), rtol=1e-15, atol=0)
