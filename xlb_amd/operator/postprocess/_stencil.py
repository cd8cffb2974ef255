"""Shared argument handling of the velocity-gradient post-processing operators."""

import torch

from xlb_amd import _lib
from xlb_amd.operator.operator import expect_field, field_extents
from xlb_amd.precision_policy import Precision


def run(op_name, entry, u, bc_mask, out_a, card_a, out_b):
    _lib.require_cuda(u, bc_mask, out_a, out_b)
    nx, ny, nz = field_extents(u)
    expect_field(u, 3, (nx, ny, nz), "u")
    if u.dtype not in (torch.float32, torch.float64):
        raise ValueError(f"{op_name} supports float32 and float64 velocity fields")
    expect_field(out_a, card_a, (nx, ny, nz), "output", u.dtype)
    expect_field(out_b, 1, (nx, ny, nz), "output", u.dtype)
    if bc_mask is not None:
        expect_field(bc_mask, 1, (nx, ny, nz), "bc_mask", torch.uint8)
    _lib.check(
        getattr(_lib.lib(), entry)(Precision.from_torch(u.dtype).field_id, _lib.ptr(u), _lib.ptr(bc_mask),
                                   _lib.ptr(out_a), _lib.ptr(out_b), nx, ny, nz, _lib.stream_of(u)),
        op_name,
    )
    _lib.mark_written(out_a, out_b)
    return out_a, out_b
