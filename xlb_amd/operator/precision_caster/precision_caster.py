"""PrecisionCaster (mirrors xlb/operator/precision_caster/precision_caster.py:15-82): casts a
field from one Precision to another, value by value (xlbh_cast).  Shifted population storage is
a property of a precision policy, not of a Precision, so this operator casts stored values as
they are."""

import torch

from xlb_amd import _lib
from xlb_amd.compute_backend import ComputeBackend
from xlb_amd.operator.operator import Operator
from xlb_amd.precision_policy import Precision


class PrecisionCaster(Operator):
    def __init__(self, input_precision: Precision, output_precision: Precision, velocity_set=None,
                 precision_policy=None, compute_backend=None):
        super().__init__(velocity_set=velocity_set, precision_policy=precision_policy, compute_backend=compute_backend)
        self.input_precision = input_precision
        self.output_precision = output_precision

    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, from_f, to_f):
        _lib.require_cuda(from_f, to_f)
        if from_f.dtype != self.input_precision.torch_dtype or to_f.dtype != self.output_precision.torch_dtype:
            raise ValueError(f"PrecisionCaster expects {self.input_precision} -> {self.output_precision}, "
                             f"got {from_f.dtype} -> {to_f.dtype}")
        if from_f.shape != to_f.shape or not from_f.is_contiguous() or not to_f.is_contiguous():
            raise ValueError("PrecisionCaster needs contiguous fields of the same shape")
        _lib.check(
            _lib.lib().xlbh_cast(self.input_precision.field_id, self.output_precision.field_id, _lib.ptr(from_f),
                                 _lib.ptr(to_f), int(from_f.numel()), _lib.stream_of(from_f)),
            "PrecisionCaster",
        )
        _lib.mark_written(to_f)
        return to_f
