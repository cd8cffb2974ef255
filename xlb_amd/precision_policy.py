"""Store vs compute precision (mirrors xlb/precision_policy.py:8-89).

The reference's members keep their names and order.  Added:

* ``Precision.BF16`` and ``PrecisionPolicy.FP32BF16`` -- bfloat16 storage, FP32 arithmetic;
* ``PrecisionPolicy.FP32FP16_SHIFTED`` / ``FP32BF16_SHIFTED`` -- populations stored as the
  deviation ``f_l - w_l`` in FP16 / BF16 with FP32 arithmetic.  A stored field of these policies
  holds deviations, so raw tensor values differ from ``f`` by the lattice weights; every operator
  of this package reads and writes them in that form.
* ``Precision.INT16`` and ``PrecisionPolicy.FP32I16_BLOCK`` -- 16-bit block floating point with
  FP32 arithmetic: int16 mantissas of ``f_l / w_l - 1`` and one exponent per row segment of 64 to
  256 cells, kept after the last plane of the field's storage (include/xlb_hip.h,
  XLBH_I16_BLOCK).  Its fields come from ``create_field(..., padded=True)`` /
  ``create_nse_fields``; ``decode_populations`` turns one into FP32 values.
* ``PrecisionPolicy.FP32FP32_DEVIATION`` -- FP32 storage and arithmetic, with the fused step's BGK
  collision evaluated on deviations from the rest state (f_l - w_l, rho - 1), the cell's moment
  sums accumulated exactly and the lattice weights applied without a systematic rounding
  (include/xlb_hip.h, XLBH_COMPUTE_F32_DEVIATION): the same update as FP32FP32 with roundings
  that do not add up coherently over long runs.  BGK only; every other operator computes as
  FP32FP32.
"""

from enum import Enum, auto

import torch


class Precision(Enum):
    FP64 = auto()
    FP32 = auto()
    FP16 = auto()
    UINT8 = auto()
    BOOL = auto()
    BF16 = auto()
    INT16 = auto()

    @property
    def torch_dtype(self):
        return {
            Precision.FP64: torch.float64,
            Precision.FP32: torch.float32,
            Precision.FP16: torch.float16,
            Precision.UINT8: torch.uint8,
            Precision.BOOL: torch.bool,
            Precision.BF16: torch.bfloat16,
            Precision.INT16: torch.int16,
        }[self]

    @property
    def field_id(self):
        """Element-type id of macroscopic fields in the C ABI (XLBH_FIELD_*)."""
        ids = {Precision.FP32: 0, Precision.FP64: 1, Precision.FP16: 2, Precision.BF16: 3}
        if self not in ids:
            raise ValueError(f"{self} is not a floating-point field type")
        return ids[self]

    @staticmethod
    def from_torch(dtype):
        for p in Precision:
            if p.torch_dtype == dtype:
                return p
        raise ValueError(f"unsupported dtype {dtype}")


class PrecisionPolicy(Enum):
    FP64FP64 = auto()
    FP64FP32 = auto()
    FP64FP16 = auto()
    FP32FP32 = auto()
    FP32FP16 = auto()
    FP32BF16 = auto()
    FP32FP16_SHIFTED = auto()
    FP32BF16_SHIFTED = auto()
    FP32I16_BLOCK = auto()
    FP32FP32_DEVIATION = auto()

    @property
    def compute_precision(self):
        if self in (PrecisionPolicy.FP64FP64, PrecisionPolicy.FP64FP32, PrecisionPolicy.FP64FP16):
            return Precision.FP64
        return Precision.FP32

    @property
    def store_precision(self):
        return {
            PrecisionPolicy.FP64FP64: Precision.FP64,
            PrecisionPolicy.FP64FP32: Precision.FP32,
            PrecisionPolicy.FP64FP16: Precision.FP16,
            PrecisionPolicy.FP32FP32: Precision.FP32,
            PrecisionPolicy.FP32FP16: Precision.FP16,
            PrecisionPolicy.FP32BF16: Precision.BF16,
            PrecisionPolicy.FP32FP16_SHIFTED: Precision.FP16,
            PrecisionPolicy.FP32BF16_SHIFTED: Precision.BF16,
            PrecisionPolicy.FP32I16_BLOCK: Precision.INT16,
            PrecisionPolicy.FP32FP32_DEVIATION: Precision.FP32,
        }[self]

    @property
    def shifted(self):
        """True when populations are stored as f_l - w_l."""
        return self in (PrecisionPolicy.FP32FP16_SHIFTED, PrecisionPolicy.FP32BF16_SHIFTED)

    @property
    def block(self):
        """True for the 16-bit block floating-point store (int16 mantissas, row-segment exponents)."""
        return self == PrecisionPolicy.FP32I16_BLOCK

    @property
    def store_format(self):
        """Population storage format id in the C ABI (XLBH_F32 ... XLBH_I16_BLOCK)."""
        if self.block:
            return 6
        base = {Precision.FP32: 0, Precision.FP64: 1, Precision.FP16: 2, Precision.BF16: 3}[self.store_precision]
        if self.shifted:
            base += 2
        return base

    @property
    def compute_id(self):
        """Compute type id in the C ABI (XLBH_COMPUTE_F32 / XLBH_COMPUTE_F64 /
        XLBH_COMPUTE_F32_DEVIATION)."""
        if self == PrecisionPolicy.FP32FP32_DEVIATION:
            return 2
        return 1 if self.compute_precision == Precision.FP64 else 0
