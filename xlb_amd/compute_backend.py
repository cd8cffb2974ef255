"""Compute backends (mirrors xlb/compute_backend.py:6-8).

``MI355X`` is this package's backend: hand-written HIP kernels for gfx950 behind the C ABI of
``libxlbhip.so``.  ``JAX`` and ``WARP`` are kept so scripts written against the reference still
parse; selecting them raises, because neither runtime is part of this package.
"""

from enum import Enum, auto


class ComputeBackend(Enum):
    JAX = auto()
    WARP = auto()
    MI355X = auto()
