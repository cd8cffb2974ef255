"""Operator base class and backend registry (mirrors xlb/operator/operator.py:10-134).

``register_backend`` / ``__call__`` keep the reference's dispatch: every implementation
registered for (class name, backend) is tried with the call's arguments; if none binds and runs,
a generic ``Exception`` carrying the last error and its traceback is raised
(operator.py:54-74).  Signatures are computed once at registration instead of on every call.
"""

import inspect
import traceback

from xlb_amd import _lib
from xlb_amd.compute_backend import ComputeBackend
from xlb_amd.default_config import DefaultConfig
from xlb_amd.precision_policy import Precision


class Operator:
    _backends = {}
    _signatures = {}

    def __init__(self, velocity_set=None, precision_policy=None, compute_backend=None):
        self.velocity_set = velocity_set or DefaultConfig.velocity_set
        self.precision_policy = precision_policy or DefaultConfig.default_precision_policy
        self.compute_backend = compute_backend or DefaultConfig.default_backend or ComputeBackend.MI355X
        if self.compute_backend not in ComputeBackend:
            raise ValueError(f"Compute_backend {compute_backend} is not supported")
        if self.compute_backend != ComputeBackend.MI355X:
            raise ValueError(f"Compute_backend {self.compute_backend} is not available (this package implements ComputeBackend.MI355X)")

    @classmethod
    def register_backend(cls, backend_name):
        def decorator(func):
            subclass_name = func.__qualname__.split(".")[0]
            signature = inspect.signature(func)
            key = (subclass_name, backend_name, str(signature))
            cls._backends[key] = func
            cls._signatures[key] = signature
            return func

        return decorator

    def __call__(self, *args, callback=None, **kwargs):
        error, traceback_str, key = None, "", None
        for klass in type(self).__mro__:
            candidates = [
                (k, m) for k, m in self._backends.items() if k[0] == klass.__name__ and k[1] == self.compute_backend
            ]
            for key, backend_method in candidates:
                try:
                    self._signatures[key].bind(self, *args, **kwargs)
                    result = backend_method(self, *args, **kwargs)
                    callback_arg = result if result is not None else (args, kwargs)
                    if callback and callable(callback):
                        callback(callback_arg)
                    return result
                except Exception as e:  # same contract as the reference: try the next candidate
                    error = e
                    traceback_str = traceback.format_exc()
                    continue
            if candidates:
                break
        raise Exception(f"Error captured for backend with key {key} for operator {self.__class__.__name__}: {error}\n {traceback_str}")

    @property
    def supported_compute_backend(self):
        return list(self._backends.keys())

    def __repr__(self):
        return f"{self.__class__.__name__}()"

    @property
    def backend(self):
        import torch

        return torch

    @property
    def compute_dtype(self):
        return self.precision_policy.compute_precision.torch_dtype

    @property
    def store_dtype(self):
        return self.precision_policy.store_precision.torch_dtype

    # ---- helpers shared by the MI355X implementations ----
    @property
    def lattice_id(self):
        return self.velocity_set.lattice_id

    def format_of(self, f):
        """Population storage format of field ``f`` (the policy's, or the plain format of its dtype)."""
        policy = self.precision_policy
        if f.dtype == policy.store_precision.torch_dtype:
            return policy.store_format
        return {Precision.FP32: _lib.F32, Precision.FP64: _lib.F64, Precision.FP16: _lib.F16, Precision.BF16: _lib.BF16}[
            Precision.from_torch(f.dtype)
        ]

    @property
    def compute_id(self):
        return self.precision_policy.compute_id


def plane_stride(f):
    """Elements between consecutive planes of a (card, nx, ny, nz) field laid out in C order within
    each plane: nx*ny*nz for a contiguous tensor, more for a plane-padded population field
    (``create_field(..., padded=True)``).  None if the field has another layout."""
    if f.dim() != 4:
        return None
    card, nx, ny, nz = (int(v) for v in f.shape)
    n = nx * ny * nz
    want = (None, ny * nz, nz, 1)
    for d in (1, 2, 3):
        if f.shape[d] > 1 and f.stride(d) != want[d]:
            return None
    if card == 1:
        return n
    p = int(f.stride(0))
    return p if p >= n else None


def field_extents(f):
    """(nx, ny, nz) of a (card, nx, ny, nz) field (contiguous, or plane-padded)."""
    if f.dim() != 4:
        raise ValueError(f"fields must have shape (cardinality, nx, ny, nz); got {tuple(f.shape)}")
    if plane_stride(f) is None:
        raise ValueError("fields must be C-ordered (z unit-stride) within each plane")
    return int(f.shape[1]), int(f.shape[2]), int(f.shape[3])


def expect_population(t, q, extents, name, dtype, lattice, store):
    """Check a population field (shape (q, *extents), C order within each plane, on the GPU, of
    dtype) and that its storage holds every plane (and, for the block format, its segment
    exponents); returns its plane stride."""
    if t is None:
        raise ValueError(f"{name} is required")
    if not t.is_cuda:
        raise ValueError(f"{name} must live on the GPU (device='cuda')")
    if tuple(t.shape) != (int(q),) + tuple(extents):
        raise ValueError(f"{name} has shape {tuple(t.shape)}; expected {(int(q),) + tuple(extents)}")
    if t.dtype != dtype:
        raise ValueError(f"{name} has dtype {t.dtype}; expected {dtype}")
    plane = plane_stride(t)
    if plane is None:
        raise ValueError(f"{name} must be C-ordered within each plane (a contiguous tensor or a plane-padded field)")
    need = int(_lib.lib().xlbh_population_bytes(lattice, store, *extents, plane))
    have = t.untyped_storage().nbytes() - t.storage_offset() * t.element_size()
    if need < 0 or have < need:
        what = " and its segment exponents" if store == _lib.I16_BLOCK else ""
        raise ValueError(f"{name}: its storage holds {have} bytes, fewer than the {need} its planes{what} need "
                         "(allocate population fields with grid.create_field(..., padded=True) or create_nse_fields)")
    return plane


def shared_layout(fields):
    """(plane, fields): the fields as given when they share one plane stride, else contiguous
    copies of them (plane N).  Callers copy outputs back when the tensor object changed."""
    planes = {plane_stride(t) for t in fields if t is not None}
    if len(planes) == 1 and None not in planes:
        return planes.pop(), list(fields)
    out = [t.contiguous() if t is not None else None for t in fields]
    return 0, out


def common_plane(planes, what):
    """The one plane stride shared by the population fields of a call."""
    ps = {p for p in planes if p is not None}
    if len(ps) > 1:
        raise ValueError(f"{what}: population fields of one call must share one plane stride; got {sorted(ps)}")
    return ps.pop() if ps else 0


def expect_field(t, cardinality, extents, name, dtype=None):
    """Check a field handed to a kernel: shape (cardinality, *extents), contiguous, on the GPU,
    and optionally of one dtype.  Kernels index fields through raw pointers, so every size is
    validated here before a launch."""
    if t is None:
        raise ValueError(f"{name} is required")
    if not t.is_cuda:
        raise ValueError(f"{name} must live on the GPU (device='cuda')")
    if tuple(t.shape) != (int(cardinality),) + tuple(extents):
        raise ValueError(f"{name} has shape {tuple(t.shape)}; expected {(int(cardinality),) + tuple(extents)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"{name} has dtype {t.dtype}; expected {dtype}")
    return t
