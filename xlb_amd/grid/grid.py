"""Grids and field allocation (mirrors xlb/grid/grid.py:9-86 and warp_grid.py:10-32).

Fields are torch tensors in HBM with the reference's WarpGrid layout: C order
(cardinality, nx, ny, nz), z unit-stride, and (cardinality, nx, ny, 1) for 2-D grids.
"""

from typing import Tuple

import numpy as np
import torch

from xlb_amd.compute_backend import ComputeBackend
from xlb_amd.default_config import DefaultConfig


def grid_factory(shape: Tuple[int, ...], compute_backend: ComputeBackend = None):
    compute_backend = compute_backend or DefaultConfig.default_backend or ComputeBackend.MI355X
    if compute_backend == ComputeBackend.MI355X:
        return MI355XGrid(shape)
    raise ValueError(f"Compute backend {compute_backend} is not supported")


class Grid:
    def __init__(self, shape: Tuple[int, ...], compute_backend: ComputeBackend):
        self.shape = tuple(int(s) for s in shape)
        self.dim = len(self.shape)
        self.compute_backend = compute_backend

    def bounding_box_indices(self, remove_edges=False):
        """Cells of each face of the box as index lists (d rows), the interface of grid.py:34-86:
        faces "bottom"/"top" (last axis at 0 / n-1), "left"/"right" (x) and, in 3-D,
        "front"/"back" (y), in that order.  A face holds every cell whose coordinate on the face's
        axis is fixed, enumerated in C order; with remove_edges the other coordinates run over
        1 .. n-2 only."""
        if self.dim not in _FACES:
            raise ValueError(f"grid dimension {self.dim} not supported")
        trim = 1 if remove_edges else 0
        faces = {}
        for name, axis, far in _FACES[self.dim]:
            spans = [np.arange(trim, n - trim) for n in self.shape]
            spans[axis] = np.array([self.shape[axis] - 1 if far else 0])
            cells = np.stack(np.meshgrid(*spans, indexing="ij")).reshape(self.dim, -1)
            faces[name] = cells.tolist()
        return faces


# (face name, axis held fixed, far side?) per dimension, in the order the faces are listed.
_FACES = {
    2: (("bottom", 1, False), ("top", 1, True), ("left", 0, False), ("right", 0, True)),
    3: (("bottom", 2, False), ("top", 2, True), ("left", 0, False), ("right", 0, True), ("front", 1, False),
        ("back", 1, True)),
}


def copy_population(dst, src):
    """dst <- src for two population fields of one layout, including what lies outside the view
    (the block format's segment exponents)."""
    if tuple(dst.shape) != tuple(src.shape) or dst.stride() != src.stride() or dst.dtype != src.dtype:
        raise ValueError("copy_population needs two fields of one shape, layout and dtype")
    if dst.storage_offset() == 0 and src.storage_offset() == 0 and \
            dst.untyped_storage().nbytes() == src.untyped_storage().nbytes():
        dst.untyped_storage().copy_(src.untyped_storage())
    else:
        dst.copy_(src)
    return dst


def block_exponents(f):
    """The segment exponents of a block-format population field (int32, one per row segment, in
    the order of include/xlb_hip.h) as a view into its storage."""
    from xlb_amd import _lib

    q, nx, ny, nz = (int(v) for v in f.shape)
    plane = int(f.stride(0))
    lattice = {9: _lib.D2Q9, 19: _lib.D3Q19, 27: _lib.D3Q27}[q]
    nbytes = int(_lib.lib().xlbh_population_bytes(lattice, _lib.I16_BLOCK, nx, ny, nz, plane))
    off = -(-(q * plane * 2) // 256) * 256
    raw = torch.empty(0, dtype=torch.uint8, device=f.device).set_(f.untyped_storage())
    return raw[f.storage_offset() * 2 + off: f.storage_offset() * 2 + nbytes].view(torch.int32)


def decode_populations(f, weights, block=False, shifted=False, dtype=torch.float64):
    """f_l of a stored population field as a (q, nx, ny, nz) tensor of ``dtype``: the value itself,
    f - w_l + w_l for the shifted formats, and w_l (1 + m 2^-k) for the block format (the kernels
    round that to float once; here it is evaluated in ``dtype``)."""
    q, nx, ny, nz = (int(v) for v in f.shape)
    w = torch.as_tensor([float(v) for v in weights], dtype=dtype, device=f.device).view(q, 1, 1, 1)
    if not block:
        v = f.to(dtype)
        return v + w if shifted else v
    fast = nz if nz > 1 else ny
    bz = next((b for b in (64, 128, 256) if b >= fast), 256)
    segs = -(-fast // bz)
    k = block_exponents(f).to(dtype)
    if nz > 1:
        k = k.view(nx, ny, segs).repeat_interleave(bz, dim=2)[:, :, :nz]
    else:
        k = k.view(nx, segs).repeat_interleave(bz, dim=1)[:, :ny].unsqueeze(-1)
    wf = w.float().to(dtype)  # the kernels use the float-rounded weights
    return wf * (1 + f.to(dtype) * torch.pow(2.0, -k).unsqueeze(0))


class MI355XGrid(Grid):
    def __init__(self, shape, device=None):
        super().__init__(shape, ComputeBackend.MI355X)
        self.device = torch.device(device) if device is not None else None

    @property
    def field_shape(self):
        return self.shape if len(self.shape) != 2 else self.shape + (1,)

    @property
    def nxyz(self):
        s = self.field_shape
        return int(s[0]), int(s[1]), int(s[2])

    def _device(self):
        if self.device is not None:
            return self.device
        return torch.device("cuda", torch.cuda.current_device())

    def create_field(self, cardinality: int, dtype=None, fill_value=None, padded=False):
        """Allocate a (cardinality, *shape) field on the GPU (warp_grid.py:17-32).

        ``padded=True`` (population fields: cardinality 9, 19 or 27) lays the planes out with the
        library's plane stride: the tensor is a (q, nx, ny, nz) view with strides
        (plane, ny*nz, nz, 1) over storage that also holds the pad after each plane and, for the
        block format (dtype INT16), the segment exponents after the last plane.  Indexing is the
        same as for a contiguous field.  Copy such a field whole with ``copy_population``."""
        from xlb_amd.precision_policy import Precision

        if dtype is None:
            policy = DefaultConfig.default_precision_policy
            dtype = policy.store_precision if policy is not None else Precision.FP32
        tdt = dtype.torch_dtype if isinstance(dtype, Precision) else dtype
        shape = (int(cardinality),) + self.field_shape
        if padded or tdt == torch.int16:
            f = self._population(int(cardinality), tdt)
            if fill_value is not None:
                f.fill_(fill_value)
            return f
        if fill_value is None:
            return torch.zeros(shape, dtype=tdt, device=self._device())
        return torch.full(shape, fill_value, dtype=tdt, device=self._device())

    def _population(self, q, tdt):
        from xlb_amd import _lib

        lattice = {9: _lib.D2Q9, 19: _lib.D3Q19, 27: _lib.D3Q27}.get(q)
        store = {torch.float32: _lib.F32, torch.float64: _lib.F64, torch.float16: _lib.F16, torch.bfloat16: _lib.BF16,
                 torch.int16: _lib.I16_BLOCK}.get(tdt)
        if lattice is None or store is None:
            raise ValueError(f"padded fields are population fields: cardinality 9, 19 or 27 and a population dtype "
                             f"(got {q}, {tdt})")
        nx, ny, nz = self.nxyz
        L = _lib.lib()
        plane = int(L.xlbh_plane_stride(lattice, store, nx, ny, nz))
        nbytes = int(L.xlbh_population_bytes(lattice, store, nx, ny, nz, plane))
        if plane < 0 or nbytes < 0:
            raise ValueError(L.xlbh_last_error().decode())
        elem = torch.empty((), dtype=tdt).element_size()
        storage = torch.zeros((nbytes + elem - 1) // elem, dtype=tdt, device=self._device())
        return storage.as_strided((q, nx, ny, nz), (plane, ny * nz, nz, 1))
