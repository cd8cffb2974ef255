"""Device-side helpers for boundary masks: packed missing bits (cached per mask tensor and
version) and ordered lists of the cells carrying one boundary id."""

import weakref

import numpy as np
import torch

from xlb_amd import _lib

_bits_cache = {}


def _n_cells(t):
    return int(t.shape[1]) * int(t.shape[2]) * int(t.shape[3])


def packed_missing(missing_mask, lattice_id):
    """(N,) int32 tensor whose bit l is missing_mask[l] (uint32 bit pattern), cached."""
    key = id(missing_mask)
    hit = _bits_cache.get(key)
    if hit is not None:
        ref, version, lat, bits = hit
        if ref() is missing_mask and version == missing_mask._version and lat == lattice_id:
            return bits
    _lib.require_cuda(missing_mask)
    nx, ny, nz = int(missing_mask.shape[1]), int(missing_mask.shape[2]), int(missing_mask.shape[3])
    bits = torch.empty(nx * ny * nz, dtype=torch.int32, device=missing_mask.device)
    mm = missing_mask if missing_mask.dtype in (torch.bool, torch.uint8) else (missing_mask != 0)
    mm = mm.contiguous()
    _lib.check(
        _lib.lib().xlbh_pack_missing(lattice_id, _lib.ptr(mm), _lib.ptr(bits), nx, ny, nz, _lib.stream_of(mm)),
        "pack_missing",
    )
    remember_bits(missing_mask, lattice_id, bits)
    return bits


def remember_bits(missing_mask, lattice_id, bits):
    for k in [k for k, v in _bits_cache.items() if v[0]() is None]:
        del _bits_cache[k]
    _bits_cache[id(missing_mask)] = (weakref.ref(missing_mask), missing_mask._version, lattice_id, bits)


def cells_with_id(bc_mask, bc_id):
    """Ordered (n,) int32 device tensor of flat cell indices with bc_mask == bc_id
    (bc_id = -1: every boundary cell, 0 < id < 255)."""
    _lib.require_cuda(bc_mask)
    nx, ny, nz = int(bc_mask.shape[1]), int(bc_mask.shape[2]), int(bc_mask.shape[3])
    N = nx * ny * nz
    L = _lib.lib()
    ws = torch.empty(int(L.xlbh_compact_workspace_bytes(N)), dtype=torch.uint8, device=bc_mask.device)
    cells = torch.empty(N, dtype=torch.int32, device=bc_mask.device)
    count = torch.zeros(1, dtype=torch.int32, device=bc_mask.device)
    _lib.check(
        L.xlbh_compact_cells(_lib.ptr(bc_mask), int(bc_id), _lib.ptr(cells), _lib.ptr(count), _lib.ptr(ws), nx, ny, nz,
                             _lib.stream_of(bc_mask)),
        "compact_cells",
    )
    n = int(count.item())
    return cells[:n]


def cell_coordinates(cells, shape, d):
    """Flat indices -> (d, n) int64 coordinates (host)."""
    nx, ny, nz = shape
    c = cells.detach().cpu().numpy().astype(np.int64)
    idx = np.stack([c // (ny * nz), (c // nz) % ny, c % nz])
    return idx[:d]
