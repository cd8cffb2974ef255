"""ctypes binding of libxlbhip.so (the C ABI declared in include/xlb_hip.h).

The library is built in-tree (``make -C xlb_amd/csrc``); a missing or stale build is an error,
never a fallback.  torch is imported first so the library binds to the HIP runtime torch already
loaded (both carry the soname libamdhip64.so.7) and shares its streams and allocations.
"""

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxlbhip.so")

# ABI constants (include/xlb_hip.h)
D2Q9, D3Q19, D3Q27 = 0, 1, 2
BGK, KBC, SMAGORINSKY_BGK = 0, 1, 2
F32, F64, F16, BF16, F16_SHIFTED, BF16_SHIFTED, I16_BLOCK = 0, 1, 2, 3, 4, 5, 6
COMPUTE_F32, COMPUTE_F64, COMPUTE_F32_DEVIATION = 0, 1, 2
FIELD_F32, FIELD_F64, FIELD_F16, FIELD_BF16 = 0, 1, 2, 3
BC_NONE = 0
BC_EQUILIBRIUM = 1
BC_HALFWAY_BB = 2
BC_FULLWAY_BB = 3
BC_DO_NOTHING = 4
BC_ZOUHE_VELOCITY = 5
BC_ZOUHE_PRESSURE = 6
BC_REGULARIZED_VELOCITY = 7
BC_REGULARIZED_PRESSURE = 8
BC_EXTRAPOLATION_OUTFLOW = 9
BC_GRADS_APPROXIMATION = 10
FORCE_NONE, FORCE_UNIFORM, FORCE_FIELD = 0, 1, 2
ADJ_RECOMPUTE, ADJ_BEGIN, ADJ_END = 1, 2, 4

ABI_VERSION = 4


class BcDesc(ctypes.Structure):
    """xlbh_bc_desc: one entry of the 256-entry boundary table (40 bytes)."""

    _fields_ = [
        ("kind", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("rho", ctypes.c_double),
        ("u", ctypes.c_double * 3),
    ]


_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_i64 = ctypes.c_int64
_dp = ctypes.POINTER(ctypes.c_double)
_bcp = ctypes.POINTER(BcDesc)

# name -> (restype, argtypes)
SIGNATURES = {
    "xlbh_last_error": (ctypes.c_char_p, []),
    "xlbh_abi_version": (_i, []),
    "xlbh_plane_stride": (_i64, [_i, _i, _i, _i, _i]),
    "xlbh_population_bytes": (_i64, [_i, _i, _i, _i, _i, _i64]),
    "xlbh_step": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _dp, _vp, _i, _i, _i, _i64, _d, _d, _i, _vp]),
    "xlbh_run": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _dp, _vp, _i, _i, _i, _i64, _d, _d, _i, _i, _vp]),
    "xlbh_step_slab": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _dp, _vp, _i, _i, _i, _i64, _d, _d, _i, _i,
                            _i, _vp]),
    "xlbh_stream": (_i, [_i, _i, _vp, _vp, _i, _i, _i, _i64, _vp]),
    "xlbh_equilibrium": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _i, _i, _i, _i64, _vp]),
    "xlbh_moments": (_i, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _i, _i, _i64, _vp]),
    "xlbh_collide": (_i, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _d, _d, _i, _dp, _vp, _i, _i, _i, _i64, _vp]),
    "xlbh_apply_bc": (_i, [_i, _i, _i, _i, _i, _bcp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i64, _vp]),
    "xlbh_mask_indices": (_i, [_i, _vp, _vp, _i, _i, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "xlbh_pack_missing": (_i, [_i, _vp, _vp, _i, _i, _i, _vp]),
    "xlbh_compact_workspace_bytes": (_i64, [_i64]),
    "xlbh_compact_cells": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "xlbh_scatter_aux": (_i, [_i, _vp, _vp, _i, _vp, _vp]),
    "xlbh_mesh_workspace_bytes": (_i64, [_i, _i, _i]),
    "xlbh_cast": (_i, [_i, _i, _vp, _vp, _i64, _vp]),
    "xlbh_vorticity": (_i, [_i, _vp, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "xlbh_q_criterion": (_i, [_i, _vp, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "xlbh_mesh_mask": (_i, [_i, _vp, _i64, _i, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "xlbh_reduce_workspace_bytes": (_i64, [_i64]),
    "xlbh_momentum_transfer": (_i, [_i, _i, _i, _i, _i, _bcp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i64, _vp]),
    "xlbh_momentum_transfer_noalloc": (_i, [_i, _i, _i, _i, _i, _bcp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i64, _vp]),
    "xlbh_checkpoint_encode": (_i, [_i, _i, _vp, _vp, _vp, _i, _i, _i, _i64, _vp]),
    "xlbh_checkpoint_decode": (_i, [_i, _i, _vp, _vp, _vp, _i, _i, _i, _i64, _vp]),
    "xlbh_mass_momentum": (_i, [_i, _i, _vp, _vp, _vp, _vp, _i, _i, _i, _i64, _vp]),
    "xlbh_adjoint_step": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _dp, _vp, _vp, _i, _i,
                               _i, _i64, _d, _d, _vp]),
    "xlbh_adjoint_run": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _dp, _vp, _vp, _i,
                              _i, _i, _i64, _d, _d, _i, _i, _vp, _i, _vp]),
    "xlbh_adjoint_workspace_bytes": (_i64, [_i, _i, _i, _i, _i]),
    "xlbh_record": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _dp, _vp, _i, _i, _i, _i64, _d, _d, _i, _i, _vp]),
}

_LIB = None


def lib():
    """Load (once) and return the bound library; raise if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built. Build it with `make -C {os.path.join(_HERE, 'csrc')} -j8` "
            "(or __graft_entry__.build()). There is no fallback implementation."
        )
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    v = L.xlbh_abi_version()
    if v != ABI_VERSION:
        raise RuntimeError(f"libxlbhip.so ABI version {v} != expected {ABI_VERSION}; rebuild it")
    _LIB = L
    return L


def check(rc, what="libxlbhip"):
    if rc != 0:
        msg = lib().xlbh_last_error().decode()
        raise RuntimeError(f"{what} failed (code {rc}): {msg}")


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def require_cuda(*tensors):
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise ValueError("MI355X backend fields must live on the GPU (device='cuda')")


def stream_of(t=None):
    """hipStream_t (as void*) of torch's current stream on the tensor's device."""
    dev = t.device if t is not None else torch.device("cuda", torch.cuda.current_device())
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def doubles(values):
    arr = (ctypes.c_double * 3)(*([float(v) for v in values] + [0.0] * (3 - len(values))))
    return ctypes.cast(arr, _dp), arr


def mark_written(*tensors):
    """Bump torch's version counter of tensors written through raw pointers."""
    bump = getattr(torch.autograd.graph, "increment_version", None)
    for t in tensors:
        if t is not None and bump is not None:
            bump(t)
