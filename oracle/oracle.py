"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the parity reference of the HIP backend and the CPU baseline timed by
bench.py.  Nothing in the product package imports this module.

The oracle takes plain NumPy arrays in the reference's layout: populations (q, nx, ny, nz) in the
store format (float32 / float64 / uint16 bit patterns for FP16 and BF16; int16 mantissas for the
block format, as the view at the start of a buffer that holds the segment exponents after the
planes, like the GPU fields), bc_mask (1, nx, ny, nz)
uint8, missing mask (q, nx, ny, nz) bool, and a list of BC descriptions
``dict(id=, kind=, rho=, u=)`` in the stepper's list order.
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# kinds (same numbering as the C ABI's XLBH_BC_*)
EQUILIBRIUM, HALFWAY_BB, FULLWAY_BB, DO_NOTHING = 1, 2, 3, 4
ZOUHE_VELOCITY, ZOUHE_PRESSURE, REGULARIZED_VELOCITY, REGULARIZED_PRESSURE = 5, 6, 7, 8
EXTRAPOLATION_OUTFLOW, GRADS_APPROXIMATION = 9, 10

# store formats
F32, F64, F16, BF16, F16_SHIFTED, BF16_SHIFTED, I16_BLOCK = 0, 1, 2, 3, 4, 5, 6


class OrcBc(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int32), ("kind", ctypes.c_int32), ("rho", ctypes.c_double), ("u", ctypes.c_double * 3)]


_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.orc_step.argtypes = [i, i, i, i, vp, vp, vp, vp, vp, i, i, vp, vp, i, i, i, d, d, vp]
        L.orc_aux_init.argtypes = [i, vp, vp, i, vp, i, i, i]
        L.orc_equilibrium.argtypes = [i, i, i, vp, vp, vp, i, i, i]
        L.orc_macroscopic.argtypes = [i, i, i, vp, vp, vp, i, i, i]
        L.orc_stream.argtypes = [i, i, vp, vp, i, i, i]
        L.orc_collide.argtypes = [i, i, i, i, vp, vp, vp, vp, vp, d, d, i, vp, vp, i, i, i]
        L.orc_apply_bc.argtypes = [i, i, i, vp, vp, vp, vp, vp, vp, i, i, i]
        L.orc_mask_indices.argtypes = [i, vp, vp, i, i, vp, vp, i, i, i]
        L.orc_mesh_mask.argtypes = [i, vp, ctypes.c_int64, i, vp, vp, i, i, i]
        L.orc_tri_box.argtypes = [vp, vp]
        L.orc_momentum_transfer.argtypes = [i, i, i, vp, vp, vp, vp, vp, i, i, i, vp]
        L.orc_lattice.argtypes = [i, vp, vp, vp, vp]
        L.orc_set_threads.argtypes = [i]
        L.orc_plant.argtypes = [i]
        L.orc_f32_to_f16.argtypes = [vp, vp, i]
        L.orc_f32_to_bf16.argtypes = [vp, vp, i]
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _bcs(bcs):
    arr = (OrcBc * max(1, len(bcs)))()
    for k, b in enumerate(bcs):
        arr[k].id = int(b["id"])
        arr[k].kind = int(b["kind"])
        arr[k].rho = float(b.get("rho", 0.0))
        u = list(b.get("u", (0.0, 0.0, 0.0))) + [0.0, 0.0, 0.0]
        for j in range(3):
            arr[k].u[j] = float(u[j])
    return arr


def _dims(f):
    return int(f.shape[1]), int(f.shape[2]), int(f.shape[3])


# Planted slips (orc_plant): each pinning test switches one on to show that its check fails.
PLANT_KBC_GAMMA_SIGN = 1
PLANT_KBC_SHEAR_INDEX = 2
PLANT_MOMENTUM_OPP_INDEX = 3
PLANT_MOMENTUM_NO_STREAM_TERM = 4
PLANT_EXTRAPOLATION_WEIGHTS = 5
PLANT_EXTRAPOLATION_SOURCE = 6


class planted:
    """``with planted(PLANT_...):`` runs the oracle with that transcription slip switched on."""

    def __init__(self, which):
        self.which = int(which)

    def __enter__(self):
        self.old = lib().orc_plant(self.which)
        return self

    def __exit__(self, *exc):
        lib().orc_plant(self.old)
        return False


def set_threads(n):
    return lib().orc_set_threads(int(n))


def lattice(lat):
    c = np.zeros((3, 27), dtype=np.int32)
    w = np.zeros(27)
    opp = np.zeros(27, dtype=np.int32)
    qi = np.zeros((27, 6))
    q = lib().orc_lattice(lat, _p(c), _p(w), _p(opp), _p(qi))
    return dict(c=c[:, :q], w=w[:q], opp=opp[:q], qi=qi[:q])


def step(lat, coll, store, compute, f0, f1, bc_mask, missing, bcs, omega, force=None, force_field=None, cparam=0.0,
         aux=None):
    """One fused step in place: reads f0 (receiving aux recovery writes), writes f1.  coll: 0 BGK,
    1 KBC, 2 Smagorinsky BGK with coefficient cparam.  aux: (N,) prescribed values for the block
    format (I16_BLOCK), whose int16 mantissas cannot carry the reference's f_1[0] stash."""
    nx, ny, nz = _dims(f0)
    bm = None if bc_mask is None else np.ascontiguousarray(bc_mask, dtype=np.uint8)
    mm = None if missing is None else np.ascontiguousarray(missing, dtype=np.uint8)
    arr = _bcs(bcs)
    mode, fu, ff = 0, None, None
    if force_field is not None:
        mode, ff = 2, np.ascontiguousarray(force_field, dtype=np.float64)
    elif force is not None:
        mode, fu = 1, np.ascontiguousarray(list(force) + [0.0] * (3 - len(force)), dtype=np.float64)
    if bm is None:
        bm = np.zeros((1, nx, ny, nz), dtype=np.uint8)
    ax = None if aux is None else np.ascontiguousarray(np.broadcast_to(aux, (nx * ny * nz,)), dtype=np.float64)
    rc = lib().orc_step(lat, coll, store, compute, _p(f0), _p(f1), _p(bm), _p(mm), arr, len(bcs), mode, _p(fu), _p(ff),
                        nx, ny, nz, float(omega), float(cparam), _p(ax))
    assert rc == 0


def aux_init(store, f1, bc_mask, bc_id, values):
    nx, ny, nz = _dims(f1)
    v = np.ascontiguousarray(np.broadcast_to(values, (nx * ny * nz,)), dtype=np.float64)
    bm = np.ascontiguousarray(bc_mask, dtype=np.uint8)
    assert lib().orc_aux_init(store, _p(f1), _p(bm), int(bc_id), _p(v), nx, ny, nz) == 0


def equilibrium(lat, store, compute, rho, u, f):
    nx, ny, nz = _dims(f)
    assert lib().orc_equilibrium(lat, store, compute, _p(np.ascontiguousarray(rho, dtype=np.float64)),
                                 _p(np.ascontiguousarray(u, dtype=np.float64)), _p(f), nx, ny, nz) == 0
    return f


def macroscopic(lat, store, compute, f, d):
    nx, ny, nz = _dims(f)
    rho = np.zeros((1, nx, ny, nz))
    u = np.zeros((d, nx, ny, nz))
    assert lib().orc_macroscopic(lat, store, compute, _p(np.ascontiguousarray(f)), _p(rho), _p(u), nx, ny, nz) == 0
    return rho, u


def stream(lat, store, fin):
    fout = np.empty_like(fin)
    nx, ny, nz = _dims(fin)
    assert lib().orc_stream(lat, store, _p(np.ascontiguousarray(fin)), _p(fout), nx, ny, nz) == 0
    return fout


def collide(lat, coll, store, compute, f, feq, rho, u, omega, force=None, force_field=None, cparam=0.0):
    fout = np.empty_like(f)
    nx, ny, nz = _dims(f)
    mode, fu, ff = 0, None, None
    if force_field is not None:
        mode, ff = 2, np.ascontiguousarray(force_field, dtype=np.float64)
    elif force is not None:
        mode, fu = 1, np.ascontiguousarray(list(force) + [0.0] * (3 - len(force)), dtype=np.float64)
    r = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
    uu = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
    assert lib().orc_collide(lat, coll, store, compute, _p(np.ascontiguousarray(f)), _p(np.ascontiguousarray(feq)), _p(fout),
                             _p(r), _p(uu), float(omega), float(cparam), mode, _p(fu), _p(ff), nx, ny, nz) == 0
    return fout


def apply_bc(lat, store, compute, bc, fpre, fpost, bc_mask, missing, aux=None):
    nx, ny, nz = _dims(fpost)
    out = np.ascontiguousarray(fpost).copy()
    a = None if aux is None else np.ascontiguousarray(np.broadcast_to(aux, (nx * ny * nz,)), dtype=np.float64)
    arr = _bcs([bc])
    assert lib().orc_apply_bc(lat, store, compute, arr, _p(np.ascontiguousarray(fpre)), _p(out),
                              _p(np.ascontiguousarray(bc_mask, dtype=np.uint8)),
                              _p(np.ascontiguousarray(missing, dtype=np.uint8)), _p(a), nx, ny, nz) == 0
    return out


def mask_indices(lat, indices, interior, bc_id, bc_mask, missing):
    """In place on bc_mask (1,nx,ny,nz) uint8 and missing (q,nx,ny,nz) uint8."""
    nx, ny, nz = _dims(bc_mask)
    idx = np.ascontiguousarray(indices, dtype=np.int32)
    n = idx.shape[1]
    assert lib().orc_mask_indices(lat, _p(idx), _p(np.ascontiguousarray(interior, dtype=np.uint8)), n, int(bc_id),
                                  _p(bc_mask), _p(missing), nx, ny, nz) == 0


def mesh_mask(lat, triangles, bc_id, bc_mask, missing):
    """MeshBoundaryMasker restatement, in place on bc_mask (1,nx,ny,nz) uint8 and missing
    (q,nx,ny,nz) uint8; triangles (T, 3, 3)."""
    nx, ny, nz = _dims(bc_mask)
    tri = np.ascontiguousarray(triangles, dtype=np.float32).reshape(-1, 9)
    assert lib().orc_mesh_mask(lat, _p(tri), ctypes.c_int64(tri.shape[0]), int(bc_id), _p(bc_mask), _p(missing),
                               nx, ny, nz) == 0


def tri_box(tri9, p):
    """Schwarz-Seidel triangle / unit-box test of the oracle (box minimum corner p)."""
    t = np.ascontiguousarray(tri9, dtype=np.float32).reshape(9)
    q = np.ascontiguousarray(p, dtype=np.float32).reshape(3)
    return bool(lib().orc_tri_box(_p(t), _p(q)))


def momentum_transfer(lat, store, compute, bc, f0, bc_mask, missing, aux=None):
    nx, ny, nz = _dims(f0)
    out = np.zeros(3)
    a = None if aux is None else np.ascontiguousarray(np.broadcast_to(aux, (nx * ny * nz,)), dtype=np.float64)
    assert lib().orc_momentum_transfer(lat, store, compute, _bcs([bc]), _p(np.ascontiguousarray(f0)),
                                       _p(np.ascontiguousarray(bc_mask, dtype=np.uint8)),
                                       _p(np.ascontiguousarray(missing, dtype=np.uint8)), _p(a), nx, ny, nz, _p(out)) == 0
    return out


def f32_to_f16_bits(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint16)
    lib().orc_f32_to_f16(_p(x), _p(out), x.size)
    return out


def f32_to_bf16_bits(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint16)
    lib().orc_f32_to_bf16(_p(x), _p(out), x.size)
    return out


def bf16_bits_to_f32(b):
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)
