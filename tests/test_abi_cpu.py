"""The C ABI library builds, loads without a GPU, and exports every entry point declared in
include/xlb_hip.h with the signature the Python binding uses."""

import ctypes
import os
import re

from conftest import ROOT


def header_functions():
    text = open(os.path.join(ROOT, "include", "xlb_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xlbh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = header_functions()
    for required in ("xlbh_step", "xlbh_run", "xlbh_stream", "xlbh_equilibrium", "xlbh_moments", "xlbh_collide",
                     "xlbh_apply_bc", "xlbh_mask_indices", "xlbh_pack_missing", "xlbh_momentum_transfer",
                     "xlbh_mass_momentum", "xlbh_last_error"):
        assert required in names


def test_library_exports_every_header_symbol():
    from xlb_amd import _lib

    L = _lib.lib()
    raw = ctypes.CDLL(_lib.LIB_PATH)
    for name in header_functions():
        assert hasattr(raw, name), f"{name} declared in xlb_hip.h but not exported"
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature in xlb_amd/_lib.py"
        getattr(L, name)


def test_error_channel_and_validation_without_gpu():
    """Argument validation happens before any HIP call, so it runs on a machine without a GPU."""
    from xlb_amd import _lib

    L = _lib.lib()
    assert L.xlbh_abi_version() == _lib.ABI_VERSION
    # in-place step is rejected
    buf = ctypes.c_void_p(0x1000)
    rc = L.xlbh_step(_lib.D3Q19, _lib.BGK, _lib.F32, _lib.COMPUTE_F32, buf, buf, None, None, None, None,
                     _lib.FORCE_NONE, None, None, 8, 8, 8, 0, 1.0, 0.0, 0, None)
    assert rc != 0 and b"in place" in L.xlbh_last_error()
    # 2-D lattice with nz > 1 is rejected
    rc = L.xlbh_step(_lib.D2Q9, _lib.BGK, _lib.F32, _lib.COMPUTE_F32, ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000),
                     None, None, None, None, _lib.FORCE_NONE, None, None, 8, 8, 4, 0, 1.0, 0.0, 0, None)
    assert rc != 0 and b"2-D" in L.xlbh_last_error()
    # KBC on D3Q19 is rejected (kbc.py:183-185)
    rc = L.xlbh_collide(_lib.D3Q19, _lib.KBC, _lib.F32, _lib.COMPUTE_F32, _lib.FIELD_F32, ctypes.c_void_p(8),
                        ctypes.c_void_p(16), ctypes.c_void_p(24), ctypes.c_void_p(32), ctypes.c_void_p(40), 1.0,
                        0.0, _lib.FORCE_NONE, None, None, 4, 4, 4, 0, None)
    assert rc != 0 and b"KBC" in L.xlbh_last_error()
    # unknown lattice
    rc = L.xlbh_stream(7, _lib.F32, ctypes.c_void_p(8), ctypes.c_void_p(16), 4, 4, 4, 0, None)
    assert rc != 0 and b"lattice" in L.xlbh_last_error()
    # grid limits: empty extents, nx above the launch grid's z range, more than 2^32 - 256 cells
    a, b = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)
    # (tiles of y rows: one row per tile once nz > 128, so ny = 65536 there is one tile too many)
    for (nx, ny, nz), msg in [((0, 8, 8), b"positive"), ((8, 8, -1), b"positive"), ((65536, 2, 2), b"65535"),
                              ((65535, 65536, 2), b"2^32 - 256"), ((2, 65536, 129), b"tiles of y rows"),
                              ((2, 262144, 64), b"tiles of y rows")]:
        rc = L.xlbh_step(_lib.D3Q19, _lib.BGK, _lib.F32, _lib.COMPUTE_F32, a, b, None, None, None, None,
                         _lib.FORCE_NONE, None, None, nx, ny, nz, 0, 1.0, 0.0, 0, None)
        assert rc != 0 and msg in L.xlbh_last_error(), (nx, ny, nz, L.xlbh_last_error())
    # the slab step: its planes need a ghost plane on each side; the block format is refused
    for xb, xe in ((0, 4), (3, 3), (5, 2), (1, 8)):
        rc = L.xlbh_step_slab(_lib.D3Q19, _lib.BGK, _lib.F32, _lib.COMPUTE_F32, a, b, None, None, None, None,
                              _lib.FORCE_NONE, None, None, 8, 4, 4, 0, 1.0, 0.0, 0, xb, xe, None)
        assert rc != 0 and b"ghost plane" in L.xlbh_last_error(), (xb, xe, L.xlbh_last_error())
    rc = L.xlbh_step_slab(_lib.D3Q19, _lib.BGK, _lib.I16_BLOCK, _lib.COMPUTE_F32, a, b, None, None, None, None,
                          _lib.FORCE_NONE, None, None, 8, 4, 4, 0, 1.0, 0.0, 0, 1, 7, None)
    assert rc != 0 and b"block format" in L.xlbh_last_error()
    # the post-processing stencils sweep the same launch grid and take the same bounds
    for fn in (L.xlbh_vorticity, L.xlbh_q_criterion):
        for (nx, ny, nz), msg in [((65536, 3, 3), b"65535"), ((3, 65536, 129), b"tiles of y rows")]:
            rc = fn(_lib.FIELD_F32, a, None, b, ctypes.c_void_p(0x3000), nx, ny, nz, None)
            assert rc != 0 and msg in L.xlbh_last_error(), (nx, ny, nz, L.xlbh_last_error())
    # the largest 2-D grid that fits passes validation: 65535 x 65535 cells = 2^32 - 2^17 + 1
    rc = L.xlbh_run(_lib.D2Q9, _lib.BGK, _lib.F16, _lib.COMPUTE_F32, a, b, None, None, None, None,
                    _lib.FORCE_NONE, None, None, 65535, 65535, 1, 0, 1.0, 0.0, 0, 0, None)
    assert rc == 0, L.xlbh_last_error()
    # negative step counts
    rc = L.xlbh_run(_lib.D3Q19, _lib.BGK, _lib.F32, _lib.COMPUTE_F32, a, b, None, None, None, None,
                    _lib.FORCE_NONE, None, None, 4, 4, 4, 0, 1.0, 0.0, 0, -1, None)
    assert rc != 0 and b"n_steps" in L.xlbh_last_error()
    # a boundary mask without its packed bits and table
    rc = L.xlbh_step(_lib.D3Q19, _lib.BGK, _lib.F32, _lib.COMPUTE_F32, a, b, ctypes.c_void_p(0x3000), None, None,
                     None, _lib.FORCE_NONE, None, None, 4, 4, 4, 0, 1.0, 0.0, 0, None)
    assert rc != 0 and b"missing bits" in L.xlbh_last_error()
    # plane strides: 0 (dense) or >= N; the block format computes in FP32 only
    rc = L.xlbh_run(_lib.D3Q19, _lib.BGK, _lib.F32, _lib.COMPUTE_F32, a, b, None, None, None, None,
                    _lib.FORCE_NONE, None, None, 4, 4, 4, 63, 1.0, 0.0, 0, 1, None)
    assert rc != 0 and b"plane stride" in L.xlbh_last_error()
    rc = L.xlbh_run(_lib.D3Q19, _lib.BGK, _lib.F32, _lib.COMPUTE_F32, a, b, None, None, None, None,
                    _lib.FORCE_NONE, None, None, 4, 4, 4, 64, 1.0, 0.0, 0, 0, None)
    assert rc == 0, L.xlbh_last_error()
    rc = L.xlbh_run(_lib.D3Q19, _lib.BGK, _lib.I16_BLOCK, _lib.COMPUTE_F64, a, b, None, None, None, None,
                    _lib.FORCE_NONE, None, None, 4, 4, 4, 0, 1.0, 0.0, 0, 1, None)
    assert rc != 0 and b"unsupported" in L.xlbh_last_error()
    # the block format is refused where its segment exponents cannot follow the values
    rc = L.xlbh_stream(_lib.D3Q19, _lib.I16_BLOCK, ctypes.c_void_p(8), ctypes.c_void_p(16), 4, 4, 4, 0, None)
    assert rc != 0 and b"block format" in L.xlbh_last_error()
    # workspace sizes
    assert L.xlbh_reduce_workspace_bytes(256) == 4 * 8
    assert L.xlbh_compact_workspace_bytes(512) == 2 * 2 * 4


def test_bc_desc_layout_matches_header():
    from xlb_amd import _lib

    assert ctypes.sizeof(_lib.BcDesc) == 40
    assert _lib.BcDesc.rho.offset == 8 and _lib.BcDesc.u.offset == 16


def test_largest_y_extent_passes_validation():
    """65535 tiles of y rows are accepted: ny = 65535 with nz > 128 (one row per tile) and
    ny = 4 * 65535 with nz <= 64 (four rows per tile); validation runs before any HIP call."""
    from xlb_amd import _lib

    L = _lib.lib()
    a, b = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)
    for nx, ny, nz in ((1, 65535, 129), (1, 4 * 65535, 2)):
        rc = L.xlbh_run(_lib.D3Q19, _lib.BGK, _lib.F32, _lib.COMPUTE_F32, a, b, None, None, None, None,
                        _lib.FORCE_NONE, None, None, nx, ny, nz, 0, 1.0, 0.0, 0, 0, None)
        assert rc == 0, L.xlbh_last_error()


def test_population_layout_sizes():
    """xlbh_population_bytes: Q * plane * S for the plain formats; the block format adds one int32
    exponent per row segment after the planes, at a 256-byte aligned offset.  xlbh_plane_stride
    returns a stride >= N."""
    from xlb_amd import _lib

    L = _lib.lib()
    for lat, q, (nx, ny, nz) in ((_lib.D3Q19, 19, (8, 6, 300)), (_lib.D2Q9, 9, (5, 70, 1)), (_lib.D3Q27, 27, (3, 4, 5))):
        n = nx * ny * nz
        for store, s in ((_lib.F32, 4), (_lib.F64, 8), (_lib.F16, 2), (_lib.BF16_SHIFTED, 2)):
            p = L.xlbh_plane_stride(lat, store, nx, ny, nz)
            assert p >= n
            assert L.xlbh_population_bytes(lat, store, nx, ny, nz, 0) == q * n * s
            assert L.xlbh_population_bytes(lat, store, nx, ny, nz, p) == q * p * s
        fast = nz if nz > 1 else ny
        bz = next(b for b in (64, 128, 256) if b >= fast) if fast <= 256 else 256
        segs = (nx * ny if nz > 1 else nx) * (-(-fast // bz))
        off = -(-(q * n * 2) // 256) * 256
        assert L.xlbh_population_bytes(lat, _lib.I16_BLOCK, nx, ny, nz, 0) == off + 4 * segs
    assert L.xlbh_population_bytes(_lib.D3Q19, _lib.F32, 4, 4, 4, 63) == -1


def test_recommended_plane_stride():
    """xlbh_plane_stride: N for 4- and 8-byte populations and N + 128 for 2-byte ones, except for
    planes of more than 16 MiB that exceed a power of two by less than 1 MiB (an x-slab of a
    power-of-two share with its two ghost planes), which are padded to 2^k + 1 MiB."""
    from xlb_amd import _lib

    L = _lib.lib()
    MiB = 1 << 20
    n3 = 256**3
    assert L.xlbh_plane_stride(_lib.D3Q19, _lib.F32, 256, 256, 256) == n3
    assert L.xlbh_plane_stride(_lib.D3Q19, _lib.F16_SHIFTED, 256, 256, 256) == n3 + 128
    assert L.xlbh_plane_stride(_lib.D3Q19, _lib.F64, 240, 240, 240) == 240**3
    # the slab of a 256^3 share: 64 MiB + 512 KiB of FP32, 32 MiB + 256 KiB of FP16
    assert L.xlbh_plane_stride(_lib.D3Q19, _lib.F32, 258, 256, 256) == (64 * MiB + MiB) // 4
    assert L.xlbh_plane_stride(_lib.D3Q19, _lib.F16_SHIFTED, 258, 256, 256) == (32 * MiB + MiB) // 2
    assert L.xlbh_plane_stride(_lib.D3Q19, _lib.F64, 130, 128, 128) == (16 * MiB + MiB) // 8
    # exceeding the power of two by 1 MiB or more, or small planes: unchanged
    assert L.xlbh_plane_stride(_lib.D3Q19, _lib.F32, 260, 256, 256) == 260 * 256 * 256
    assert L.xlbh_plane_stride(_lib.D3Q19, _lib.F32, 130, 128, 128) == 130 * 128 * 128
