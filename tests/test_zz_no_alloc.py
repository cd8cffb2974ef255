"""xlbh_momentum_transfer_noalloc (no allocation: its partial sums live in a per-device array the
library holds) against xlbh_momentum_transfer with a caller-allocated workspace: bit-identical
forces on two calls in a row, the second on a smaller grid, through the C ABI."""

import numpy as np
import pytest
import torch

import xlb_amd as xa
from xlb_amd import _lib
from xlb_amd.operator.boundary_masker.packing import packed_missing
from xlb_amd.workloads import wind_tunnel_sphere

pytestmark = pytest.mark.gpu


def _forces(shape, radius, steps):
    wl = wind_tunnel_sphere(shape, xa.PrecisionPolicy.FP32FP32, radius=radius, center=(shape[0] // 4, shape[1] // 2,
                                                                                      shape[2] // 2))
    st = wl.stepper
    f0, f1 = st.run(wl.f_0, wl.f_1, wl.bc_mask, wl.missing_mask, wl.omega, steps)
    sphere = wl.extra["sphere_bc"]
    bits = packed_missing(wl.missing_mask, st.lattice_id)
    nx, ny, nz = shape
    L = _lib.lib()
    ws = torch.empty(int(L.xlbh_reduce_workspace_bytes(nx * ny * nz)), dtype=torch.uint8, device="cuda")
    with_ws = torch.zeros(4, dtype=torch.float64, device="cuda")
    no_ws = torch.full((4,), float("nan"), dtype=torch.float64, device="cuda")
    args = (st.lattice_id, st.precision_policy.store_format, st.compute_id, sphere.kind, int(sphere.id),
            sphere.bc_desc(), _lib.ptr(f0), _lib.ptr(wl.bc_mask), _lib.ptr(bits), None)
    s = _lib.stream_of(f0)
    _lib.check(L.xlbh_momentum_transfer(*args, _lib.ptr(with_ws), _lib.ptr(ws), nx, ny, nz, f0.stride(0), s))
    _lib.check(L.xlbh_momentum_transfer_noalloc(*args, _lib.ptr(no_ws), nx, ny, nz, f0.stride(0), s))
    torch.cuda.synchronize()
    return with_ws[:3].cpu().numpy(), no_ws[:3].cpu().numpy()


def test_noalloc_momentum_transfer_matches_the_workspace_entry_point():
    for shape, radius, steps in (((96, 48, 48), 6, 60), ((48, 32, 24), 4, 30)):
        a, b = _forces(shape, radius, steps)
        print(f"{shape}: drag with workspace {a}, without {b}")
        assert np.all(np.isfinite(a)) and abs(a[0]) > 0
        np.testing.assert_array_equal(a, b)


def test_noalloc_rejects_grids_beyond_its_scratch():
    L = _lib.lib()
    rc = L.xlbh_momentum_transfer_noalloc(_lib.D3Q19, _lib.F32, _lib.COMPUTE_F32, _lib.BC_HALFWAY_BB, 1, None,
                                          _lib.ptr(torch.zeros(1, device="cuda")), None, None, None, None,
                                          4096, 256, 257, 0, None)
    assert rc != 0 and b"2^28" in L.xlbh_last_error()
