"""The checkpoint codec (xlbh_checkpoint_encode / xlbh_checkpoint_decode) against a plain torch
FP64 restatement of the same encoding, and its round trip.

Encoding: moments rho - 1 and u rounded to float32; neq_l = bfloat16(f_l - feq_l(rho32, u32)) with
feq evaluated in FP64 from the float32 moments.  Decoding: f_l = feq_l(rho32, u32) + neq_l -
w_l (dr + 3 c_l.dj) with dr, dj the mass and momentum of the rounded neq, rounded to the store
format.  Bars: moments within one float32 ulp of the restatement (the two sum the populations in
different orders), neq within one bfloat16 ulp; the decoded field within 2^-8 (|neq_l| +
8 w_l sum_k |neq_k|) plus one ulp of the store format, with the kept moments.
"""

import numpy as np
import pytest
import torch

import xlb_amd as xa
from xlb_amd import _lib
from xlb_amd.autodiff import equilibrium_torch, macroscopic_torch
from xlb_amd.workloads import decaying_turbulence_3d, turbulence_velocity

pytestmark = pytest.mark.gpu

P = xa.PrecisionPolicy
LATTICES = {"D2Q9": (_lib.D2Q9, xa.velocity_set.D2Q9), "D3Q19": (_lib.D3Q19, xa.velocity_set.D3Q19),
            "D3Q27": (_lib.D3Q27, xa.velocity_set.D3Q27)}


def _field(lat, shape, dtype, seed):
    """A smooth flow plus a small non-equilibrium part, as a (q, nx, ny, nz) field."""
    _, cls = LATTICES[lat]
    vs = cls(P.FP64FP64, xa.ComputeBackend.MI355X)
    g = torch.Generator(device="cuda").manual_seed(seed)
    d = vs.d
    grid = [torch.arange(n, dtype=torch.float64, device="cuda") for n in shape]
    mesh = torch.meshgrid(*grid, indexing="ij")
    u = torch.stack([0.04 * torch.sin(2 * np.pi * (mesh[(i + 1) % len(shape)] + 0.5) / shape[(i + 1) % len(shape)] + i)
                     for i in range(d)])
    rho = 1 + 0.01 * torch.cos(2 * np.pi * mesh[0] / shape[0]).unsqueeze(0)
    f = equilibrium_torch(rho, u, vs)
    f = f * (1 + 1e-3 * torch.randn(f.shape, generator=g, device="cuda", dtype=torch.float64))
    if len(shape) == 2:
        f = f.unsqueeze(-1)
    return f.to(dtype).contiguous(), vs


def _codec(lat, store, f):
    lid, _ = LATTICES[lat]
    q, *ext = f.shape
    d = 2 if lat == "D2Q9" else 3
    neq = torch.empty((q, *ext), dtype=torch.bfloat16, device="cuda")
    mom = torch.empty((d + 1, *ext), dtype=torch.float32, device="cuda")
    L = _lib.lib()
    _lib.check(L.xlbh_checkpoint_encode(lid, store, _lib.ptr(f), _lib.ptr(neq), _lib.ptr(mom), *ext, f.stride(0), None))
    back = torch.empty_like(f)
    _lib.check(L.xlbh_checkpoint_decode(lid, store, _lib.ptr(neq), _lib.ptr(mom), _lib.ptr(back), *ext, back.stride(0),
                                        None))
    torch.cuda.synchronize()
    return neq, mom, back


@pytest.mark.parametrize("lat,shape,store", [("D2Q9", (96, 40), _lib.F32), ("D3Q19", (24, 20, 70), _lib.F32),
                                             ("D3Q27", (16, 12, 40), _lib.F32), ("D3Q19", (16, 16, 64), _lib.F64)])
def test_codec_matches_torch_restatement(lat, shape, store):
    dtype = torch.float64 if store == _lib.F64 else torch.float32
    f, vs = _field(lat, shape, dtype, seed=5)
    neq, mom, back = _codec(lat, store, f)
    fd = f.double().squeeze(-1) if len(shape) == 2 else f.double()
    rho, u = macroscopic_torch(fd, vs)
    m32 = torch.cat([rho - 1, u]).float()
    ulp32 = torch.abs(torch.nextafter(m32, torch.full_like(m32, np.inf)) - m32)
    got = mom.squeeze(-1) if len(shape) == 2 else mom
    assert torch.all(torch.abs(got - m32) <= ulp32 + 1e-15), "moments differ from the restatement by more than one float32 ulp"
    r64 = 1 + got[:1].double()
    feq = equilibrium_torch(r64, got[1:].double(), vs)
    want = (fd - feq).float().to(torch.bfloat16).float()
    have = (neq.squeeze(-1) if len(shape) == 2 else neq).float()
    bulp = torch.abs(want) * 2.0 ** -7 + 1e-12
    assert torch.all(torch.abs(have - want) <= bulp), "neq differs from the restatement by more than one bfloat16 ulp"
    # round trip: the moments survive, the non-equilibrium part to bfloat16 precision
    err = torch.abs(back.double() - f.double())
    store_ulp = torch.abs(f.double()) * (2.0 ** -23 if dtype == torch.float32 else 2.0 ** -52)
    neq_abs = torch.abs(fd - feq)
    w = torch.as_tensor(np.asarray(vs.w), dtype=torch.float64, device="cuda").view(-1, *([1] * (neq_abs.dim() - 1)))
    bound = ((neq_abs + 8 * w * neq_abs.sum(0, keepdim=True)) * 2.0 ** -8).reshape(err.shape) + store_ulp + 1e-30
    print(f"{lat} {tuple(shape)}: round trip max|df| {err.max().item():.2e}, max |neq| {have.abs().max().item():.2e}")
    assert torch.all(err <= bound)
    rb, ub = macroscopic_torch(back.double().squeeze(-1) if len(shape) == 2 else back.double(), vs)
    tol = 1e-6 if dtype == torch.float32 else 1e-12
    assert torch.allclose(rb - 1, got[:1].double(), rtol=0, atol=tol) and torch.allclose(ub, got[1:].double(), rtol=0, atol=tol)


def test_codec_keeps_turbulence_moments():
    """On the 128^3 decaying-turbulence field after 200 FP32 steps, decoding changes u by far less
    than rounding f - w_l to bfloat16 does."""
    n = 128
    wl = decaying_turbulence_3d(n, P.FP32FP32, omega=1.95)
    st, vs = wl.stepper, wl.stepper.velocity_set
    rho = torch.ones((1, n, n, n), dtype=torch.float64, device="cuda")
    f0 = equilibrium_torch(rho, turbulence_velocity(n, seed=3), vs).float().contiguous()
    f, _ = st.run(f0, torch.empty_like(f0), wl.bc_mask, wl.missing_mask, wl.omega, 200)
    _, _, back = _codec("D3Q19", _lib.F32, f)
    w = torch.as_tensor(np.asarray(vs.w), dtype=torch.float32, device="cuda").view(-1, 1, 1, 1)
    shifted = ((f - w).to(torch.bfloat16).float() + w)
    _, u = macroscopic_torch(f.double(), vs)
    _, u_codec = macroscopic_torch(back.double(), vs)
    _, u_shift = macroscopic_torch(shifted.double(), vs)
    e_codec = float((u_codec - u).norm() / u.norm())
    e_shift = float((u_shift - u).norm() / u.norm())
    print(f"rel L2 of u after decoding: codec {e_codec:.2e}, bfloat16 of f - w {e_shift:.2e}")
    assert e_codec <= 1e-6 and e_codec * 100 <= e_shift
