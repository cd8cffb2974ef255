"""Pins the CPU oracle before it is trusted as the parity reference.

Checks come from (a) the reference's own known-answer tests (tests/kernels/*, tests/boundary_conditions/*),
(b) lattice identities, (c) conservation laws and closed-form flows (plane Poiseuille, Taylor-Green
decay), and (d) special cases that reduce to a library routine (numpy.roll, numpy/torch rounding).
Each is chosen so that a plausible slip -- a dropped term, a wrong sign or index, a transposed
operand -- fails at least one of them.
"""

import itertools

import numpy as np
import pytest
import torch

from oracle import oracle as orc

LATS = {0: (2, 9), 1: (3, 19), 2: (3, 27)}


def ref_lattice(lat):
    """The reference's constructors (d2q9.py:16-24, d3q19.py:17-30, d3q27.py:17-32)."""
    if lat == 0:
        cx = [0, 0, 0, 1, -1, 1, -1, 1, -1]
        cy = [0, 1, -1, 0, 1, -1, 0, 1, -1]
        c = np.array(tuple(zip(cx, cy))).T
        w = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 9, 1 / 36, 1 / 36])
        return c, w
    full = list(itertools.product([0, -1, 1], repeat=3))
    if lat == 1:
        c = np.array([ci for ci in full if np.sum(np.abs(ci)) <= 2]).T
        w = np.array([{0: 1 / 3, 1: 1 / 18, 2: 1 / 36}[int(np.sum(np.abs(c[:, i])))] for i in range(19)])
    else:
        c = np.array(full).T
        w = np.array([{0: 8 / 27, 1: 2 / 27, 2: 1 / 54, 3: 1 / 216}[int(np.sum(np.abs(c[:, i])))] for i in range(27)])
    return c, w


def grid_shape(lat, n):
    return (n, n, 1) if lat == 0 else (n, n, n)


@pytest.mark.parametrize("lat", [0, 1, 2])
def test_lattice_tables(lat):
    d, q = LATS[lat]
    t = orc.lattice(lat)
    c, w = ref_lattice(lat)
    assert t["c"][:d].tolist() == c.tolist()
    np.testing.assert_allclose(t["w"], w, rtol=0, atol=1e-16)
    opp = t["opp"]
    assert (c[:, opp] == -c).all() and (opp[opp] == np.arange(q)).all()
    # isotropy: sum w = 1, sum w c_a c_b = delta/3, odd moments vanish
    assert abs(w.sum() - 1) < 1e-15
    np.testing.assert_allclose((w * c) @ c.T, np.eye(d) / 3, atol=1e-15)
    np.testing.assert_allclose(w @ c.T, 0, atol=1e-16)
    # qi = cc - delta/3 with off-diagonal doubled
    k = 0
    for a in range(d):
        for b in range(a, d):
            expect = c[a] * c[b] - (1 / 3 if a == b else 0)
            if a != b:
                expect = 2 * expect
            np.testing.assert_allclose(t["qi"][:, k], expect, atol=1e-15)
            k += 1


def test_f16_bf16_rounding_matches_numpy_and_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-9, 5, 20000),
        np.array([0.0, -0.0, 65504, 65519.99, 65520, 1e-8, 2.0**-24, 2.0**-25, 1.5 * 2.0**-24, np.inf, -np.inf], np.float32),
    ]).astype(np.float32)
    with np.errstate(over="ignore"):
        ref16 = x.astype(np.float16).view(np.uint16)
    assert (orc.f32_to_f16_bits(x) == ref16).all()
    tb = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert (orc.f32_to_bf16_bits(x) == tb).all()


# ------------------------------------------------------------------ reference known answers
@pytest.mark.parametrize("lat,n", [(0, 50), (1, 20), (2, 12)])
def test_equilibrium_rest_state_is_weights(lat, n):
    # tests/kernels/equilibrium/test_equilibrium_warp.py: feq(1, 0) = w, sum = 1
    d, q = LATS[lat]
    shp = grid_shape(lat, n)
    f = np.zeros((q,) + shp)
    orc.equilibrium(lat, orc.F64, 1, np.ones(shp), np.zeros((d,) + shp), f)
    _, w = ref_lattice(lat)
    np.testing.assert_allclose(f, np.broadcast_to(w.reshape(-1, 1, 1, 1), f.shape), atol=1e-16)
    np.testing.assert_allclose(f.sum(0), 1.0, atol=1e-15)


@pytest.mark.parametrize("lat,rho,vel", [(0, 1.0, 0.0), (0, 1.1, 1.0), (0, 1.1, 2.0), (1, 1.0, 0.0), (1, 1.1, 1.0), (1, 1.1, 2.0), (2, 1.1, 1.0)])
def test_macroscopic_inverts_equilibrium(lat, rho, vel):
    # tests/kernels/macroscopic/test_macroscopic_warp.py (atol 1e-6 on u), FP32 store/compute
    d, q = LATS[lat]
    shp = grid_shape(lat, 10)
    f = np.zeros((q,) + shp, dtype=np.float32)
    orc.equilibrium(lat, orc.F32, 0, np.full(shp, rho), np.full((d,) + shp, vel), f)
    r, u = orc.macroscopic(lat, orc.F32, 0, f, d)
    np.testing.assert_allclose(r, rho, rtol=1e-6)
    np.testing.assert_allclose(u, vel, atol=1e-6)


@pytest.mark.parametrize("lat,omega", [(0, 0.6), (0, 1.0), (1, 0.6), (1, 1.0), (2, 1.0)])
def test_bgk_closed_form(lat, omega):
    # tests/kernels/collision/test_bgk_collision_warp.py: fout = f - omega (f - feq)
    d, q = LATS[lat]
    shp = grid_shape(lat, 8)
    rng = np.random.default_rng(1)
    f = rng.random((q,) + shp).astype(np.float32)
    feq = rng.random((q,) + shp).astype(np.float32)
    out = orc.collide(lat, 0, orc.F32, 0, f, feq, None, None, omega)
    np.testing.assert_allclose(out, f - np.float32(omega) * (f - feq), atol=1e-6)


@pytest.mark.parametrize("lat", [0, 1, 2])
def test_stream_is_roll(lat):
    # tests/kernels/stream/test_stream_warp.py: stream = np.roll(f_i, c_i)
    d, q = LATS[lat]
    shp = grid_shape(lat, 9) if lat else (11, 7, 1)
    rng = np.random.default_rng(2)
    f = rng.random((q,) + shp)
    out = orc.stream(lat, orc.F64, f)
    c, _ = ref_lattice(lat)
    for i in range(q):
        shift = tuple(int(v) for v in c[:, i]) + ((0,) if d == 2 else ())
        np.testing.assert_array_equal(out[i], np.roll(f[i], shift, axis=(0, 1, 2)))


def _sphere(n, d):
    x = np.arange(n)
    grids = np.meshgrid(*([x] * d), indexing="xy")
    r2 = sum((g - n // 2) ** 2 for g in grids)
    return np.where(r2 < (n // 4) ** 2)


@pytest.mark.parametrize("lat,n", [(0, 50), (1, 20), (2, 16)])
def test_masker_no_padding_marks_exactly_the_indices(lat, n):
    # tests/boundary_conditions/mask/test_bc_indices_masker_warp.py (FullwayBB: no padding)
    d, q = LATS[lat]
    shp = grid_shape(lat, n)
    ind = _sphere(n, d)
    idx = np.array(ind + ((np.zeros(len(ind[0]), int),) if d == 2 else ()))
    bc_mask = np.zeros((1,) + shp, np.uint8)
    miss = np.zeros((q,) + shp, np.uint8)
    orc.mask_indices(lat, idx, np.zeros(idx.shape[1], np.uint8), 5, bc_mask, miss)
    expect = np.zeros(shp, np.uint8)
    expect[tuple(idx)] = 5
    np.testing.assert_array_equal(bc_mask[0], expect)
    assert miss.sum() == 0  # sphere strictly interior: no pull leaves the domain


@pytest.mark.parametrize("lat", [1, 2])
def test_masker_padding_marks_solid_neighbours(lat):
    """Halfway-BB semantics (indices_boundary_masker.py:148-156): missing[l, x] iff x - c_l is solid."""
    d, q = LATS[lat]
    n = 12
    shp = (n, n, n)
    ind = _sphere(n, 3)
    idx = np.array(ind)
    solid = np.zeros(shp, bool)
    solid[tuple(idx)] = True
    bc_mask = np.zeros((1,) + shp, np.uint8)
    miss = np.zeros((q,) + shp, np.uint8)
    orc.mask_indices(lat, idx, np.ones(idx.shape[1], np.uint8), 3, bc_mask, miss)
    c, _ = ref_lattice(lat)
    for l in range(q):
        expect = np.roll(solid, tuple(int(v) for v in c[:, l]), axis=(0, 1, 2))  # solid at x - c_l
        np.testing.assert_array_equal(miss[l].astype(bool), expect)
    dil = np.zeros(shp, bool)
    for l in range(q):
        dil |= np.roll(solid, tuple(int(v) for v in c[:, l]), axis=(0, 1, 2))
    np.testing.assert_array_equal(bc_mask[0] == 3, dil)


@pytest.mark.parametrize("lat", [0, 1])
def test_equilibrium_bc_sets_weights_inside_only(lat):
    # tests/boundary_conditions/bc_equilibrium/test_bc_equilibrium_warp.py
    d, q = LATS[lat]
    n = 20
    shp = grid_shape(lat, n)
    ind = _sphere(n, d)
    idx = np.array(ind + ((np.zeros(len(ind[0]), int),) if d == 2 else ()))
    bc_mask = np.zeros((1,) + shp, np.uint8)
    miss = np.zeros((q,) + shp, np.uint8)
    orc.mask_indices(lat, idx, np.zeros(idx.shape[1], np.uint8), 1, bc_mask, miss)
    fpre = np.zeros((q,) + shp, np.float32)
    fpost = np.full((q,) + shp, 2.0, np.float32)
    out = orc.apply_bc(lat, orc.F32, 0, dict(id=1, kind=orc.EQUILIBRIUM, rho=1.0, u=(0, 0, 0)), fpre, fpost, bc_mask, miss)
    _, w = ref_lattice(lat)
    inside = bc_mask[0] == 1
    for i in range(q):
        np.testing.assert_allclose(out[i][inside], w[i], rtol=1e-7)
        np.testing.assert_array_equal(out[i][~inside], 2.0)


@pytest.mark.parametrize("lat", [0, 1, 2])
def test_fullway_bounce_back_flips_inside_only(lat):
    # tests/boundary_conditions/bc_fullway_bounce_back/*: the flip the reference computes but never asserts
    d, q = LATS[lat]
    n = 16
    shp = grid_shape(lat, n)
    ind = _sphere(n, d)
    idx = np.array(ind + ((np.zeros(len(ind[0]), int),) if d == 2 else ()))
    bc_mask = np.zeros((1,) + shp, np.uint8)
    miss = np.zeros((q,) + shp, np.uint8)
    orc.mask_indices(lat, idx, np.zeros(idx.shape[1], np.uint8), 2, bc_mask, miss)
    rng = np.random.default_rng(3)
    fpre = rng.random((q,) + shp).astype(np.float32)
    fpost = np.full((q,) + shp, 2.0, np.float32)
    out = orc.apply_bc(lat, orc.F32, 0, dict(id=2, kind=orc.FULLWAY_BB), fpre, fpost, bc_mask, miss)
    opp = orc.lattice(lat)["opp"]
    inside = bc_mask[0] == 2
    for i in range(q):
        np.testing.assert_array_equal(out[i][inside], fpre[opp[i]][inside])
        np.testing.assert_array_equal(out[i][~inside], 2.0)


# ------------------------------------------------------------------ conservation and closed forms
def _run(lat, coll, f0, omega, steps, bc_mask=None, miss=None, bcs=(), force=None, store=orc.F64, compute=1):
    f0 = f0.copy()
    f1 = f0.copy()
    for _ in range(steps):
        orc.step(lat, coll, store, compute, f0, f1, bc_mask, miss, list(bcs), omega, force=force)
        f0, f1 = f1, f0
    return f0


@pytest.mark.parametrize("lat,coll", [(1, 0), (2, 1), (0, 1), (0, 0)])
def test_periodic_mass_and_momentum_conserved(lat, coll):
    d, q = LATS[lat]
    shp = grid_shape(lat, 10)
    rng = np.random.default_rng(4)
    f = np.zeros((q,) + shp)
    orc.equilibrium(lat, orc.F64, 1, 1 + 0.05 * rng.standard_normal(shp), 0.05 * rng.standard_normal((d,) + shp), f)
    f += 1e-3 * rng.standard_normal(f.shape) * ref_lattice(lat)[1].reshape(-1, 1, 1, 1)
    c, _ = ref_lattice(lat)
    m0, p0 = f.sum(), np.tensordot(c, f, axes=(1, 0)).sum(axis=(1, 2, 3))
    out = _run(lat, coll, f, 1.7, 20)
    assert abs(out.sum() - m0) < 1e-11 * m0
    np.testing.assert_allclose(np.tensordot(c, out, axes=(1, 0)).sum(axis=(1, 2, 3)), p0, atol=1e-11)


@pytest.mark.parametrize("lat,coll", [(1, 0), (2, 1), (0, 1)])
def test_uniform_flow_is_steady(lat, coll):
    d, q = LATS[lat]
    shp = grid_shape(lat, 6)
    u0 = np.array([0.03, -0.02, 0.01][:d]).reshape((d, 1, 1, 1)) * np.ones((d,) + shp)
    f = np.zeros((q,) + shp)
    orc.equilibrium(lat, orc.F64, 1, np.ones(shp), u0, f)
    out = _run(lat, coll, f, 1.3, 10)
    np.testing.assert_allclose(out, f, atol=1e-15)


def test_closed_box_halfway_conserves_mass_and_zero_wall_flux():
    lat, (d, q) = 1, LATS[1]
    n = 10
    shp = (n, n, n)
    g = np.indices(shp)
    faces = (g[0] == 0) | (g[0] == n - 1) | (g[1] == 0) | (g[1] == n - 1) | (g[2] == 0) | (g[2] == n - 1)
    idx = np.array(np.where(faces))
    bc_mask = np.zeros((1,) + shp, np.uint8)
    miss = np.zeros((q,) + shp, np.uint8)
    orc.mask_indices(lat, idx, np.zeros(idx.shape[1], np.uint8), 1, bc_mask, miss)
    rng = np.random.default_rng(5)
    f = np.zeros((q,) + shp)
    orc.equilibrium(lat, orc.F64, 1, 1 + 0.02 * rng.standard_normal(shp), 0.03 * rng.standard_normal((d,) + shp), f)
    m0 = f.sum()
    out = _run(lat, 0, f, 1.4, 30, bc_mask, miss, [dict(id=1, kind=orc.HALFWAY_BB)])
    assert abs(out.sum() - m0) < 1e-11 * m0


def test_plane_poiseuille_d2q9():
    """Body-force-driven channel with halfway bounce-back walls at y = -1/2 and ny - 1/2; at
    tau = 1/2 + sqrt(3/16) the BGK halfway wall is exactly mid-link, so the steady profile is
    the parabola u(y) = g/(2 nu) (y + 1/2)(ny - 1/2 - y)."""
    lat, (d, q) = 0, LATS[0]
    nx, ny = 4, 16
    shp = (nx, ny, 1)
    tau = 0.5 + np.sqrt(3 / 16)
    omega, nu = 1 / tau, (tau - 0.5) / 3
    # The masker marks every pull that leaves the box (also across the periodic x ends), so the
    # periodic channel's mask is built directly: missing iff the pull crosses a y wall.
    c, _ = ref_lattice(lat)
    bc_mask = np.zeros((1,) + shp, np.uint8)
    bc_mask[0, :, 0, :] = 1
    bc_mask[0, :, ny - 1, :] = 1
    miss = np.zeros((q,) + shp, np.uint8)
    ys = np.arange(ny).reshape(1, ny, 1)
    for l in range(q):
        src = ys - c[1, l]
        miss[l] = np.broadcast_to((src < 0) | (src >= ny), shp)
    f = np.zeros((q,) + shp)
    orc.equilibrium(lat, orc.F64, 1, np.ones(shp), np.zeros((d,) + shp), f)
    F = 1e-6
    out = _run(lat, 0, f, omega, 6000, bc_mask, miss, [dict(id=1, kind=orc.HALFWAY_BB)], force=(F, 0.0))
    rho = out.sum(0)
    ux = (np.tensordot(c, out, axes=(1, 0))[0] / rho)[0, :, 0]
    y = np.arange(ny)
    expect = F / (2 * nu) * (y + 0.5) * (ny - 0.5 - y)
    # With exact-difference forcing the pre-collision velocity sits F/2 above the parabola.
    np.testing.assert_allclose(ux - F / 2, expect, rtol=1e-6)


@pytest.mark.parametrize("omega,tol", [(1.0, 1e-3), (1.5, 2.5e-2)])
def test_taylor_green_decay_rate(omega, tol):
    """Linear-regime decay of a divergence-free Taylor-Green vortex: KE ~ exp(-2 nu |k|^2 t).
    The BGK discretisation error is O(k^2) and vanishes at omega = 1."""
    lat, (d, q) = 1, LATS[1]
    n = 24
    shp = (n, n, n)
    k = 2 * np.pi / n
    x = np.indices(shp) * k
    u0 = 0.005
    u = np.stack([u0 * np.sin(x[0]) * np.cos(x[1]) * np.cos(x[2]),
                  -u0 * np.cos(x[0]) * np.sin(x[1]) * np.cos(x[2]), np.zeros(shp)])
    f = np.zeros((q,) + shp)
    orc.equilibrium(lat, orc.F64, 1, np.ones(shp), u, f)
    nu = (1 / omega - 0.5) / 3
    c, _ = ref_lattice(lat)

    def ke(ff):
        r = ff.sum(0)
        v = np.tensordot(c, ff, axes=(1, 0)) / r
        return 0.5 * (v**2).sum()

    t1, t2 = 50, 250
    a = _run(lat, 0, f, omega, t1)
    b = _run(lat, 0, a, omega, t2 - t1)
    rate = np.log(ke(a) / ke(b)) / (t2 - t1)
    expect = 2 * nu * 3 * k**2
    assert abs(rate - expect) / expect < tol


@pytest.mark.parametrize("lat", [0, 2])
def test_kbc_relaxes_shear_like_bgk(lat):
    """KBC (kbc.py:252-296) relaxes the shear part at omega whatever gamma is: the off-diagonal
    second moments and the normal-stress differences of f_out - feq are (1 - omega) times those of
    f - feq (at rho = 1); mass and momentum are conserved; at omega = 1 (beta = 1/2, gamma = 2) f_out = feq;
    equilibrium input is a fixed point."""
    d, q = LATS[lat]
    shp = grid_shape(lat, 5)
    rng = np.random.default_rng(6)
    c, w = ref_lattice(lat)
    f = np.zeros((q,) + shp)
    orc.equilibrium(lat, orc.F64, 1, 1 + 0.05 * rng.standard_normal(shp), 0.05 * rng.standard_normal((d,) + shp), f)
    f += 2e-3 * rng.standard_normal(f.shape) * w.reshape(-1, 1, 1, 1)
    # The reference scales the shear part by rho (delta_s = shear * rho), so the identity is
    # exact at rho = 1: normalise every cell.
    f /= f.sum(0)
    rho = f.sum(0)
    u = np.tensordot(c, f, axes=(1, 0)) / rho
    feq = np.zeros_like(f)
    orc.equilibrium(lat, orc.F64, 1, rho, u, feq)

    def pi(g):
        return np.einsum("al,bl,l...->ab...", c, c, g)

    for omega in (0.8, 1.3, 1.9):
        out = orc.collide(lat, 1, orc.F64, 1, f, feq, rho, u, omega)
        np.testing.assert_allclose(out.sum(0), rho, atol=1e-14)
        np.testing.assert_allclose(np.tensordot(c, out, axes=(1, 0)), rho * u, atol=1e-14)
        p_in, p_out = pi(f - feq), pi(out - feq)
        for a in range(d):
            for b in range(d):
                if a != b:
                    np.testing.assert_allclose(p_out[a, b], (1 - omega) * p_in[a, b], atol=1e-14)
        for a in range(d - 1):
            na_in = p_in[a, a] - p_in[d - 1, d - 1]
            na_out = p_out[a, a] - p_out[d - 1, d - 1]
            np.testing.assert_allclose(na_out, (1 - omega) * na_in, atol=1e-14)
    np.testing.assert_allclose(orc.collide(lat, 1, orc.F64, 1, f, feq, rho, u, 1.0), feq, atol=1e-15)
    np.testing.assert_allclose(orc.collide(lat, 1, orc.F64, 1, feq, feq, rho, u, 1.6), feq, atol=1e-15)


def _natural_shear(c, fneq, rho):
    """The shear part of fneq in KBC's natural-moment decomposition, built from its definition
    rather than the reference's index list (kbc.py:192-250): with Pi the second moment of fneq and
    D = Pi - tr(Pi)/d its deviatoric part, a main direction along axis a carries rho D_aa / 2, a
    diagonal in the (a, b) plane carries rho c_a c_b Pi_ab / 4, every other direction nothing.  The
    second moment of the result is rho D (both lattices: the reference's D3Q27 rho s and D2Q9
    rho s / 4 land on this same vector)."""
    d = c.shape[0]
    pi = np.einsum("al,bl,l...->ab...", c, c, fneq)
    tr = np.trace(pi)
    ds = np.zeros_like(fneq)
    for l in range(c.shape[1]):
        axes = np.flatnonzero(c[:, l])
        if len(axes) == 1:
            a = axes[0]
            ds[l] = rho * (pi[a, a] - tr / d) / 2
        elif len(axes) == 2:
            a, b = axes
            ds[l] = rho * c[a, l] * c[b, l] * pi[a, b] / 4
    return ds


def _kbc_residuals(c, omega, f, feq, rho, out):
    """KBC writes f_out = f - beta (2 ds + gamma dh) with dh = fneq - ds, beta = omega / 2, and
    chooses gamma so that the post-collision state is stationary for the entropy along dh:
    <dh, f_out - feq> = 0 in the scalar product weighted by 1/feq (kbc.py:252-296).  Returns
    (distance of f_out - feq - (1 - 2 beta) ds from the line through dh, the stationarity
    residual), both relative.  Together they fix gamma."""
    fneq = f - feq
    ds = _natural_shear(c, fneq, rho)
    dh = fneq - ds
    r = out - feq - (1 - omega) * ds  # = (1 - beta gamma) dh
    k = (r * dh).sum(0) / (dh * dh).sum(0)
    line = np.abs(r - k * dh).max() / np.abs(r).max()
    stationary = np.abs((dh * (out - feq) / feq).sum(0)).max() / np.abs((dh * dh / feq).sum(0)).max()
    return line, stationary


@pytest.mark.parametrize("lat", [0, 2])
def test_kbc_gamma_is_the_entropic_stationary_point(lat):
    """Pins KBC's entropic stabiliser gamma (kbc.py:252-296) for any rho, at omega in {0.8, 1.3,
    1.9}: f_out - feq - (1 - omega) ds lies on the line through dh and is orthogonal to dh in the
    1/feq-weighted product, with ds from the natural-moment definition above.  A flipped sign of the
    (2 - 1/beta) term or two exchanged shear assignments, planted in the oracle, fail it."""
    d, q = LATS[lat]
    shp = grid_shape(lat, 5)
    rng = np.random.default_rng(16)
    c, w = ref_lattice(lat)
    f = np.zeros((q,) + shp)
    orc.equilibrium(lat, orc.F64, 1, 1 + 0.05 * rng.standard_normal(shp), 0.05 * rng.standard_normal((d,) + shp), f)
    f += 2e-3 * rng.standard_normal(f.shape) * w.reshape(-1, 1, 1, 1)
    rho = f.sum(0)
    u = np.tensordot(c, f, axes=(1, 0)) / rho
    feq = np.zeros_like(f)
    orc.equilibrium(lat, orc.F64, 1, rho, u, feq)
    for omega in (0.8, 1.3, 1.9):
        line, stationary = _kbc_residuals(c, omega, f, feq, rho, orc.collide(lat, 1, orc.F64, 1, f, feq, rho, u, omega))
        assert line < 1e-10 and stationary < 1e-12, (omega, line, stationary)
        for slip in (orc.PLANT_KBC_GAMMA_SIGN, orc.PLANT_KBC_SHEAR_INDEX):
            with orc.planted(slip):
                bad = orc.collide(lat, 1, orc.F64, 1, f, feq, rho, u, omega)
            line, stationary = _kbc_residuals(c, omega, f, feq, rho, bad)
            assert max(line, stationary) > 1e-3, (slip, omega, line, stationary)


@pytest.mark.parametrize("lat", [0, 1, 2])
def test_smagorinsky_pinned(lat):
    """SmagorinskyLESBGK (smagorinsky_les_bgk.py:73-89) pinned by what its definition fixes:
    * Cs = 0 is BGK with omega, and fneq = 0 leaves feq unchanged for any Cs;
    * a single-direction perturbation fneq = eps e_l relaxes by 1/tau with
      tau (tau - tau0) = 9 Cs^2 sqrt(S), the quadratic the reference's
      tau = tau0 + (sqrt(tau0^2 + 36 Cs^2 sqrt(S)) - tau0) / 2 solves, where S is eps^2 when the
      signed component sum of c_l is 1, 2 eps^2 when it is >= 2 and 0 otherwise (tau = tau0);
    * the other directions, whose fneq is zero, stay at feq."""
    d, q = LATS[lat]
    L = orc.lattice(lat)
    c = L["c"]
    shp = grid_shape(lat, 4)
    rng = np.random.default_rng(5)
    rho = 1 + 0.05 * rng.standard_normal(shp)
    u = 0.05 * rng.standard_normal((d,) + shp)
    feq = np.zeros((q,) + shp)
    orc.equilibrium(lat, orc.F64, 1, rho, u, feq)
    f = feq + 1e-3 * rng.standard_normal(feq.shape)
    for omega in (0.7, 1.9):
        bgk = orc.collide(lat, 0, orc.F64, 1, f, feq, None, None, omega)
        np.testing.assert_allclose(orc.collide(lat, 2, orc.F64, 1, f, feq, None, None, omega, cparam=0.0), bgk,
                                   rtol=0, atol=1e-15)
        np.testing.assert_allclose(orc.collide(lat, 2, orc.F64, 1, feq, feq, None, None, omega, cparam=0.3), feq,
                                   rtol=0, atol=1e-15)
    cs, omega, eps = 0.17, 1.8, 2e-3
    tau0 = 1.0 / omega
    seen = set()
    for l in range(q):
        fl = feq.copy()
        fl[l] += eps
        out = orc.collide(lat, 2, orc.F64, 1, fl, feq, None, None, omega, cparam=cs)
        tau = eps / (fl[l] - out[l])
        csum = int(c[:, l].sum())
        strain = eps * eps if csum == 1 else (2 * eps * eps if csum >= 2 else 0.0)
        seen.add(min(max(csum, 0), 2))
        np.testing.assert_allclose(tau * (tau - tau0), 9 * cs * cs * np.sqrt(strain), rtol=1e-9, atol=1e-14)
        others = np.delete(np.arange(q), l)
        np.testing.assert_allclose(out[others], feq[others], rtol=0, atol=1e-15)
    assert seen == {0, 1, 2}


def test_momentum_exchange_symmetric_body_has_no_lift():
    lat, (d, q) = 1, LATS[1]
    n = 16
    shp = (n, n, n)
    ind = _sphere(n, 3)
    idx = np.array(ind)
    bc_mask = np.zeros((1,) + shp, np.uint8)
    miss = np.zeros((q,) + shp, np.uint8)
    orc.mask_indices(lat, idx, np.ones(idx.shape[1], np.uint8), 1, bc_mask, miss)
    f = np.zeros((q,) + shp)
    orc.equilibrium(lat, orc.F64, 1, np.ones(shp), np.stack([np.full(shp, 0.02), np.zeros(shp), np.zeros(shp)]), f)
    out = _run(lat, 0, f, 1.2, 20, bc_mask, miss, [dict(id=1, kind=orc.HALFWAY_BB)])
    F = orc.momentum_transfer(lat, orc.F64, 1, dict(id=1, kind=orc.HALFWAY_BB), out, bc_mask, miss)
    assert F[0] > 0  # drag along the flow
    assert abs(F[1]) < 1e-12 * abs(F[0]) + 1e-15 and abs(F[2]) < 1e-12 * abs(F[0]) + 1e-15


def _momentum_balance(lat, coll):
    """A periodic FP64 box with a halfway-bounce-back body (every voxel an interior index, so the
    masker pads it) in a perturbed flow, 3 steps in.  Returns (change of the fluid's momentum over
    one more step, momentum-exchange force of the state before that step)."""
    d, q = LATS[lat]
    shp = grid_shape(lat, 14)
    g = np.indices(shp).astype(float)
    ctr = (6.3, 7.1, 6.6)
    inside = sum((g[a] - ctr[a]) ** 2 for a in range(d)) < 3.2**2
    idx = np.array(np.where(inside), dtype=np.int32)
    bc_mask = np.zeros((1,) + shp, np.uint8)
    miss = np.zeros((q,) + shp, np.uint8)
    orc.mask_indices(lat, idx, np.ones(idx.shape[1], np.uint8), 1, bc_mask, miss)
    rng = np.random.default_rng(3)
    c, w = ref_lattice(lat)
    u = np.array([0.03, -0.01, 0.02][:d]).reshape((d, 1, 1, 1)) + 0.01 * rng.standard_normal((d,) + shp)
    f = np.zeros((q,) + shp)
    orc.equilibrium(lat, orc.F64, 1, 1 + 0.01 * rng.standard_normal(shp), u, f)
    f += 1e-3 * rng.standard_normal(f.shape) * w.reshape(-1, 1, 1, 1)
    bcs = [dict(id=1, kind=orc.HALFWAY_BB)]
    f = _run(lat, coll, f, 1.6, 3, bc_mask, miss, bcs)
    # the fluid: id 0 and the body's neighbours (boundary cells whose rest direction is not missing)
    fluid = (bc_mask[0] == 0) | ((bc_mask[0] == 1) & (miss[0] == 0))
    force = orc.momentum_transfer(lat, orc.F64, 1, bcs[0], f, bc_mask, miss)[:d]
    p0 = np.tensordot(c, f, axes=(1, 0))[:, fluid].sum(1)
    p1 = np.tensordot(c, _run(lat, coll, f, 1.6, 1, bc_mask, miss, bcs), axes=(1, 0))[:, fluid].sum(1)
    return p1 - p0, force


@pytest.mark.parametrize("lat,coll", [(1, 0), (2, 1), (0, 0), (0, 1)])
def test_momentum_exchange_balances_fluid_momentum(lat, coll):
    """Pins the momentum-exchange magnitude (momentum_transfer.py:145-160): the force is the
    momentum the body takes from the fluid, so over one step of a periodic box without body force
    the fluid's momentum changes by exactly -F(state before the step).  Streaming between fluid
    cells and the momentum-conserving collision change nothing; each population f_pc[opp l] that
    would enter the body is bounced back into l, taking 2 c_{opp l} f_pc[opp l], which is what
    f_pc[opp l] + f_ps[l] sums for halfway bounce-back.  Taking f_pc[l] for f_pc[opp l], or dropping
    the f_ps term, planted in the oracle, fails it."""
    dp, force = _momentum_balance(lat, coll)
    assert np.abs(force).max() > 1e-3
    np.testing.assert_allclose(dp, -force, rtol=0, atol=1e-12 * np.abs(force).max())
    for slip in (orc.PLANT_MOMENTUM_OPP_INDEX, orc.PLANT_MOMENTUM_NO_STREAM_TERM):
        with orc.planted(slip):
            dp, force = _momentum_balance(lat, coll)
        assert np.abs(dp + force).max() > 0.1 * np.abs(force).max(), slip


# ------------------------------------------------------------------ boundary-condition closed forms
def _face_case(lat, n, face_x0=True):
    """A face BC on x = 0 (or x = n-1) without edges; returns idx, bc_mask, missing."""
    d, q = LATS[lat]
    shp = grid_shape(lat, n)
    g = np.indices(shp)
    xs = 0 if face_x0 else n - 1
    sel = (g[0] == xs) & (g[1] > 0) & (g[1] < n - 1)
    if d == 3:
        sel &= (g[2] > 0) & (g[2] < n - 1)
    idx = np.array(np.where(sel), dtype=np.int32)
    bc_mask = np.zeros((1,) + shp, np.uint8)
    miss = np.zeros((q,) + shp, np.uint8)
    orc.mask_indices(lat, idx, np.zeros(idx.shape[1], np.uint8), 1, bc_mask, miss)
    return shp, sel, bc_mask, miss


def _random_pops(lat, shp, seed):
    rng = np.random.default_rng(seed)
    _, w = ref_lattice(lat)
    return w.reshape(-1, 1, 1, 1) * (1 + 0.1 * rng.standard_normal((len(w),) + shp))


@pytest.mark.parametrize("lat", [0, 1, 2])
@pytest.mark.parametrize("kind", [orc.ZOUHE_VELOCITY, orc.REGULARIZED_VELOCITY])
def test_velocity_bc_imposes_density_and_normal_momentum(lat, kind):
    """Zou-He velocity (bc_zouhe.py:279-312): on the x = 0 face the normal is n = -e_x, u = v e_x,
    rho = fsum / (1 - v) with fsum = 2 sum(known) + sum(middle).  Non-equilibrium bounce-back then
    gives sum f = rho and sum c_x f = rho v exactly; regularisation keeps both."""
    d, q = LATS[lat]
    c, _ = ref_lattice(lat)
    shp, sel, bc_mask, miss = _face_case(lat, 8)
    fpost = _random_pops(lat, shp, 11)
    v = 0.04
    out = orc.apply_bc(lat, orc.F64, 1, dict(id=1, kind=kind), fpost, fpost, bc_mask, miss, aux=np.full(bc_mask.size, v))
    opp = orc.lattice(lat)["opp"]
    mm = miss.astype(bool)
    known = mm[opp]
    middle = ~(mm | known)
    fsum = 2 * (fpost * known).sum(0) + (fpost * middle).sum(0)
    rho = fsum / (1 - v)
    np.testing.assert_allclose(out.sum(0)[sel], rho[sel], rtol=1e-13)
    np.testing.assert_allclose(np.tensordot(c[0], out, axes=(0, 0))[sel], rho[sel] * v, rtol=1e-12, atol=1e-15)
    # cells off the face are untouched
    np.testing.assert_array_equal(out[:, ~sel], fpost[:, ~sel])


@pytest.mark.parametrize("lat", [0, 1])
@pytest.mark.parametrize("kind", [orc.ZOUHE_PRESSURE, orc.REGULARIZED_PRESSURE])
def test_pressure_bc_imposes_density(lat, kind):
    """Zou-He pressure (bc_zouhe.py:314-338): rho is the prescribed value and u = (fsum/rho - 1) n;
    the result carries exactly that density and normal momentum."""
    d, q = LATS[lat]
    c, _ = ref_lattice(lat)
    shp, sel, bc_mask, miss = _face_case(lat, 8, face_x0=False)
    fpost = _random_pops(lat, shp, 12)
    rho_p = 1.02
    out = orc.apply_bc(lat, orc.F64, 1, dict(id=1, kind=kind), fpost, fpost, bc_mask, miss, aux=np.full(bc_mask.size, rho_p))
    opp = orc.lattice(lat)["opp"]
    mm = miss.astype(bool)
    known = mm[opp]
    middle = ~(mm | known)
    fsum = 2 * (fpost * known).sum(0) + (fpost * middle).sum(0)
    un = -1 + fsum / rho_p  # along n = +e_x on the x = n-1 face
    np.testing.assert_allclose(out.sum(0)[sel], rho_p, rtol=1e-13)
    np.testing.assert_allclose(np.tensordot(c[0], out, axes=(0, 0))[sel], (rho_p * un)[sel], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("lat", [0, 1, 2])
def test_regularized_reproduces_nonequilibrium_stress(lat):
    """Regularisation (helper_functions_bc.py:99-122, Latt et al. 2008 eq. 45) rebuilds
    f = feq + 9/2 w Q:Pi_neq; by fourth-order lattice isotropy the rebuilt populations carry
    exactly the second moment Pi_neq of the Zou-He populations they replace (and rho, u)."""
    d, q = LATS[lat]
    c, w = ref_lattice(lat)
    shp, sel, bc_mask, miss = _face_case(lat, 8)
    fpost = _random_pops(lat, shp, 13)
    v = 0.03
    aux = np.full(bc_mask.size, v)
    zh = orc.apply_bc(lat, orc.F64, 1, dict(id=1, kind=orc.ZOUHE_VELOCITY), fpost, fpost, bc_mask, miss, aux=aux)
    reg = orc.apply_bc(lat, orc.F64, 1, dict(id=1, kind=orc.REGULARIZED_VELOCITY), fpost, fpost, bc_mask, miss, aux=aux)
    rho = zh.sum(0)
    u = np.tensordot(c, zh, axes=(1, 0)) / rho
    feq = np.zeros_like(zh)
    orc.equilibrium(lat, orc.F64, 1, rho, u, feq)
    pi = lambda g: np.einsum("al,bl,l...->ab...", c, c, g)
    np.testing.assert_allclose(pi(reg - feq)[..., sel], pi(zh - feq)[..., sel], atol=1e-14)
    np.testing.assert_allclose(reg.sum(0)[sel], rho[sel], rtol=1e-13)


@pytest.mark.parametrize("lat", [1, 2])
def test_grads_bc_pinned(lat):
    """GradsApproximationBC (bc_grads_approximation.py:181-210, wall weight 1/2) pinned by what
    fixes it:
    * the rest equilibrium w_l is a fixed point for any missing set;
    * with weight 1/2, ((1 - q) a + q (b + c)) / (1 + q) is the plain mean of the three sources, so
      each missing l first becomes mean(f_post[opp l], f_pre[l], f_pre[opp l]), in ascending l and in
      place (a later missing direction sees an earlier one's new value; the random mask has cells
      with both l and opp l missing);
    * the result carries exactly that intermediate's density, momentum and second moment (fourth-
      order isotropy of the regularisation), and lies in the span of the ten Hermite vectors
      w_l {1, c_l, c_l c_l - I/3}, which together fix it;
    * cells with another id are untouched."""
    d, q = LATS[lat]
    c, w = ref_lattice(lat)
    opp = orc.lattice(lat)["opp"]
    shp = (6, 5, 4)
    rng = np.random.default_rng(21)
    bc_mask = (rng.random((1,) + shp) < 0.5).astype(np.uint8)
    miss = (rng.random((q,) + shp) < 0.3).astype(np.uint8)
    miss[0] = 0
    both = miss.astype(bool) & miss[opp].astype(bool)
    assert both[:, bc_mask[0] == 1].any()
    wv = w.reshape(-1, 1, 1, 1)
    rest = np.broadcast_to(wv, (q,) + shp).copy()
    desc = dict(id=1, kind=orc.GRADS_APPROXIMATION)
    np.testing.assert_allclose(orc.apply_bc(lat, orc.F64, 1, desc, rest, rest, bc_mask, miss), rest, rtol=0,
                               atol=1e-15)
    fpre = wv * (1 + 0.1 * rng.standard_normal((q,) + shp))
    fpost = wv * (1 + 0.1 * rng.standard_normal((q,) + shp))
    out = orc.apply_bc(lat, orc.F64, 1, desc, fpre, fpost, bc_mask, miss)
    sel = bc_mask[0] == 1
    np.testing.assert_array_equal(out[:, ~sel], fpost[:, ~sel])
    ft = fpost.copy()
    for l in range(q):
        m = miss[l].astype(bool) & sel
        ft[l][m] = ((ft[opp[l]] + fpre[l] + fpre[opp[l]]) / 3)[m]
    pi = lambda g: np.einsum("al,bl,l...->ab...", c, c, g)
    np.testing.assert_allclose(out.sum(0)[sel], ft.sum(0)[sel], rtol=1e-13)
    np.testing.assert_allclose(np.tensordot(c, out, axes=(1, 0))[:, sel], np.tensordot(c, ft, axes=(1, 0))[:, sel],
                               atol=1e-14)
    np.testing.assert_allclose(pi(out)[..., sel], pi(ft)[..., sel], atol=1e-14)
    basis = [np.ones(q)] + [c[a].astype(float) for a in range(3)]
    basis += [c[a] * c[b] - (a == b) / 3.0 for a in range(3) for b in range(a, 3)]
    B = (w[:, None] * np.array(basis).T)
    cells = out[:, sel]
    coef, *_ = np.linalg.lstsq(B, cells, rcond=None)
    assert np.abs(B @ coef - cells).max() < 1e-14


@pytest.mark.parametrize("lat", [0, 1, 2])
def test_extrapolation_outflow_keeps_uniform_flow(lat):
    """A uniform equilibrium flow through an EquilibriumBC inlet and an ExtrapolationOutflowBC
    outlet (bc_extrapolation_outflow.py:152-195) stays exactly uniform: every extrapolation mixes
    equal values, so an index or sign slip (e.g. f[opp l] for f[l]) would show."""
    d, q = LATS[lat]
    n = 10
    shp = grid_shape(lat, n)
    g = np.indices(shp)
    inlet = g[0] == 0
    outlet = g[0] == n - 1
    bc_mask = np.zeros((1,) + shp, np.uint8)
    miss = np.zeros((q,) + shp, np.uint8)
    # periodic in y (and z): missing only across the x faces
    c, _ = ref_lattice(lat)
    xs = np.arange(n).reshape(n, 1, 1)
    for l in range(q):
        src = xs - c[0, l]
        miss[l] = np.broadcast_to((src < 0) | (src >= n), shp)
    bc_mask[0][inlet] = 1
    bc_mask[0][outlet] = 2
    u0 = np.array([0.03, 0.01, -0.02][:d]).reshape((d,) + (1,) * 3) * np.ones((d,) + shp)
    f = np.zeros((q,) + shp)
    orc.equilibrium(lat, orc.F64, 1, np.ones(shp), u0, f)
    bcs = [dict(id=1, kind=orc.EQUILIBRIUM, rho=1.0, u=tuple(u0[:, 0, 0, 0])), dict(id=2, kind=orc.EXTRAPOLATION_OUTFLOW)]
    # The outlet keeps, in the outgoing slot opp(l) of each missing l, the extrapolated value of
    # direction l that the next step bounces back; for a uniform flow its fixed point is feq_l.
    opp = orc.lattice(lat)["opp"]
    expect = f.copy()
    for l in range(q):
        m = miss[l].astype(bool) & outlet
        expect[opp[l]][m] = f[l][m]
    out = _run(lat, 0, expect, 1.3, 12, bc_mask, miss, bcs)
    np.testing.assert_allclose(out, expect, atol=1e-14)


def _extrapolation_error(lat, axis, side, seed=21):
    """Largest distance, over the outlet's missing directions, between the oracle's stashed value
    and its definition as an upwind interpolation (bc_extrapolation_outflow.py:163-195): the
    post-streaming value of a missing direction l at the outlet (bounce-back: the own-cell f[opp l])
    and at the cell one step upstream along -n (numpy.roll of f[l] by c_l, then by n) are the
    endpoints of a segment of length 1, and the stash in slot opp l is the segment's value at
    distance c_s from the outlet, c_s = sqrt(sum_l w_l c_lx^2) the lattice sound speed."""
    d, q = LATS[lat]
    c, w = ref_lattice(lat)
    opp = orc.lattice(lat)["opp"]
    cs = np.sqrt((w * c[0] ** 2).sum())
    n = 7
    shp = grid_shape(lat, n)
    g = np.indices(shp)
    face = g[axis] == (n - 1 if side > 0 else 0)
    for b in range(d):
        if b != axis:
            face &= (g[b] > 0) & (g[b] < n - 1)
    idx = np.array(np.where(face), dtype=np.int32)
    bc_mask = np.zeros((1,) + shp, np.uint8)
    miss = np.zeros((q,) + shp, np.uint8)
    orc.mask_indices(lat, idx, np.zeros(idx.shape[1], np.uint8), 1, bc_mask, miss)
    f = _random_pops(lat, shp, seed)
    out = _run(lat, 0, f, 1.3, 1, bc_mask, miss, [dict(id=1, kind=orc.EXTRAPOLATION_OUTFLOW)])
    normal = tuple(side if a == axis else 0 for a in range(d))
    err, checked = 0.0, 0
    for l in range(q):
        m = miss[l].astype(bool) & face
        post = np.roll(f[l], shift=tuple(int(v) for v in c[:, l]), axis=tuple(range(d)))
        upstream = np.roll(post, shift=normal, axis=tuple(range(d)))
        for x in zip(*np.where(m)):
            expect = np.interp(-cs, [-1.0, 0.0], [upstream[x], f[opp[l]][x]])
            err = max(err, abs(out[opp[l]][x] - expect))
            checked += 1
    assert checked == face.sum() * (3 if d == 2 else (5 if q == 19 else 9))
    return err


@pytest.mark.parametrize("lat", [0, 1, 2])
@pytest.mark.parametrize("axis,side", [(0, 1), (0, -1), (1, 1)])
def test_extrapolation_outflow_is_upwind_interpolation(lat, axis, side):
    """Pins ExtrapolationOutflowBC's c_s weight and its x - c_l - n source on random populations
    (outlets on x = n-1, x = 0 and y = n-1, without edges, so n is the face normal).  Exchanging
    the weights c_s and 1 - c_s, or dropping n from the source, planted in the oracle, fails it."""
    assert _extrapolation_error(lat, axis, side) < 1e-15
    for slip in (orc.PLANT_EXTRAPOLATION_WEIGHTS, orc.PLANT_EXTRAPOLATION_SOURCE):
        with orc.planted(slip):
            assert _extrapolation_error(lat, axis, side) > 1e-4, slip


def test_do_nothing_skips_streaming():
    """DoNothingBC (bc_do_nothing.py:52-62): the cell collides its own pre-stream populations."""
    lat, (d, q) = 1, LATS[1]
    shp, sel, bc_mask, miss = _face_case(lat, 6)
    f = _random_pops(lat, shp, 14)
    out = _run(lat, 0, f, 1.4, 1, bc_mask, miss, [dict(id=1, kind=orc.DO_NOTHING)])
    c, _ = ref_lattice(lat)
    rho = f.sum(0)
    u = np.tensordot(c, f, axes=(1, 0)) / rho
    feq = np.zeros_like(f)
    orc.equilibrium(lat, orc.F64, 1, rho, u, feq)
    expect = f - 1.4 * (f - feq)
    np.testing.assert_allclose(out[:, sel], expect[:, sel], atol=1e-15)
