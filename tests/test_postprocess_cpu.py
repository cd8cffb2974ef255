"""The post-processing oracle (oracle/postprocess.py) pinned by closed forms: central differences
are exact on a linear velocity field u = A x, so the curl, |curl u|, |S|^2 and |Omega|^2 follow from
A; edge cells and cells next to a boundary cell keep their previous values."""

import numpy as np

from oracle import postprocess as pp


def linear_field(shape, A, b):
    g = np.indices(shape).astype(np.float64)
    return np.einsum("ab,b...->a...", A, g) + np.asarray(b).reshape(3, 1, 1, 1)


def test_linear_field_closed_forms():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((3, 3))
    shape = (7, 6, 5)
    u = linear_field(shape, A, rng.standard_normal(3))
    vort = np.full((3,) + shape, -7.0)
    mag = np.full((1,) + shape, -7.0)
    mu = np.full((1,) + shape, -7.0)
    q = np.full((1,) + shape, -7.0)
    pp.vorticity(u, None, vort, mag)
    pp.q_criterion(u, None, mu, q)
    curl = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    S, W = 0.5 * (A + A.T), 0.5 * (A - A.T)
    inner = (slice(1, -1),) * 3
    for k in range(3):
        np.testing.assert_allclose(vort[k][inner], curl[k], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(mag[0][inner], np.linalg.norm(curl), rtol=1e-12)
    np.testing.assert_allclose(mu[0][inner], np.linalg.norm(curl), rtol=1e-12)
    np.testing.assert_allclose(q[0][inner], 0.5 * ((W * W).sum() - (S * S).sum()), rtol=1e-12, atol=1e-12)
    # rigid rotation: Q = |omega|^2 / 2 with curl = 2 omega
    w = np.array([0.3, -0.2, 0.5])
    R = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    u = linear_field(shape, R, np.zeros(3))
    pp.q_criterion(u, None, mu, q)
    np.testing.assert_allclose(mu[0][inner], 2 * np.linalg.norm(w), rtol=1e-12)
    np.testing.assert_allclose(q[0][inner], (w * w).sum(), rtol=1e-12)
    # edges untouched
    assert np.all(vort[:, 0] == -7.0) and np.all(q[0, :, -1] == -7.0)


def test_cells_next_to_boundary_are_skipped():
    shape = (8, 8, 8)
    u = linear_field(shape, np.eye(3), np.zeros(3))
    bc = np.zeros((1,) + shape, np.uint8)
    bc[0, 4, 4, 4] = 3
    q = np.full((1,) + shape, -7.0)
    mu = np.zeros((1,) + shape)
    pp.q_criterion(u, bc, mu, q)
    for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        assert q[0, 4 + d[0], 4 + d[1], 4 + d[2]] == -7.0
    assert q[0, 4, 4, 4] != -7.0  # the boundary cell itself is computed, as in the reference
    assert q[0, 2, 2, 2] != -7.0
