"""Oracle (test infrastructure only): NumPy restatement of the reference's post-processing kernels,
Vorticity (xlb/operator/postprocess/vorticity.py:30-80) and QCriterion (q_criterion.py:35-132).

Interior cells 1..n-2 whose six face neighbours all have bc_mask == 0 get central-difference
derivatives (a - b) / 2; every other cell keeps its previous output value.  Computed in the dtype
of u with the reference's term order.
"""

import numpy as np


def _gradient(u, bc_mask):
    """(sel, g) with sel the (nx, ny, nz) computed-cell mask and g[a][b] = d u_a / d x_b on it."""
    _, nx, ny, nz = u.shape
    sel = np.zeros((nx, ny, nz), dtype=bool)
    sel[1:-1, 1:-1, 1:-1] = True
    if bc_mask is not None:
        m = bc_mask[0] != 0
        near = np.zeros_like(sel)
        near[1:-1, 1:-1, 1:-1] = (m[2:, 1:-1, 1:-1] | m[:-2, 1:-1, 1:-1] | m[1:-1, 2:, 1:-1] | m[1:-1, :-2, 1:-1]
                                  | m[1:-1, 1:-1, 2:] | m[1:-1, 1:-1, :-2])
        sel &= ~near
    two = u.dtype.type(2)
    g = [[None] * 3 for _ in range(3)]
    for a in range(3):
        ua = u[a]
        g[a][0] = (ua[2:, 1:-1, 1:-1] - ua[:-2, 1:-1, 1:-1]) / two
        g[a][1] = (ua[1:-1, 2:, 1:-1] - ua[1:-1, :-2, 1:-1]) / two
        g[a][2] = (ua[1:-1, 1:-1, 2:] - ua[1:-1, 1:-1, :-2]) / two
    return sel, g


def vorticity(u, bc_mask, vort, mag):
    """In place on vort (3, ...) and mag (1, ...)."""
    sel, g = _gradient(u, bc_mask)
    s = sel[1:-1, 1:-1, 1:-1]
    vx = g[2][1] - g[1][2]
    vy = g[0][2] - g[2][0]
    vz = g[1][0] - g[0][1]
    m = np.sqrt(vx * vx + vy * vy + vz * vz)
    for k, v in enumerate((vx, vy, vz)):
        vort[k, 1:-1, 1:-1, 1:-1][s] = v[s]
    mag[0, 1:-1, 1:-1, 1:-1][s] = m[s]


def q_criterion(u, bc_mask, norm_mu, q):
    """In place on norm_mu (1, ...) and q (1, ...)."""
    sel, g = _gradient(u, bc_mask)
    s = sel[1:-1, 1:-1, 1:-1]
    half = u.dtype.type(0.5)
    mx, my, mz = g[2][1] - g[1][2], g[0][2] - g[2][0], g[1][0] - g[0][1]
    mu = np.sqrt(mx * mx + my * my + mz * mz)
    s00, s11, s22 = g[0][0], g[1][1], g[2][2]
    s01, s02, s12 = half * (g[0][1] + g[1][0]), half * (g[0][2] + g[2][0]), half * (g[1][2] + g[2][1])
    ss = s00 * s00 + s01 * s01 + s02 * s02 + s01 * s01 + s11 * s11 + s12 * s12 + s02 * s02 + s12 * s12 + s22 * s22
    o01, o02, o12 = half * (g[0][1] - g[1][0]), half * (g[0][2] - g[2][0]), half * (g[1][2] - g[2][1])
    oo = 2 * (o01 * o01 + o02 * o02 + o12 * o12)
    norm_mu[0, 1:-1, 1:-1, 1:-1][s] = mu[s]
    q[0, 1:-1, 1:-1, 1:-1][s] = (half * (oo - ss))[s]
