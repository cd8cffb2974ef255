"""The oracle's mesh masker (restatement of mesh_boundary_masker.py:49-236), pinned on the CPU:
its Schwarz-Seidel triangle / box test against an independent float64 separating-axis test, and
the voxelisation of an axis-aligned box surface against its closed form."""

import numpy as np
import pytest

from mesh_cases import box_mesh, sat_overlap_margin
from oracle import oracle as orc


def test_tri_box_matches_separating_axis_test():
    rng = np.random.default_rng(3)
    checked = hits = 0
    for _ in range(20000):
        tri = rng.uniform(-1.0, 3.0, size=(3, 3))
        p = rng.integers(-1, 3, size=3)
        m = sat_overlap_margin(tri, p)
        if abs(m) < 1e-4:  # touching within rounding: either answer is defensible
            continue
        got = orc.tri_box(tri.astype(np.float32), p)
        assert got == (m < 0), f"tri {tri.tolist()} box {p.tolist()}: oracle {got}, SAT margin {m}"
        checked += 1
        hits += got
    assert checked > 19000 and 2000 < hits < checked - 2000


@pytest.mark.parametrize("lo,hi", [((2.3, 3.6, 1.2), (9.4, 7.7, 8.1)), ((3.5, 3.5, 3.5), (8.5, 10.5, 6.5))])
def test_box_surface_voxelisation_closed_form(lo, hi):
    shape = (12, 14, 11)
    lat, q = 1, 19
    bc_mask = np.zeros((1,) + shape, np.uint8)
    missing = np.zeros((q,) + shape, np.uint8)
    orc.mesh_mask(lat, box_mesh(lo, hi), 7, bc_mask, missing)
    g = np.indices(shape).astype(float)
    lo, hi = np.asarray(lo).reshape(3, 1, 1, 1), np.asarray(hi).reshape(3, 1, 1, 1)
    overlap = np.all((g <= hi) & (g + 1 >= lo), axis=0)
    strictly_inside = np.all((g > lo) & (g + 1 < hi), axis=0)
    solid = overlap & ~strictly_inside
    np.testing.assert_array_equal(bc_mask[0] == 255, solid)
    # non-solid voxels with a solid neighbour c_l carry the id and missing[opp l]
    lt = orc.lattice(lat)
    c, opp = lt["c"], lt["opp"]
    pad = np.pad(solid, 1)
    expect_missing = np.zeros_like(missing, dtype=bool)
    for l in range(1, q):
        sx, sy, sz = (int(c[0][l]), int(c[1][l]), int(c[2][l]))
        nb = pad[1 + sx:1 + sx + shape[0], 1 + sy:1 + sy + shape[1], 1 + sz:1 + sz + shape[2]]
        expect_missing[opp[l]] |= nb & ~solid
    np.testing.assert_array_equal(missing.astype(bool), expect_missing)
    np.testing.assert_array_equal(bc_mask[0] == 7, expect_missing.any(axis=0))
