"""Overlap check of the stepper's boundary-condition list, with the contract of
xlb/helper/check_boundary_overlaps.py:5-24 under its Warp backend: a cell listed twice by one
boundary condition, or by two of them, raises ValueError, because the masks would then depend on
the order in which the masker writes them."""

import numpy as np


def _repeats(columns):
    """True if some column of the (dim, n) integer array occurs more than once."""
    if columns.shape[1] < 2:
        return False
    ordered = columns[:, np.lexsort(columns[::-1])]
    return bool((ordered[:, 1:] == ordered[:, :-1]).all(axis=0).any())


def check_bc_overlaps(bclist, dim, compute_backend):
    """Raise ValueError if a boundary condition's `indices` name a cell twice, or two boundary
    conditions share a cell.  Boundary conditions without indices (mesh-defined) are skipped;
    `compute_backend` is accepted for interface parity."""
    listed = []
    for bc in bclist:
        if bc.indices is None:
            continue
        cells = np.asarray(bc.indices, dtype=np.int64).reshape(len(bc.indices), -1)[:dim]
        if _repeats(cells):
            raise ValueError(f"Boundary condition {type(bc).__name__} has duplicate indices!")
        listed.append(cells)
    if listed and _repeats(np.concatenate(listed, axis=1)):
        raise ValueError("Boundary condition list contains duplicate indices!")
