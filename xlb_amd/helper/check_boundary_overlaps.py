"""Duplicate-index check (mirrors xlb/helper/check_boundary_overlaps.py:5-24, Warp behaviour:
duplicates raise)."""

import numpy as np


def check_bc_overlaps(bclist, dim, compute_backend):
    index_list = [[] for _ in range(dim)]
    for bc in bclist:
        if bc.indices is None:
            continue
        index_arr = np.unique(np.asarray(bc.indices), axis=-1)
        if index_arr.shape[-1] != len(bc.indices[0]):
            raise ValueError(f"Boundary condition {bc.__class__.__name__} has duplicate indices!")
        for d in range(dim):
            index_list[d] += list(bc.indices[d])
    if len(index_list[0]) == 0:
        return
    index_arr = np.unique(np.asarray(index_list), axis=-1)
    if index_arr.shape[-1] != len(index_list[0]):
        raise ValueError("Boundary condition list containes duplicate indices!")
