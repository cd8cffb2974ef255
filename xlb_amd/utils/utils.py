"""Field output (mirrors xlb/utils/utils.py:93-145, save_fields_vtk).

The reference hands the fields to pyvista (``ImageData`` with one point more than cells per axis,
cell data flattened in Fortran order, ``save(binary=True)``), which writes a legacy binary VTK
file.  pyvista is not a dependency here: the same file is written directly, in the legacy VTK
4.2 format (``STRUCTURED_POINTS``, ``CELL_DATA``, big-endian binary scalars), which ParaView and
VTK read like the reference's output.
"""

import os
from time import time

import numpy as np


def _to_numpy(value):
    if hasattr(value, "detach"):  # torch tensor, possibly on the GPU
        value = value.detach().cpu().numpy()
    return np.asarray(value)


def save_fields_vtk(fields, timestep, output_dir=".", prefix="fields"):
    """Write ``fields`` (name -> array of shape (nx, ny) or (nx, ny, nz), numpy or torch) as cell
    data of a uniform grid to ``<output_dir>/<prefix>_<timestep:07d>.vtk`` and return the path.
    All fields must have the same shape (utils.py:119-124).  float64 fields are written as
    ``double``, every other dtype as ``float``."""
    arrays = {str(k): _to_numpy(v) for k, v in fields.items()}
    if not arrays:
        raise ValueError("no fields to save")
    dims = None
    for key, value in arrays.items():
        if dims is None:
            dims = value.shape
        else:
            assert value.shape == dims, "All fields must have the same dimensions!"
        if " " in key:
            raise ValueError(f"field name {key!r} contains a space, which the VTK format does not allow")
    if len(dims) not in (2, 3):
        raise ValueError(f"fields must be 2-D or 3-D, got shape {dims}")
    output_filename = os.path.join(output_dir, prefix + "_" + f"{timestep:07d}.vtk")
    points = tuple(d + 1 for d in dims) + ((1,) if len(dims) == 2 else ())
    n_cells = int(np.prod(dims))
    start = time()
    with open(output_filename, "wb") as fh:
        header = (
            "# vtk DataFile Version 4.2\n"
            "vtk output\n"
            "BINARY\n"
            "DATASET STRUCTURED_POINTS\n"
            f"DIMENSIONS {points[0]} {points[1]} {points[2]}\n"
            "SPACING 1 1 1\n"
            "ORIGIN 0 0 0\n"
            f"CELL_DATA {n_cells}\n"
        )
        fh.write(header.encode("ascii"))
        for key, value in arrays.items():
            big = ">f8" if value.dtype == np.float64 else ">f4"
            kind = "double" if big == ">f8" else "float"
            fh.write(f"SCALARS {key} {kind} 1\nLOOKUP_TABLE default\n".encode("ascii"))
            fh.write(np.asarray(value.flatten(order="F"), dtype=big).tobytes())
            fh.write(b"\n")
    print(f"Saved {output_filename} in {time() - start:.6f} seconds.")
    return output_filename
