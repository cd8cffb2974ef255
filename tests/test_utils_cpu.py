"""save_fields_vtk (utils.py:93-145): the legacy binary VTK file it writes is read back by an
independent parser of the format (header grammar, big-endian payload); grid dimensions are the
cell counts + 1 and values are in Fortran (x-fastest) order, as the reference's pyvista call
stores them.  pyvista itself is not installed, so byte-identity with its output is parity
unpinned."""

import numpy as np
import pytest

from xlb_amd.utils import save_fields_vtk


def read_legacy_vtk(path):
    raw = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        end = raw.index(b"\n", pos)
        text = raw[pos:end].decode("ascii")
        pos = end + 1
        return text

    assert line().startswith("# vtk DataFile Version")
    line()
    assert line() == "BINARY"
    assert line() == "DATASET STRUCTURED_POINTS"
    header = {}
    while True:
        t = line().split()
        header[t[0]] = t[1:]
        if t[0] == "CELL_DATA":
            break
    n = int(header["CELL_DATA"][0])
    out = {}
    while pos < len(raw):
        t = line().split()
        assert t[0] == "SCALARS" and t[3] == "1"
        assert line() == "LOOKUP_TABLE default"
        dt = {"float": ">f4", "double": ">f8"}[t[2]]
        size = n * np.dtype(dt).itemsize
        out[t[1]] = np.frombuffer(raw[pos:pos + size], dtype=dt)
        pos += size
        assert raw[pos:pos + 1] == b"\n"
        pos += 1
    return header, out


@pytest.mark.parametrize("shape", [(5, 4, 3), (6, 7)])
def test_save_fields_vtk_round_trip(tmp_path, shape):
    rng = np.random.default_rng(0)
    rho = rng.random(shape).astype(np.float32)
    ux = rng.standard_normal(shape)
    path = save_fields_vtk({"rho": rho, "u_x": ux}, 42, output_dir=str(tmp_path), prefix="fields")
    assert path.endswith("fields_0000042.vtk")
    header, data = read_legacy_vtk(path)
    dims = [int(v) for v in header["DIMENSIONS"]]
    assert dims == [s + 1 for s in shape] + ([1] if len(shape) == 2 else [])
    assert int(header["CELL_DATA"][0]) == rho.size
    np.testing.assert_array_equal(data["rho"].reshape(shape, order="F"), rho)
    np.testing.assert_array_equal(data["u_x"].reshape(shape, order="F"), ux)
    # x fastest: the second value is cell (1, 0, ...)
    assert data["rho"][1] == rho[(1,) + (0,) * (len(shape) - 1)]


def test_save_fields_vtk_rejects_mismatched_shapes(tmp_path):
    with pytest.raises(AssertionError):
        save_fields_vtk({"a": np.zeros((3, 3)), "b": np.zeros((3, 4))}, 0, output_dir=str(tmp_path))
