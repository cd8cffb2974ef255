"""Pull streaming (mirrors xlb/operator/stream/stream.py:13-115)."""

from xlb_amd import _lib
from xlb_amd.compute_backend import ComputeBackend
from xlb_amd.operator.operator import Operator, common_plane, expect_population, field_extents


class Stream(Operator):
    """f_1[l](x) = f_0[l](x - c_l) with periodic wrap (stream.py:59-99)."""

    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, f_0, f_1):
        _lib.require_cuda(f_0, f_1)
        nx, ny, nz = field_extents(f_0)
        fmt, q = self.format_of(f_0), self.velocity_set.q
        plane = common_plane([expect_population(f_0, q, (nx, ny, nz), "f_0", f_0.dtype, self.lattice_id, fmt),
                              expect_population(f_1, q, (nx, ny, nz), "f_1", f_0.dtype, self.lattice_id, fmt)], "Stream")
        if f_0.data_ptr() == f_1.data_ptr():
            raise ValueError("Stream cannot run in place")
        _lib.check(
            _lib.lib().xlbh_stream(self.lattice_id, fmt, _lib.ptr(f_0), _lib.ptr(f_1), nx, ny, nz, plane, _lib.stream_of(f_0)),
            "Stream",
        )
        _lib.mark_written(f_1)
        return f_1
