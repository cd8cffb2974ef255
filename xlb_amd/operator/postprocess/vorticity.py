"""Vorticity (mirrors xlb/operator/postprocess/vorticity.py:12-90): curl of the velocity by central
differences over the interior cells not adjacent to a boundary cell (xlbh_vorticity)."""

from xlb_amd.compute_backend import ComputeBackend
from xlb_amd.operator.operator import Operator
from xlb_amd.operator.postprocess._stencil import run


class Vorticity(Operator):
    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, u, bc_mask, vorticity, vorticity_magnitude):
        return run("Vorticity", "xlbh_vorticity", u, bc_mask, vorticity, 3, vorticity_magnitude)
