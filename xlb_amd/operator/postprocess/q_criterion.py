"""QCriterion (mirrors xlb/operator/postprocess/q_criterion.py:12-142): |curl u| and
Q = (|Omega|^2 - |S|^2) / 2 from central differences over the interior cells not adjacent to a
boundary cell (xlbh_q_criterion)."""

from xlb_amd.compute_backend import ComputeBackend
from xlb_amd.operator.operator import Operator
from xlb_amd.operator.postprocess._stencil import run


class QCriterion(Operator):
    @Operator.register_backend(ComputeBackend.MI355X)
    def mi355x_implementation(self, u, bc_mask, norm_mu, q):
        return run("QCriterion", "xlbh_q_criterion", u, bc_mask, norm_mu, 1, q)
