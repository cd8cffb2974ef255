"""xlb_amd: an MI355X (gfx950) compute backend with the operator API of XLB.

Scripts written for the reference switch by importing ``xlb_amd`` in place of ``xlb`` and
selecting ``ComputeBackend.MI355X``.  All arithmetic runs in the hand-written HIP kernels of
``libxlbhip.so`` (C ABI in include/xlb_hip.h); fields are torch tensors in HBM.
"""

__version__ = "0.1.0"

from xlb_amd.compute_backend import ComputeBackend as ComputeBackend
from xlb_amd.precision_policy import PrecisionPolicy as PrecisionPolicy, Precision as Precision
from xlb_amd.physics_type import PhysicsType as PhysicsType
from xlb_amd.default_config import init as init, DefaultConfig as DefaultConfig

import xlb_amd.velocity_set
import xlb_amd.grid
import xlb_amd.operator
import xlb_amd.operator.stream
import xlb_amd.operator.equilibrium
import xlb_amd.operator.macroscopic
import xlb_amd.operator.collision
import xlb_amd.operator.boundary_condition
import xlb_amd.operator.boundary_masker
import xlb_amd.operator.force
import xlb_amd.operator.postprocess
import xlb_amd.operator.precision_caster
import xlb_amd.operator.stepper
import xlb_amd.helper
import xlb_amd.autodiff
import xlb_amd.utils
