from xlb_amd.operator.postprocess.vorticity import Vorticity as Vorticity
from xlb_amd.operator.postprocess.q_criterion import QCriterion as QCriterion
