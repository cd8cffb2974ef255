from xlb_amd.operator.force.force import MomentumTransfer as MomentumTransfer
from xlb_amd.operator.force.force import ExactDifference as ExactDifference
