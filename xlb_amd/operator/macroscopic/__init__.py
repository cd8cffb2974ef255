from xlb_amd.operator.macroscopic.macroscopic import Macroscopic as Macroscopic
from xlb_amd.operator.macroscopic.macroscopic import ZeroMoment as ZeroMoment
from xlb_amd.operator.macroscopic.macroscopic import FirstMoment as FirstMoment
from xlb_amd.operator.macroscopic.macroscopic import SecondMoment as SecondMoment
