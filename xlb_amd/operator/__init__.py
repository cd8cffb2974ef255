from xlb_amd.operator.operator import Operator as Operator
