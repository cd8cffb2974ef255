from xlb_amd.operator.equilibrium.equilibrium import Equilibrium as Equilibrium
from xlb_amd.operator.equilibrium.equilibrium import QuadraticEquilibrium as QuadraticEquilibrium
