from xlb_amd.velocity_set.velocity_set import VelocitySet as VelocitySet
from xlb_amd.velocity_set.d2q9 import D2Q9 as D2Q9
from xlb_amd.velocity_set.d3q19 import D3Q19 as D3Q19
from xlb_amd.velocity_set.d3q27 import D3Q27 as D3Q27
