from xlb_amd.grid.grid import grid_factory as grid_factory, Grid as Grid, MI355XGrid as MI355XGrid
from xlb_amd.grid.grid import block_exponents as block_exponents, copy_population as copy_population
from xlb_amd.grid.grid import decode_populations as decode_populations
