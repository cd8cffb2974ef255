from xlb_amd.grid.grid import grid_factory as grid_factory, Grid as Grid, MI355XGrid as MI355XGrid
