from xlb_amd.helper.nse_solver import create_nse_fields as create_nse_fields
from xlb_amd.helper.initializers import initialize_eq as initialize_eq
from xlb_amd.helper.check_boundary_overlaps import check_bc_overlaps as check_bc_overlaps
