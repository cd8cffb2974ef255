from xlb_amd.utils.utils import save_fields_vtk as save_fields_vtk
