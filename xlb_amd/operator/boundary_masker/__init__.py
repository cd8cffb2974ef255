from xlb_amd.operator.boundary_masker.indices_boundary_masker import IndicesBoundaryMasker as IndicesBoundaryMasker


class MeshBoundaryMasker:
    """STL voxelisation masker (mesh_boundary_masker.py) -- not provided by this backend yet."""

    def __init__(self, *args, **kwargs):
        raise NotImplementedError("MeshBoundaryMasker is not implemented by the MI355X backend")
