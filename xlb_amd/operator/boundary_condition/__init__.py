from xlb_amd.operator.boundary_condition.boundary_condition import (
    ImplementationStep as ImplementationStep,
    BoundaryCondition as BoundaryCondition,
    BoundaryConditionRegistry as BoundaryConditionRegistry,
    boundary_condition_registry as boundary_condition_registry,
    EquilibriumBC as EquilibriumBC,
    HalfwayBounceBackBC as HalfwayBounceBackBC,
    FullwayBounceBackBC as FullwayBounceBackBC,
    DoNothingBC as DoNothingBC,
    ZouHeBC as ZouHeBC,
    RegularizedBC as RegularizedBC,
    ExtrapolationOutflowBC as ExtrapolationOutflowBC,
    GradsApproximationBC as GradsApproximationBC,
)
