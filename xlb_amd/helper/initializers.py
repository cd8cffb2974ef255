"""Equilibrium initialisation (mirrors xlb/helper/initializers.py:5-20)."""

from xlb_amd.operator.equilibrium import QuadraticEquilibrium


def initialize_eq(f, grid, velocity_set, precision_policy, compute_backend, rho=None, u=None):
    if rho is None:
        rho = grid.create_field(cardinality=1, fill_value=1.0, dtype=precision_policy.compute_precision)
    if u is None:
        u = grid.create_field(cardinality=velocity_set.d, fill_value=0.0, dtype=precision_policy.compute_precision)
    equilibrium = QuadraticEquilibrium(velocity_set, precision_policy, compute_backend)
    f = equilibrium(rho, u, f)
    del rho, u
    return f
