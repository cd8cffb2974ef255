"""Global defaults (mirrors xlb/default_config.py:7-26)."""

from dataclasses import dataclass

from xlb_amd.compute_backend import ComputeBackend


@dataclass
class DefaultConfig:
    default_precision_policy = None
    velocity_set = None
    default_backend = None


def init(velocity_set, default_backend, default_precision_policy):
    """Set the global velocity set, backend and precision policy (default_config.py:14-26).

    Only ``ComputeBackend.MI355X`` is implemented; the HIP library is loaded here so a missing
    build fails at initialisation rather than at the first kernel launch.
    """
    if default_backend != ComputeBackend.MI355X:
        raise ValueError(f"Unsupported compute backend: {default_backend} (this package implements ComputeBackend.MI355X)")
    DefaultConfig.velocity_set = velocity_set
    DefaultConfig.default_backend = default_backend
    DefaultConfig.default_precision_policy = default_precision_policy
    from xlb_amd import _lib

    _lib.lib()


def default_backend() -> ComputeBackend:
    return DefaultConfig.default_backend
