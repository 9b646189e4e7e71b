"""grains3dgpu_amd -- MI355X-native DEM time step (linked-cell broad phase,
GJK narrow phase, Hooke contact force, explicit integration) behind a C-ABI
(include/grains_hip.h, libgrains_hip.so).

`engine.Engine` drives the library; `scene` builds workloads.
"""
from . import _abi as abi  # noqa: F401
from .scene import Scene, load  # noqa: F401
