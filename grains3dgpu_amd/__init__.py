"""grains3dgpu_amd -- MI355X-native DEM time step (broad phase, GJK narrow
phase, Hooke contact force, explicit integration) behind a C-ABI.

`Engine` (engine.py) drives libgrains_hip.so; `scene` builds workloads.
"""
from . import _abi as abi  # noqa: F401
from .scene import Scene, load  # noqa: F401
