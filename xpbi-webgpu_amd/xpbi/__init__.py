"""MI355X XPBI substep engine -- Python host surface over libxpbi.so.

Import with the package directory (``xpbi-webgpu_amd``) on ``sys.path``.
"""
from ._abi import FIELD, MAT, STAGE, XpbiError, lib
from .sim import Simulation, pack_materials, params_from_scene
from . import scenes

__all__ = ["FIELD", "MAT", "STAGE", "XpbiError", "lib", "Simulation", "pack_materials", "params_from_scene",
           "scenes"]
