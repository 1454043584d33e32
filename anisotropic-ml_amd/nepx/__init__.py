"""nepx: MI355X-native rigid-body NEP molecular-dynamics step.

Host side of libnepx.so.  The reference call surface lives in the modules
named after the reference classes (DescriptorGenerator, ForceEvaluator, Sim,
NVT); `system.Arena` is the device-resident engine underneath.
"""
from . import _lib  # noqa: F401  (raises if the HIP library is missing)
from .model import EnergyNet, make_pair_model, n_descriptor  # noqa: F401
from .shapes import SHAPES, Shape, get_shape  # noqa: F401
from .system import Arena, SystemSpec  # noqa: F401

__all__ = ["Arena", "SystemSpec", "EnergyNet", "make_pair_model", "n_descriptor", "SHAPES", "Shape",
           "get_shape"]
