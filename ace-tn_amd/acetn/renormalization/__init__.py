from .ctmrg import ctmrg
from .directional_mover import DirectionalMover
from .projectors import ProjectorCalculator

__all__ = ["ctmrg", "DirectionalMover", "ProjectorCalculator"]
