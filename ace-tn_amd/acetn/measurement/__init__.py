from .measure import measure
from .rdm import RDM

__all__ = ["measure", "RDM"]
