import torch

from .bond import Bond
from .ipeps import Ipeps

# acetn/ipeps/__init__.py:7 -- the reference disables autograd globally on
# import; kept for drop-in behaviour.
torch.set_grad_enabled(False)

__all__ = ["Ipeps", "Bond"]
