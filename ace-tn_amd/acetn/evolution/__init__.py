"""Imaginary-time evolution (acetn/evolution/__init__.py:1-28) on the MI355X
engine: the full-update bond step (QR split, norm tensor, positive
approximation, gauge fix, ALS, finalisation) and the simple update run as
HIP kernels behind the C-ABI of include/acetn_hip.h."""
from .als_solver import ALSSolver
from .evolve import evolve
from .gate import Gate
from .tensor_update import TensorUpdater
from .full_update import FullUpdater
from .fast_full_update import FastFullUpdater
from .simple_update import SimpleUpdater

__all__ = [
    'ALSSolver',
    'evolve',
    'Gate',
    'TensorUpdater',
    'FullUpdater',
    'FastFullUpdater',
    'SimpleUpdater',
]
