from .model import Model
from .model_factory import model_factory
from .pauli_matrix import PauliMatrix, pauli_matrices

__all__ = ["Model", "model_factory", "PauliMatrix", "pauli_matrices"]
