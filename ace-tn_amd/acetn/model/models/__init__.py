from .heisenberg import HeisenbergModel
from .ising import IsingModel

__all__ = ['IsingModel', 'HeisenbergModel']
models = {'ising': IsingModel, 'heisenberg': HeisenbergModel}
