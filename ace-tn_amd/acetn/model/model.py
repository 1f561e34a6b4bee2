"""Model base class (acetn/model/model.py:3-171)."""
from abc import ABC, abstractmethod


class Model(ABC):
    def __init__(self, config):
        self.config = config
        self.name = config.name
        self.params = config.params
        self.dtype = config.dtype
        self.device = "cpu"  # operators are d x d host matrices
        self.dim = config.dim

    @staticmethod
    def bond_direction(bond):
        _, _, k = bond
        return {0: '-x', 1: '+y', 2: '+x', 3: '-y'}.get(k, None)

    def initial_site_state(self, site):
        return [1, ] + [0, ] * (self.dim - 1)

    def one_site_observables(self, site=None):
        return {}

    def two_site_observables(self, bond=None):
        return {}

    def one_site_hamiltonian(self, site=None):
        return None

    @abstractmethod
    def two_site_hamiltonian(self, bond=None):
        pass
