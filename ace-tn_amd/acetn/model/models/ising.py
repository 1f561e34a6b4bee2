from ..model import Model
from ..pauli_matrix import pauli_matrices


class IsingModel(Model):
    """acetn/model/models/ising.py:4-21."""

    def one_site_observables(self, site):
        X, Y, Z, I = pauli_matrices(self.dtype)
        return {"sx": X, "sz": Z}

    def one_site_hamiltonian(self, site):
        X, Y, Z, I = pauli_matrices(self.dtype)
        return -self.params.get('hx') * X

    def two_site_hamiltonian(self, bond):
        X, Y, Z, I = pauli_matrices(self.dtype)
        return -self.params.get('jz') * Z * Z
