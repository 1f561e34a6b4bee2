from ..model import Model
from ..pauli_matrix import pauli_matrices


class HeisenbergModel(Model):
    """acetn/model/models/heisenberg.py:4-26."""

    def initial_site_state(self, site):
        xi, yi = site
        return [1., 0.] if (xi + yi) % 2 == 0 else [0., 1.]

    def one_site_observables(self, site):
        xi, yi = site
        X, Y, Z, I = pauli_matrices(self.dtype)
        return {"sx": X, "sz": Z} if (xi + yi) % 2 == 0 else {"sx": -X, "sz": -Z}

    def two_site_hamiltonian(self, bond):
        J = self.params.get('J')
        X, Y, Z, I = pauli_matrices(self.dtype)
        return 0.25 * J * (X * X + Y * Y + Z * Z)
