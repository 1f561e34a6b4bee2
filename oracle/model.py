"""Hamiltonians, observables and imaginary-time gates (oracle side).

Restates acetn/model/pauli_matrix.py, acetn/model/models/{heisenberg,ising}.py
and acetn/evolution/gate.py for real FP64 arithmetic.
"""
import numpy as np
import scipy.linalg

# acetn/model/pauli_matrix.py:29-38 -- the real-valued Y is [[0,-1],[1,0]].
X = np.array([[0., 1.], [1., 0.]])
Y = np.array([[0., -1.], [1., 0.]])
Z = np.array([[1., 0.], [0., -1.]])
I2 = np.eye(2)


def kron_real_pauli(a_key, a, b_key, b):
    """Pauli product with the real-Y phase rule of pauli_matrix.py:64-72:
    flip the sign when the combined key holds an even number of Y and the
    left factor an odd number."""
    key = a_key + b_key
    phase = -1.0 if (key.count('Y') % 2 == 0 and a_key.count('Y') % 2 != 0) else 1.0
    return phase * np.kron(a, b)


class Heisenberg:
    """acetn/model/models/heisenberg.py:4-26."""
    name = "heisenberg"

    def __init__(self, J=1.0):
        self.J = float(J)
        self.dim = 2

    def initial_site_state(self, site):
        x, y = site
        return [1., 0.] if (x + y) % 2 == 0 else [0., 1.]

    def one_site_hamiltonian(self, site):
        return None

    def two_site_hamiltonian(self, bond):
        XX = kron_real_pauli('X', X, 'X', X)
        YY = kron_real_pauli('Y', Y, 'Y', Y)
        ZZ = kron_real_pauli('Z', Z, 'Z', Z)
        return 0.25 * self.J * (XX + YY + ZZ)

    def one_site_observables(self, site):
        x, y = site
        s = 1.0 if (x + y) % 2 == 0 else -1.0
        return {"sx": s * X, "sz": s * Z}

    def two_site_observables(self, bond):
        return {}


class Ising:
    """acetn/model/models/ising.py:4-21 (transverse-field Ising)."""
    name = "ising"

    def __init__(self, jz=1.0, hx=0.0):
        self.jz = float(jz)
        self.hx = float(hx)
        self.dim = 2

    def initial_site_state(self, site):
        return [1., 0.]

    def one_site_hamiltonian(self, site):
        return -self.hx * X

    def two_site_hamiltonian(self, bond):
        return -self.jz * kron_real_pauli('Z', Z, 'Z', Z)

    def one_site_observables(self, site):
        return {"sx": X, "sz": Z}

    def two_site_observables(self, bond):
        return {}


def bond_gate(model, dtau, bond):
    """Two-site gate with one-site terms folded in at weight 1/4
    (acetn/evolution/gate.py:63-90, wrap_one_site=True)."""
    pD = model.dim
    h = np.array(model.two_site_hamiltonian(bond), dtype=np.float64)
    h1 = model.one_site_hamiltonian(bond[0])
    h2 = model.one_site_hamiltonian(bond[1])
    if h1 is not None and h2 is not None:
        h = h + 0.25 * (np.kron(h1, np.eye(pD)) + np.kron(np.eye(pD), h2))
    return scipy.linalg.expm(-dtau * h).reshape(pD, pD, pD, pD)
