"""Pauli algebra with the real-valued Y convention
(acetn/model/pauli_matrix.py:4-182).  Host-side d x d matrices."""
import torch


class PauliMatrix:
    def __init__(self, key, dtype=torch.float64, device="cpu", mat=None):
        self.key = key
        self.dtype = dtype
        self.device = device
        if mat is not None:
            self.mat = mat
        else:
            base = {'X': [[0., 1.], [1., 0.]], 'Y': [[0., -1.], [1., 0.]],
                    'Z': [[1., 0.], [0., -1.]], 'I': [[1., 0.], [0., 1.]]}[key]
            self.mat = torch.tensor(base, dtype=dtype, device=device)
            if key == 'Y' and torch.is_complex(self.mat):
                self.mat = self.mat * 1j

    def __mul__(self, other):
        """Kronecker product (sign flip for an even product of real Y,
        pauli_matrix.py:64-72) or scalar scaling."""
        if isinstance(other, PauliMatrix):
            new_key = self.key + other.key
            phase = 1.
            if not torch.is_complex(self.mat):
                phase = -1. if (new_key.count('Y') % 2 == 0 and self.key.count('Y') % 2 != 0) else 1.
            return PauliMatrix(new_key, self.dtype, self.device, mat=phase * torch.kron(self.mat, other.mat))
        return PauliMatrix(self.key, self.dtype, self.device, mat=other * self.mat)

    def __imul__(self, other):
        """``m *= other`` (pauli_matrix.py:74-90): with a PauliMatrix the
        plain Kronecker product of the two matrices as a bare tensor (no
        real-Y sign, unlike ``*``); with a scalar, the scaled PauliMatrix."""
        if isinstance(other, PauliMatrix):
            return torch.kron(self.mat, other.mat)
        return PauliMatrix(self.key, self.dtype, self.device, mat=other * self.mat)

    def __rmul__(self, other):
        return PauliMatrix(self.key, self.dtype, self.device, mat=other * self.mat)

    def __add__(self, other):
        return PauliMatrix(self.key + other.key, self.dtype, self.device, mat=self.mat + other.mat)

    def __sub__(self, other):
        return PauliMatrix(self.key + other.key, self.dtype, self.device, mat=self.mat - other.mat)

    def __neg__(self):
        return PauliMatrix(self.key, self.dtype, self.device, mat=-self.mat)

    def __truediv__(self, scalar):
        return PauliMatrix(self.key, self.dtype, self.device, mat=self.mat / scalar)


def pauli_matrices(dtype=torch.float64, device="cpu"):
    return tuple(PauliMatrix(k, dtype, device) for k in "XYZI")
