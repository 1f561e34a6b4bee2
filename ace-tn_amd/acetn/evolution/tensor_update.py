"""Bond-frame handling of a tensor update (acetn/evolution/tensor_update.py:7-79).

The rotation into the bond frame ("lurdp" legs cyclically shifted by k) and
back is folded into the HIP decomposition / finalisation kernels
(acetn_hip_bond_decompose, acetn_hip_finalize_bond), so update() hands the
site-frame tensors straight to tensor_update()."""
from abc import ABC, abstractmethod

from .. import _hip


class TensorUpdater(ABC):
    def __init__(self, ipeps, gate):
        self.ipeps = ipeps
        self.dims = ipeps.dims
        self.gate = gate

    @abstractmethod
    def tensor_update(self, a1, a2, bond):
        """a1, a2 in the site frame -> updated a1, a2 in the site frame."""

    def update(self, bond):
        """tensor_update.py:18-38."""
        s1, s2, k = bond
        a1, a2 = self.tensor_update(self.ipeps[s1]['A'], self.ipeps[s2]['A'], bond)
        return a1, a2

    @staticmethod
    def bond_permutation(k):
        """tensor_update.py:46-50."""
        return [(i + k) % 4 for i in range(4)] + [4, ]

    @staticmethod
    def decompose_site_tensors(a1, a2, k=0):
        """tensor_update.py:52-68 (a1, a2 in the site frame, rotated by k)."""
        return _hip.bond_decompose(a1, a2, k)
