"""Bond-frame handling of a tensor update (acetn/evolution/tensor_update.py:7-79).

update() hands the site-frame tensors to site_frame_update(), which a
subclass may implement with the rotation into the bond frame and back
folded into its kernels (the full update's decomposition and finalisation
kernels do); the default rotates explicitly around tensor_update(), as the
reference's update() does."""
from abc import ABC, abstractmethod

from .. import _hip


class TensorUpdater(ABC):
    def __init__(self, ipeps, gate):
        self.ipeps = ipeps
        self.dims = ipeps.dims
        self.gate = gate

    @abstractmethod
    def tensor_update(self, a1, a2, bond):
        """a1, a2 in the bond frame -> updated a1, a2 in the bond frame."""

    def update(self, bond):
        """tensor_update.py:18-38 (the reference's one-site gates are wrapped
        into the bond gates, gate.wrap_one_site)."""
        s1, s2, _ = bond
        return self.site_frame_update(self.ipeps[s1]['A'], self.ipeps[s2]['A'], bond)

    def site_frame_update(self, a1, a2, bond):
        """tensor_update() on site-frame tensors, rotating explicitly."""
        _, _, k = bond
        a1, a2 = self.permute_bond_tensors(a1, a2, k)
        a1, a2 = self.tensor_update(a1, a2, bond)
        return self.permute_bond_tensors(a1, a2, 4 - k)

    def permute_bond_tensors(self, a1, a2, k):
        """tensor_update.py:40-44."""
        p = self.bond_permutation(k)
        return a1.permute(p), a2.permute(p)

    @staticmethod
    def bond_permutation(k):
        """tensor_update.py:46-50."""
        return [(i + k) % 4 for i in range(4)] + [4, ]

    @staticmethod
    def decompose_site_tensors(a1, a2, k=0):
        """tensor_update.py:52-68: (a1q, a1r, a2q, a2r) of bond-frame a1, a2
        (k != 0: site-frame tensors rotated by k on the way in)."""
        return _hip.bond_decompose(a1, a2, k)

    @staticmethod
    def recompose_site_tensors(a1q, a1r, a2q, a2r):
        """tensor_update.py:70-75: bond-frame (a1, a2)."""
        return _hip.recompose(a1r, a2r, a1q, a2q, 0, normalize=False)

    def finalize(self):
        """Called after the evolution loop (tensor_update.py:77-79)."""
