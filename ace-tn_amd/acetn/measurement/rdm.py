"""Reduced density matrices (acetn/measurement/rdm.py:6-154) on the HIP
engine: acetn_hip_site_rdm / acetn_hip_bond_rdm."""
from .. import _hip
from ..ipeps.bond import Bond


class RDM:
    def __init__(self, ipeps):
        self.ipeps = ipeps

    def __getitem__(self, key):
        """Bond -> bond RDM (d,d,d,d); site tuple -> site RDM (d,d)."""
        if isinstance(key, Bond):
            return self.build_bond_rdm(key)
        return self.build_site_rdm(key)

    def _site(self, s):
        st = self.ipeps[s]
        return st['A'], st['C'], st['E']

    def build_site_rdm(self, site):
        return _hip.site_rdm(self._site(tuple(site)))

    def build_bond_rdm(self, bond):
        s1, s2, k = bond
        return _hip.bond_rdm(self._site(s1), self._site(s2), k)
