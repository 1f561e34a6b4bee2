"""Reduced density matrices (acetn/measurement/rdm.py:6-154) on the HIP
engine: acetn_hip_site_rdm / acetn_hip_bond_rdm.  Sites held on the host
are computed on the GPU and their RDMs returned on the host."""
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
        eng = _hip.to_engine
        return eng(st['A']), [eng(c) for c in st['C']], [eng(e) for e in st['E']]

    def build_site_rdm(self, site):
        """rdm.py:35-67."""
        site = tuple(site)
        home = self.ipeps[site]['A']
        return _hip.site_rdm(self._site(site)).to(device=home.device, dtype=home.dtype)

    def build_bond_rdm(self, bond):
        """rdm.py:69-99."""
        s1, s2, k = bond
        home = self.ipeps[s1]['A']
        rho = _hip.bond_rdm(self._site(s1), self._site(s2), k)
        return rho.to(device=home.device, dtype=home.dtype)
