"""Bond weights of the simple update (acetn/evolution/lambda_tensors.py:5-49),
kept on the device; absorbing and stripping them is done inside
acetn_hip_simple_update_bond."""
import torch


class LambdaTensors:
    CUTOFF = 1e-12

    def __init__(self, bond_list, bD, dtype=torch.float64, device="cuda"):
        self._data = {}
        self._site_bonds = {}
        for bond in bond_list:
            self._data[bond] = torch.ones(bD, dtype=dtype, device=device)
            self._site_bonds.setdefault(bond.s1, {})[bond.k] = bond
            self._site_bonds.setdefault(bond.s2, {})[(bond.k + 2) % 4] = bond

    def __getitem__(self, bond):
        return self._data[bond]

    def __setitem__(self, bond, value):
        self._data[bond] = value

    def get(self, site, direction):
        """The lambda vector of a site along a direction (0=l, 1=u, 2=r, 3=d)."""
        return self._data[self._site_bonds[site][direction]]

    def legs(self, site, exclude_dir):
        """The four leg vectors of a site with the excluded leg as None."""
        return [None if d == exclude_dir else self.get(site, d) for d in range(4)]
