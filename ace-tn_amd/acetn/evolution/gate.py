"""Gates of the imaginary-time evolution (acetn/evolution/gate.py:4-125).

A gate is exp(-dtau H) of a d x d (one-site) or d^2 x d^2 (two-site)
Hamiltonian, computed once per evolve() call as parameter set-up and kept
on the device.  With wrap_one_site (the reference's fixed choice) the bond
gate carries 1/4 of each one-site term."""
import torch


class Gate:
    def __init__(self, model, dtau, bond_list=None, site_list=None, device=None):
        self.dtype = torch.float64
        self.device = device if device is not None else torch.device("cuda")
        self.wrap_one_site = True
        self._gate = {}
        if site_list is not None:
            self.build_one_site_gates(model, dtau, site_list)
        if bond_list is not None:
            self.build_two_site_gates(model, dtau, bond_list)

    def __getitem__(self, key):
        return self._gate[key]

    def __setitem__(self, key, gate):
        self._gate[key] = gate

    @staticmethod
    def _mat(op):
        op = op.mat if hasattr(op, 'mat') else op
        return torch.as_tensor(op, dtype=torch.float64).cpu()

    def build_one_site_gates(self, model, dtau, site_list):
        """gate.py:63-78."""
        for site in site_list:
            site_ham = model.one_site_hamiltonian(site)
            if site_ham is not None:
                scale = 1. if self.wrap_one_site else 0.25
                self[site] = self.calculate_gate(self._mat(site_ham), scale * dtau)

    def build_two_site_gates(self, model, dtau, bond_list):
        """gate.py:80-107."""
        for bond in bond_list:
            pD = model.dim
            bond_ham = self._mat(model.two_site_hamiltonian(bond)).clone()
            if self.wrap_one_site:
                s1, s2, _ = bond
                site_ham_1 = model.one_site_hamiltonian(s1)
                site_ham_2 = model.one_site_hamiltonian(s2)
                if site_ham_1 is not None and site_ham_2 is not None:
                    id_mat = torch.eye(pD, dtype=torch.float64)
                    bond_ham += 0.25 * (torch.kron(self._mat(site_ham_1), id_mat) +
                                        torch.kron(id_mat, self._mat(site_ham_2)))
            self[bond] = self.calculate_gate(bond_ham, dtau).reshape(pD, pD, pD, pD)

    def calculate_gate(self, ham, dtau):
        """gate.py:109-125."""
        return torch.matrix_exp(-dtau * ham).to(dtype=self.dtype, device=self.device).contiguous()
