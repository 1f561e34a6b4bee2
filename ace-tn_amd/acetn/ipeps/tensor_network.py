"""Lattice container (acetn/ipeps/tensor_network.py:5-193).

Unlike the reference (whose site dictionary is a class attribute shared by
every instance, tensor_network.py:12), each network owns its sites, so
independent iPEPS instances (e.g. a parameter sweep) can be co-resident."""
import torch

from .bond import Bond
from .site_tensor import SiteTensor


class TensorNetwork:
    def __init__(self, tensor_network, config, dtype, device):
        self.tn_config = config
        self.nx = config.nx
        self.ny = config.ny
        self.dims = config.dims
        self.dtype = dtype
        self.device = device
        self._tensor_network = {}
        self.setup_tensor_network(tensor_network)

    def __getitem__(self, site):
        st = self._tensor_network.get(tuple(site))
        if st is None:
            raise ValueError(f"Site tensor not defined at site {site}.")
        return st

    def __setitem__(self, site, site_tensor):
        self._tensor_network[tuple(site)] = site_tensor

    def copy_from(self, tensor_network):
        self.site_list = tensor_network.site_list.copy()
        self.bond_list = tensor_network.bond_list.copy()
        for site in self.site_list:
            self[site] = SiteTensor(self.dims, tensor_network[site], dtype=self.dtype,
                                    device=self.device)

    def copy(self):
        return TensorNetwork(self, self.tn_config, self.dtype, self.device)

    def to(self, dtype=None, device=None):
        return TensorNetwork(self, self.tn_config, dtype or self.dtype, device or self.device)

    def save(self, filename='ipeps.pt'):
        """Plain-tensor checkpoint: {site: {'A','C','E'}} on the CPU
        (tensor_network.py:94-109 stores pickled objects instead)."""
        state = {"format": "acetn-mi355x/1", "nx": self.nx, "ny": self.ny, "dims": dict(self.dims),
                 "tensors": {f"{s[0]},{s[1]}": {"A": self[s]['A'].cpu(),
                                                  "C": [c.cpu() for c in self[s]['C']],
                                                  "E": [e.cpu() for e in self[s]['E']]}
                             for s in self.site_list}}
        torch.save(state, filename)

    def load(self, filename):
        """tensor_network.py:111-122: reads the engine's plain-tensor format of
        save() (torch.load with weights_only=True), or a checkpoint written by
        the reference's save(), whose pickled acetn objects are read without
        executing the pickle (acetn/ipeps/checkpoint.py)."""
        from .checkpoint import is_reference_checkpoint, read_reference_checkpoint
        if is_reference_checkpoint(filename):
            self.load_arrays(read_reference_checkpoint(filename, self.nx, self.ny))
            return
        state = torch.load(filename, weights_only=True)
        for key, t in state["tensors"].items():
            site = tuple(int(v) for v in key.split(","))
            st = self[site]
            st['A'] = t['A']
            st['C'] = t['C']
            st['E'] = t['E']
        self.site_states_initialized = True

    def load_arrays(self, arrays):
        """Load {site: {'A','C','E'}} of numpy/torch arrays."""
        for site, t in arrays.items():
            st = self[site]
            st['A'] = torch.as_tensor(t['A'])
            st['C'] = [torch.as_tensor(c) for c in t['C']]
            st['E'] = [torch.as_tensor(e) for e in t['E']]
        self.site_states_initialized = True

    def setup_tensor_network(self, tensor_network):
        if tensor_network is None:
            uniform = lambda site: [1.] + [0.] * (self.dims['phys'] - 1)
            self.initialize_site_tensors(uniform)
        else:
            self.copy_from(tensor_network)

    def initialize_site_tensors(self, site_state_map):
        self.bond_list = self.build_bond_list()
        self.site_list = self.build_site_list()
        for site in self.site_list:
            self[site] = SiteTensor(self.dims, site_state=site_state_map(site), dtype=self.dtype,
                                    device=self.device)
        self.site_states_initialized = True

    def build_bond_list(self):
        """tensor_network.py:153-177: horizontal (k=2) then vertical (k=1)."""
        bonds = []
        for xi in range(self.nx):
            for yi in range(self.ny):
                bonds.append(Bond((xi, yi), ((xi + 1) % self.nx, yi), 2))
        for yi in range(self.ny):
            for xi in range(self.nx):
                bonds.append(Bond((xi, yi), (xi, (yi + 1) % self.ny), 1))
        return bonds

    def build_site_list(self):
        return [(xi, yi) for xi in range(self.nx) for yi in range(self.ny)]
