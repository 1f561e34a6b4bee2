"""Site tensor with its corner and edge environment
(acetn/ipeps/site_tensor.py:4-202).  A is (l,u,r,d,p); C[k] is (chi,chi);
E[k] is (chi,chi,D_ket,D_bra).  A stand-alone SiteTensor lives on the host
by default, as the reference's does; the networks of the engine put theirs
on the GPU in float64."""
import torch


class SiteTensor:
    def __init__(self, dims, site_tensor=None, site_state=(1., 0.), dtype=torch.float64,
                 device=torch.device("cpu")):
        self.dims = dims
        self.dtype = dtype
        self.device = device
        self.initial_condition = None
        self._site_tensor = None
        self._corner_tensors = None
        self._edge_tensors = None
        if site_tensor is None:
            self.initialize_site_tensor(site_state, noise=1e-2)
            self.initialize_corner_tensors()
            self.initialize_edge_tensors()
        else:
            self.copy_from(site_tensor)

    def __getitem__(self, key):
        if key == 'A':
            return self._site_tensor
        if key == 'C':
            return self._corner_tensors
        if key == 'E':
            return self._edge_tensors
        raise ValueError(f"Invalid key: '{key}' provided.")

    def __setitem__(self, key, val):
        conv = lambda v: torch.clone(v).detach().to(dtype=self.dtype, device=self.device).contiguous()
        if key == 'A':
            self._site_tensor = conv(val)
        elif key == 'C':
            self._corner_tensors = [conv(v) for v in val]
        elif key == 'E':
            self._edge_tensors = [conv(v) for v in val]
        else:
            raise ValueError(f"Invalid key: '{key}' provided.")

    def copy_from(self, site_tensor):
        self['A'] = site_tensor['A']
        self['C'] = site_tensor['C']
        self['E'] = site_tensor['E']

    def copy(self):
        return SiteTensor(self.dims, site_tensor=self, dtype=self.dtype, device=self.device)

    def to(self, dtype=None, device=None):
        return SiteTensor(self.dims, site_tensor=self, dtype=dtype or self.dtype,
                          device=device or self.device)

    def initialize_site_tensor(self, site_state=(1.,), noise=0.0):
        """site_tensor.py:120-136: noise*U[0,1) + site_state at [0,0,0,0,:],
        normalised.  Drawn with torch's CPU generator (torch.manual_seed)."""
        bD, pD = self.dims['bond'], self.dims['phys']
        a = noise * torch.rand(bD, bD, bD, bD, pD, dtype=torch.float64)
        for n in range(len(site_state)):
            a[0, 0, 0, 0, n] += site_state[n]
        self['A'] = a / a.norm()

    def bond_permute(self, k):
        """site_tensor.py:138-150."""
        return self['A'].permute([(i + k) % 4 for i in range(4)] + [4])

    def initialize_corner_tensors(self):
        """site_tensor.py:152-176 (setup-time host contraction of D^4 size)."""
        bD = self.dims['bond']
        a = self['A'].cpu()
        cs = []
        for k in range(4):
            ak = a.permute([(i + k) % 4 for i in range(4)] + [4])
            ck = torch.einsum("lurdp,luRDp->dDrR", ak, ak.conj()).reshape(bD ** 2, bD ** 2)
            cs.append(ck / ck.norm())
        self['C'] = cs

    def initialize_edge_tensors(self):
        """site_tensor.py:178-202."""
        bD = self.dims['bond']
        a = self['A'].cpu()
        es = []
        for k in range(4):
            ak = a.permute([(i + k) % 4 for i in range(4)] + [4])
            ek = torch.einsum("lurdp,LuRDp->lLrRdD", ak, ak.conj()).reshape(bD ** 2, bD ** 2, bD, bD)
            es.append(ek / ek.norm())
        self['E'] = es

    def reset_environment(self):
        """Re-initialise C/E from the current A."""
        self.initialize_corner_tensors()
        self.initialize_edge_tensors()
