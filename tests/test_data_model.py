"""The data model on the host: SiteTensor, TensorNetwork, Bond and the Pauli
algebra, with the behaviour the reference's unit tests ask of them
(ace-tn tests/unit/test_site_tensor.py, test_tensor_network.py,
test_pauli_matrices.py), restated here.  No GPU: a stand-alone SiteTensor and
a TensorNetwork built with device='cpu' keep their tensors on the host, as
the reference's do."""
import pytest
import torch

from conftest import ROOT  # noqa: F401  (puts the package on sys.path)
from acetn.ipeps.bond import Bond
from acetn.ipeps.ipeps_config import TNConfig
from acetn.ipeps.site_tensor import SiteTensor
from acetn.ipeps.tensor_network import TensorNetwork
from acetn.model.pauli_matrix import PauliMatrix, pauli_matrices

DIMS = {'bond': 4, 'phys': 2, 'chi': 3}


@pytest.fixture
def site():
    torch.manual_seed(11)
    return SiteTensor(dims=DIMS)


def test_site_tensor_layout_and_copy(site):
    """test_site_tensor.py:17-39: A is (D,D,D,D,d) with four corners and
    four edges; a SiteTensor built from another holds the same tensors."""
    D, d = DIMS['bond'], DIMS['phys']
    assert site['A'].shape == (D, D, D, D, d) and site['A'].device.type == "cpu"
    assert len(site['C']) == 4 and len(site['E']) == 4
    # initial environment from A: C (D^2, D^2), E (D^2, D^2, D, D), unit norm
    assert all(c.shape == (D * D, D * D) for c in site['C'])
    assert all(e.shape == (D * D, D * D, D, D) for e in site['E'])
    assert all(abs(float(t.norm()) - 1.0) < 1e-12 for t in site['C'] + site['E'])
    twin = SiteTensor(dims=DIMS, site_tensor=site)
    assert torch.equal(twin['A'], site['A'])
    assert all(torch.equal(a, b) for a, b in zip(twin['C'], site['C']))
    assert all(torch.equal(a, b) for a, b in zip(twin['E'], site['E']))
    twin['A'] = torch.zeros_like(site['A'])   # a copy, not a view
    assert not torch.equal(twin['A'], site['A'])


def test_site_tensor_keys(site):
    """test_site_tensor.py:41-64: 'A', 'C', 'E' get and set; others raise."""
    ones = torch.ones(4, 4, 4, 4, 2)
    site['A'] = ones
    assert torch.equal(site['A'], ones.double())
    with pytest.raises(ValueError):
        site['InvalidKey']
    with pytest.raises(ValueError):
        site['InvalidKey'] = ones


def test_site_tensor_bond_permute_and_noise(site):
    """test_site_tensor.py:66-95: bond_permute(k) rotates the four virtual
    legs cyclically; re-initialising with noise changes A."""
    A = site['A']
    for k in range(4):
        P = site.bond_permute(k)
        assert P.shape == A.shape
        assert torch.equal(P, A.permute([(i + k) % 4 for i in range(4)] + [4]))
    assert not torch.equal(site.bond_permute(1), A)
    before = A.clone()
    site.initialize_site_tensor(site_state=[1., 0.], noise=0.1)
    assert not torch.equal(site['A'], before)
    assert abs(float(site['A'].norm()) - 1.0) < 1e-12
    site.initialize_corner_tensors()
    site.initialize_edge_tensors()
    assert len(site['C']) == 4 and len(site['E']) == 4


@pytest.fixture
def network():
    torch.manual_seed(12)
    cfg = TNConfig(nx=2, ny=2, dims={'phys': 2, 'bond': 3, 'chi': 10})
    return TensorNetwork(None, cfg, torch.float32, torch.device('cpu')), cfg


def test_tensor_network_construction(network):
    """test_tensor_network.py:19-30."""
    tn, cfg = network
    assert (tn.nx, tn.ny) == (cfg.nx, cfg.ny) and tn.dims == cfg.dims
    assert tn.dtype == torch.float32 and tn.device == torch.device('cpu')
    assert len(tn.site_list) == 4 and len(tn.bond_list) == 8
    assert tn[(0, 0)]['A'].dtype == torch.float32


def test_tensor_network_set_copy_save_load(network, tmp_path):
    """test_tensor_network.py:33-78: set a site, copy a network, save and
    load it."""
    tn, cfg = network
    st = SiteTensor(tn.dims, site_state=[1.0, 0.0], dtype=tn.dtype, device=tn.device)
    tn[(0, 0)] = st
    assert tn[(0, 0)] is st
    other = TensorNetwork(None, cfg, tn.dtype, tn.device)
    other.copy_from(tn)
    for s in tn.site_list:
        assert torch.equal(tn[s]['A'], other[s]['A'])
    path = str(tmp_path / "network.pt")
    tn.save(path)
    back = TensorNetwork(None, cfg, tn.dtype, tn.device)
    back.load(path)
    for s in tn.site_list:
        assert torch.equal(tn[s]['A'], back[s]['A'])
        assert all(torch.equal(a, b) for a, b in zip(tn[s]['C'], back[s]['C']))


def test_bond_list(network):
    """test_tensor_network.py:81-96: eight bonds on the 2x2 cell, each
    (s1, s2, k) with k = 1 (up) or 2 (right) from s1, s2 the neighbour."""
    tn, _ = network
    bonds = tn.build_bond_list()
    assert len(bonds) == 8
    for b in bonds:
        assert isinstance(b, Bond) and isinstance(b.s1, tuple) and isinstance(b.s2, tuple)
        assert b.k in (1, 2)
        s1, s2, k = b
        # on the 2 x 2 cell the neighbour one step along x (k = 2) or y (k = 1)
        step = (1, 0) if k == 2 else (0, 1)
        assert s2 == ((s1[0] + step[0]) % 2, (s1[1] + step[1]) % 2)
    assert len({(b.s1, b.s2, b.k) for b in bonds}) == 8


def test_pauli_algebra():
    """test_pauli_matrices.py: 2 x 2 matrices of the requested dtype/device;
    +, -, unary -, scalar * and /; X * Y the Kronecker product; X *= Y the
    bare Kronecker product tensor."""
    X, Y, Z, I = pauli_matrices(torch.float32, torch.device('cpu'))
    for m in (X, Y, Z, I):
        assert isinstance(m, PauliMatrix) and m.mat.shape == (2, 2)
        assert m.mat.dtype == torch.float32 and m.mat.device == torch.device('cpu')
    assert X.dtype == torch.float32 and X.device == torch.device('cpu')
    assert torch.equal((X + Y).mat, X.mat + Y.mat)
    assert torch.equal((X - Z).mat, X.mat - Z.mat)
    assert torch.equal((-X).mat, -X.mat)
    assert torch.equal((X * 2).mat, 2 * X.mat) and torch.equal((2 * X).mat, 2 * X.mat)
    assert torch.equal((X / 2).mat, X.mat / 2)
    assert (X * Y).mat.shape == (4, 4)
    # real-Y convention: Y = [[0,-1],[1,0]] = -i sigma_y, so Y*Y carries a -1
    # to stay the real form of sigma_y (x) sigma_y
    assert torch.equal((Y * Y).mat, -torch.kron(Y.mat, Y.mat))
    x = X
    x *= Y
    assert isinstance(x, torch.Tensor) and x.shape == (4, 4)
    assert torch.equal(x, torch.kron(X.mat, Y.mat))
