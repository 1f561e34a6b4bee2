"""The reference's own unit and integration tests, restated against the HIP
engine through the acetn API: ace-tn tests/unit/test_als.py, test_rdm.py and
tests/integration/test_full_update.py, test_ctmrg.py,
test_directional_mover.py, test_projectors.py, test_ground_states.py.  The
same configurations, calls and assertions (the host tensors those tests hand
over are moved to the GPU by the API and the results come back on the host);
where a reference test checks only shapes, the HIP result is also compared
with the oracle on the same inputs."""
import numpy as np
import pytest
import torch

from conftest import GOLDEN_CASES, load_golden
from oracle import ctmrg as oc
from oracle import full_update as ofu
from oracle import measure as om
from oracle import model as omod
from oracle.ctmrg import ein

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    from acetn import _hip
    _hip.require_gpu()


def npy(t):
    return t.detach().cpu().numpy()


# ---- ALS (tests/unit/test_als.py) -----------------------------------------

def als_case(seed=3):
    """test_als.py:8-33: reduced tensors (nD, bD, pD) = (8, 5, 3), so nD is
    not bD pD; positive_approx and gauge_fix are called with nD as their
    cutoff / atol, as there."""
    from acetn.evolution.full_update import gauge_fix, positive_approx
    g = torch.Generator().manual_seed(seed)
    nD, bD, pD = 8, 5, 3
    a1r = torch.rand(nD, bD, pD, dtype=torch.float64, generator=g)
    a2r = torch.rand(nD, bD, pD, dtype=torch.float64, generator=g)
    a12g = torch.einsum("yup,xuq->yxpq", a1r, a2r)
    a12g += 1e-4 * torch.rand(a12g.shape, dtype=torch.float64, generator=g)
    n12 = torch.rand(nD, nD, nD, nD, dtype=torch.float64, generator=g)
    nz = positive_approx(n12, nD)
    n12g, a12gg, *_ = gauge_fix(nz, a12g, nD)
    return dict(n12_raw=n12, a12g_raw=a12g, nz=nz, n12=n12g, a12g=a12gg, shape=(nD, bD, pD))


def test_als_solver_reference_case():
    """test_als.py:35-63: operand shapes, solution shapes, and the cost after
    99 and 100 sweeps equal to within als_tol.  The inputs and results stay
    on the host, as the reference test's do."""
    from acetn.evolution.als_solver import ALSSolver
    from acetn.ipeps.ipeps_config import EvolutionConfig
    c = als_case()
    config = EvolutionConfig(als_niter=20, als_tol=1e-15, backend="torch")
    solver = ALSSolver(c['n12'], c['a12g'], c['shape'], config)
    assert solver.n12.shape == (8, 8, 8, 8) and solver.a12g.shape == (8, 8, 3, 3)
    a1r, a2r = solver.solve()
    assert a1r.shape == (8, 5, 3) and a2r.shape == (8, 5, 3)
    assert a1r.device.type == "cpu" and a1r.dtype == torch.float64
    solver.niter = 99
    a1r, a2r = solver.solve()
    prev = solver.calculate_cost(a1r, a2r, solver.a12g, solver.n12)
    solver.niter = 100
    a1r, a2r = solver.solve()
    nxt = solver.calculate_cost(a1r, a2r, solver.a12g, solver.n12)
    print(f"ALS nD=8 D=5 d=3: sweeps {solver.sweeps}, cost {float(nxt):.15e}")
    assert abs(nxt - prev) < solver.tol


def test_als_general_reduced_dimension_matches_oracle():
    """The reference test's shapes (nD = 8 with D = 5, d = 3) through the
    oracle: positive_approx with the test's cutoff nD as the square root it
    is (nz nz^T = the shifted n12, whatever the eigenvector signs); then, with
    the default cutoffs, gauge_fix from the same nz, the plain start, the ALS
    sweep count and cost, calculate_cost and n12g."""
    from acetn.evolution.als_solver import ALSSolver
    from acetn.evolution.full_update import gauge_fix, positive_approx
    c = als_case()
    nD, bD, pD = c['shape']
    nz_o = ofu.positive_approx(npy(c['n12_raw']), float(nD))
    nzh = npy(c['nz']).reshape(nD * nD, nD * nD)
    nzo = nz_o.reshape(nD * nD, nD * nD)
    np.testing.assert_allclose(nzh @ nzh.T, nzo @ nzo.T, rtol=0, atol=1e-11 * np.abs(nzo).max() ** 2)
    nz_o = ofu.positive_approx(npy(c['n12_raw']), 1e-12)
    nzh = npy(positive_approx(c['n12_raw'])).reshape(nD * nD, nD * nD)
    nzo = nz_o.reshape(nD * nD, nD * nD)
    np.testing.assert_allclose(nzh @ nzh.T, nzo @ nzo.T, rtol=0, atol=1e-11 * np.abs(nzo).max() ** 2)
    # gauge_fix on the oracle's nz: Householder R factors follow LAPACK's signs
    rn12, ra12g, rxi, ryi = ofu.gauge_fix(nz_o, npy(c['a12g_raw']), 1e-12)
    hn12, ha12g, hxi, hyi = (npy(t) for t in gauge_fix(torch.as_tensor(nz_o), c['a12g_raw']))
    np.testing.assert_allclose(hn12, rn12, rtol=0, atol=1e-10 * np.abs(rn12).max())
    np.testing.assert_allclose(ha12g, ra12g, rtol=0, atol=1e-10 * np.abs(ra12g).max())
    # the plain start and the ALS on the oracle's preconditioned operands
    s1, s2 = ALSSolver.initialize_reduced_tensors(torch.as_tensor(ra12g), (nD, bD, pD))
    o1, o2 = ofu.initialize_reduced_tensors(ra12g, (nD, bD, pD))
    np.testing.assert_allclose(ein("yup,xuq->yxpq", npy(s1), npy(s2)),
                               ein("yup,xuq->yxpq", o1, o2), rtol=0,
                               atol=1e-11 * np.abs(ra12g).max())
    r1, r2, sweeps, cost = ofu.als_solve(rn12, ra12g, (nD, bD, pD), 100, 1e-15, "cholesky",
                                         1e-12)
    from acetn.ipeps.ipeps_config import EvolutionConfig
    solver = ALSSolver(torch.as_tensor(rn12), torch.as_tensor(ra12g), (nD, bD, pD),
                       EvolutionConfig())
    h1, h2 = solver.solve()
    print(f"ALS nD=8 D=5 d=3: oracle {sweeps} sweeps cost {cost:.15e}; "
          f"hip {solver.sweeps} sweeps cost {solver.cost:.15e}")
    assert solver.sweeps == sweeps
    assert solver.cost == pytest.approx(cost, rel=1e-9)
    c_api = float(ALSSolver.calculate_cost(h1, h2, torch.as_tensor(ra12g), torch.as_tensor(rn12)))
    assert c_api == pytest.approx(ofu.calculate_cost(npy(h1), npy(h2), ra12g, rn12), rel=1e-11)
    n12g = ein("yxYX,yxpq->YXpq", rn12, ra12g)
    _, _, h12g = solver.initialize_tensors()
    np.testing.assert_allclose(npy(h12g), n12g, rtol=0, atol=1e-12 * np.abs(n12g).max())


def test_als_float32_operands():
    """The reference's float32 ALS case (tests/unit/test_als_cutensor.py:137-140,
    cost within 5 % of the torch backend's): float32 operands are solved in
    float64 and the results come back in float32; their cost equals the
    oracle's float64 solve of the same (float32-valued) operands."""
    from acetn.evolution.als_solver import ALSSolver
    from acetn.ipeps.ipeps_config import EvolutionConfig
    c = als_case(seed=8)
    n12 = c['n12'].to(torch.float32).cuda()
    a12g = c['a12g'].to(torch.float32).cuda()
    solver = ALSSolver(n12, a12g, c['shape'], EvolutionConfig(als_niter=50, als_tol=1e-15))
    a1r, a2r = solver.solve()
    assert a1r.dtype == torch.float32 and a1r.device.type == "cuda"
    on12, oa12g = npy(n12).astype(np.float64), npy(a12g).astype(np.float64)
    r1, r2, sweeps, cost = ofu.als_solve(on12, oa12g, c['shape'], 50, 1e-15, "cholesky", 1e-12)
    assert solver.sweeps == sweeps
    assert solver.cost == pytest.approx(cost, rel=1e-9)
    c32 = float(ALSSolver.calculate_cost(a1r, a2r, a12g, n12))
    assert c32 == pytest.approx(cost, rel=1e-5)   # the float32 rounding of the results


# ---- full update (tests/integration/test_full_update.py) -------------------

class FuConfig:
    """The reference test's stand-in config (test_full_update.py:22-33)."""
    backend = "torch"
    als_niter = 100
    als_tol = 1e-15
    als_method = "cholesky"
    als_epsilon = 1e-12
    use_gauge_fix = False
    gauge_fix_atol = 1e-12
    positive_approx_cutoff = 1e-12


@pytest.fixture(scope="module")
def fu_setup():
    from acetn.evolution.gate import Gate
    from acetn.ipeps import Ipeps
    from acetn.model.model_factory import model_factory
    torch.manual_seed(5)
    config = {'TN': {'nx': 2, 'ny': 2, 'dims': {'phys': 2, 'bond': 5, 'chi': 8}},
              'model': {'name': 'heisenberg', 'params': {'J': 1.0}},
              'ctmrg': {'disable_progressbar': True}}
    ipeps = Ipeps(config)
    model = model_factory.create(ipeps.config.model)
    gate = Gate(model, 0.01, ipeps.bond_list, ipeps.site_list)
    return ipeps, gate


def test_tensor_update_reference_case(fu_setup):
    """test_full_update.py:36-70: host tensors in, unit-norm tensors of the
    same shape out, and the same result twice."""
    from acetn.evolution.full_update import FullUpdater
    ipeps, gate = fu_setup
    upd = FullUpdater(ipeps, gate, FuConfig())
    g = torch.Generator().manual_seed(6)
    a1 = torch.rand(5, 5, 5, 5, 2, dtype=ipeps.dtype, generator=g)
    a2 = torch.rand(5, 5, 5, 5, 2, dtype=ipeps.dtype, generator=g)
    bond = ipeps.bond_list[0]
    u1, u2 = upd.tensor_update(a1, a2, bond)
    assert u1.shape == a1.shape and u2.shape == a2.shape and u1.device.type == "cpu"
    assert torch.isclose(u1.norm(), torch.tensor(1.0, dtype=ipeps.dtype), atol=1e-6)
    assert torch.isclose(u2.norm(), torch.tensor(1.0, dtype=ipeps.dtype), atol=1e-6)
    v1, v2 = upd.tensor_update(a1, a2, bond)
    assert torch.norm(u1 - v1) < 1e-6 and torch.norm(u2 - v2) < 1e-6
    assert torch.equal(u1, v1) and torch.equal(u2, v2)


@pytest.mark.parametrize("use_gauge_fix", [False, True])
def test_bond_frame_api_equals_fused_site_frame_path(fu_setup, use_gauge_fix):
    """tensor_update on bond-frame tensors (the reference's convention) and
    the evolution's fused site-frame path give the same tensors for every
    bond direction; update_reduced_tensors + recompose_site_tensors (the
    reference's step by step) gives them too."""
    from acetn.evolution.full_update import FullUpdater, build_norm_tensor
    ipeps, gate = fu_setup
    cfg = FuConfig()
    cfg.use_gauge_fix = use_gauge_fix
    upd = FullUpdater(ipeps, gate, cfg)
    g = torch.Generator().manual_seed(7)
    for bond in ipeps.bond_list[:4]:
        s1, s2, k = bond
        a1 = torch.rand(5, 5, 5, 5, 2, dtype=torch.float64, generator=g).cuda()
        a2 = torch.rand(5, 5, 5, 5, 2, dtype=torch.float64, generator=g).cuda()
        f1, f2 = upd.site_frame_update(a1, a2, bond)
        b1, b2 = upd.permute_bond_tensors(a1, a2, k)
        t1, t2 = upd.tensor_update(b1, b2, bond)
        t1, t2 = upd.permute_bond_tensors(t1, t2, 4 - k)
        np.testing.assert_allclose(npy(t1), npy(f1), rtol=0, atol=1e-14)
        np.testing.assert_allclose(npy(t2), npy(f2), rtol=0, atol=1e-14)
        a1q, a1r, a2q, a2r = upd.decompose_site_tensors(b1.contiguous(), b2.contiguous())
        n12 = build_norm_tensor(ipeps, bond, a1q, a2q)
        r1, r2 = upd.update_reduced_tensors(a1r, a2r, n12, gate[bond])
        c1, c2 = upd.recompose_site_tensors(a1q, r1, a2q, r2)
        c1, c2 = upd.permute_bond_tensors(c1 / c1.norm(), c2 / c2.norm(), 4 - k)
        np.testing.assert_allclose(npy(c1), npy(f1), rtol=0, atol=1e-13)
        np.testing.assert_allclose(npy(c2), npy(f2), rtol=0, atol=1e-13)


def test_update_reduced_tensors_matches_oracle():
    """FullUpdater.update_reduced_tensors (gate, positive approximation,
    gauge fix, ALS, finalize_reduced_tensors) on a bond of the converged
    D=3 chi=16 state against the oracle's same steps."""
    from acetn.evolution.full_update import FullUpdater
    from acetn.ipeps.ipeps_config import EvolutionConfig
    st, _, D, chi = load_golden("heisenberg_dims_3_16_dtau_001")
    bond = oc.bond_list(2, 2)[3]
    s1, s2, k = bond
    gate = omod.bond_gate(omod.Heisenberg(1.0), 0.01, bond)
    a1 = np.transpose(st[s1]['A'], ofu.bond_permutation(k))
    a2 = np.transpose(st[s2]['A'], ofu.bond_permutation(k))
    a1q, a1r, a2q, a2r = ofu.decompose_site_tensors(a1, a2)
    n12 = ofu.build_norm_tensor(st, bond, a1q, a2q)
    a12g = ein("yxpq,pqrs->yxrs", ein("yup,xuq->yxpq", a1r, a2r), gate)
    nz = ofu.positive_approx(n12, 1e-12)
    gn12, ga12g, gx, gy = ofu.gauge_fix(nz, a12g, 1e-12)
    b1, b2, _, _ = ofu.als_solve(gn12, ga12g, a1r.shape, 100, 1e-15, "cholesky", 1e-12)
    r1, r2 = ofu.finalize_reduced_tensors(b1, b2, gy, gx)

    class Net:
        dims = {'bond': D, 'phys': 2, 'chi': chi}
    upd = FullUpdater(Net(), None, EvolutionConfig())
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    h1, h2 = upd.update_reduced_tensors(dev(a1r), dev(a2r), dev(n12), dev(gate))
    np.testing.assert_allclose(npy(h1), r1, rtol=0, atol=1e-9)
    np.testing.assert_allclose(npy(h2), r2, rtol=0, atol=1e-9)
    q1, q2 = upd.recompose_site_tensors(dev(a1q), h1, dev(a2q), h2)
    o1, o2 = ofu.recompose_site_tensors(a1q, r1, a2q, r2)
    np.testing.assert_allclose(npy(q1), o1, rtol=0, atol=1e-9)
    np.testing.assert_allclose(npy(q2), o2, rtol=0, atol=1e-9)


# ---- CTMRG pieces (tests/integration/test_ctmrg.py, test_directional_mover.py,
# test_projectors.py) --------------------------------------------------------

def make_ipeps(D, chi, seed, **ctmrg):
    from acetn.ipeps import Ipeps
    torch.manual_seed(seed)
    return Ipeps({'dtype': torch.float64, 'device': torch.device('cpu'),
                  'TN': {'nx': 2, 'ny': 2, 'dims': {'phys': 2, 'bond': D, 'chi': chi}},
                  'ctmrg': dict(ctmrg, disable_progressbar=True)})


def test_ctmrg_updates_every_corner():
    """test_ctmrg.py:30-46: three CTMRG steps with half- and full-system
    projectors change every corner tensor whose shape is kept (the reference
    test asks for device 'cpu'; the engine's network lives on the GPU)."""
    from acetn.renormalization.ctmrg import ctmrg
    ipeps = make_ipeps(3, 9, 21, steps=3)
    assert ipeps[(0, 0)]['A'].device.type == "cuda"
    config = ipeps.config.ctmrg
    for projectors in ("half-system", "full-system"):
        config.projectors = projectors
        before = {(s, k): ipeps[s]['C'][k].clone() for s in ipeps.site_list for k in range(4)}
        ctmrg(ipeps, config)
        for (s, k), c0 in before.items():
            c = ipeps[s]['C'][k]
            if c.shape == c0.shape:
                assert not torch.equal(c, c0)


@pytest.fixture(scope="module")
def mover_setup():
    from acetn.renormalization.directional_mover import DirectionalMover
    ipeps = make_ipeps(6, 36, 22, projectors='half-system')
    return ipeps, DirectionalMover(ipeps.config.ctmrg)


def test_directional_mover_moves_and_projectors(mover_setup):
    """test_directional_mover.py:35-76: the four moves run; the four
    projector calculations return tensors; renormalize_boundary writes the
    C tensor of s2."""
    from acetn.renormalization.directional_mover import DirectionalMover
    ipeps, mover = mover_setup
    assert isinstance(mover, DirectionalMover) and callable(mover.calculate_projectors)
    for name in ("left_move", "up_move", "right_move", "down_move"):
        getattr(mover, name)(ipeps, 0)
    sites = [(0, 0), (1, 0), (1, 1), (0, 1)]
    for name, i in (("calculate_left_projectors", 0), ("calculate_right_projectors", 1),
                    ("calculate_up_projectors", 2), ("calculate_down_projectors", 3)):
        p1, p2 = getattr(mover, name)(ipeps, *sites[i])
        assert isinstance(p1, torch.Tensor) and isinstance(p2, torch.Tensor)
        assert p1.shape[1:3] == (6, 6) and p1.shape == p2.shape
    proj1, proj2 = {}, {}
    for yi in range(2):
        proj1[yi], proj2[yi] = mover.calculate_left_projectors(ipeps, *sites[yi])
    c_before = ipeps[(1, 0)]['C'][0].clone()
    mover.renormalize_boundary(ipeps, proj1, proj2, (0, 0), (1, 0), 0, 1, 0)
    assert isinstance(ipeps[(1, 0)]['C'][0], torch.Tensor)
    assert not torch.equal(ipeps[(1, 0)]['C'][0], c_before)
    assert torch.isfinite(ipeps[(1, 0)]['C'][0]).all()


def test_projector_calculator_reference_cases(mover_setup):
    """test_projectors.py:33-67: calculate() for k = 0..3 with half- and
    full-system projectors; the half-system output (chi, D, D, chi) on the
    default start; a ValueError for sites outside the cell.  (The
    full-system matrix of that start squares the half-system spectrum, so it
    keeps fewer than chi values -- 12 here, as the oracle's does too.)"""
    from acetn.renormalization.projectors import ProjectorCalculator
    ipeps = make_ipeps(6, 36, 23, projectors='half-system')
    sites = [(0, 0), (1, 0), (1, 1), (0, 1)]
    config = ipeps.config.ctmrg
    for proj_type in ("half-system", "full-system"):
        config.projectors = proj_type
        calc = ProjectorCalculator(config)
        assert isinstance(calc, ProjectorCalculator)
        for k in range(4):
            p1, p2 = calc.calculate(ipeps, sites, k)
            assert p1.shape[:3] == (36, 6, 6) and p1.shape == p2.shape
            if proj_type == "half-system":
                assert p1.shape == (36, 6, 6, 36) and p2.shape == (36, 6, 6, 36)
    config.projectors = "half-system"
    calc = ProjectorCalculator(config)
    with pytest.raises(ValueError):
        calc.calculate(ipeps, [(100, 100), (200, 200), (300, 300), (400, 400)], k=0)


# ---- RDMs (tests/unit/test_rdm.py) -----------------------------------------

class MockIpeps:
    """test_rdm.py:7-16: a dict of sites; a missing site reads as {}."""

    def __init__(self):
        self.site_tensors = {}

    def __getitem__(self, key):
        return self.site_tensors.get(key, {})

    def __setitem__(self, key, value):
        self.site_tensors[key] = value


def test_rdm_reference_case():
    """test_rdm.py:18-59: site RDM (d, d) and bond RDM (d,)*4 of host site
    tensors with d = 3, D = 4, returned on the host and equal to the
    oracle's; a missing site raises KeyError."""
    from acetn.ipeps.bond import Bond
    from acetn.ipeps.site_tensor import SiteTensor
    from acetn.measurement.rdm import RDM
    torch.manual_seed(24)
    ipeps = MockIpeps()
    ipeps[(0, 0)] = SiteTensor({'phys': 3, 'bond': 4, 'chi': 5})
    ipeps[(1, 0)] = SiteTensor({'phys': 3, 'bond': 4, 'chi': 5})
    rdm = RDM(ipeps)
    rho = rdm[(0, 0)]
    assert isinstance(rho, torch.Tensor) and rho.shape == (3, 3) and rho.device.type == "cpu"
    bond = Bond((0, 0), (1, 0), 0)
    rho2 = rdm[bond]
    assert isinstance(rho2, torch.Tensor) and rho2.shape == (3, 3, 3, 3)
    with pytest.raises(KeyError):
        rdm[(5, 5)]
    st = {s: {'A': npy(ipeps[s]['A']), 'C': [npy(c) for c in ipeps[s]['C']],
              'E': [npy(e) for e in ipeps[s]['E']]} for s in [(0, 0), (1, 0)]}
    ref = om.site_rdm(st[(0, 0)])
    np.testing.assert_allclose(npy(rho), ref, rtol=1e-11, atol=1e-14 * np.abs(ref).max())
    ref2 = om.bond_rdm(st, ((0, 0), (1, 0), 0))
    np.testing.assert_allclose(npy(rho2), ref2, rtol=1e-11, atol=1e-13 * np.abs(ref2).max())


# ---- ground states (tests/integration/test_ground_states.py) --------------

@pytest.mark.parametrize("case", list(GOLDEN_CASES))
def test_ground_state_energy_and_evolution(case):
    """test_ground_states.py:7-35: the converged state's energy to rel 1e-10,
    then evolve(dtau=0.01, steps=10) (CTMRG of 40 steps with half-system
    projectors, the input files' settings) keeps it to rel 1e-4."""
    from acetn.ipeps import Ipeps
    st, e_ref, D, chi = load_golden(case)
    name, params = GOLDEN_CASES[case]
    ipeps = Ipeps({'dtype': 'float64', 'device': 'cpu',
                   'TN': {'nx': 2, 'ny': 2, 'dims': {'phys': 2, 'bond': D, 'chi': chi}},
                   'model': {'name': name, 'params': params},
                   'ctmrg': {'steps': 40, 'projectors': 'half-system',
                             'disable_progressbar': True},
                   'evolution': {'disable_progressbar': True}})
    ipeps.load_arrays(st)
    assert ipeps.site_states_initialized
    e0 = ipeps.measure()['Energy'].item()
    assert e0 == pytest.approx(e_ref, rel=1e-10)
    ipeps.evolve(dtau=0.01, steps=10)
    e1 = ipeps.measure()['Energy'].item()
    print(f"{case}: fixture {e_ref:.12f} measured {e0:.12f} after evolve(0.01, 10) {e1:.12f}")
    assert e1 == pytest.approx(e_ref, rel=1e-4)
