"""FP64 parity of the HIP full update and simple update against the oracle
(oracle/full_update.py, a restatement of acetn/evolution/*), through the
C-ABI.  Householder QR follows LAPACK's sign conventions, so the QR factors
are compared directly; quantities behind an eigen- or singular-vector sign
or order choice are compared through gauge-invariant combinations."""
import copy

import numpy as np
import pytest
import torch

from conftest import GOLDEN_CASES, load_golden
from oracle import ctmrg as oc
from oracle import full_update as ofu
from oracle import measure as om
from oracle import model as omod
from oracle.ctmrg import ein
from oracle.rng import SketchStream

pytestmark = pytest.mark.gpu

CASE = "heisenberg_dims_3_16_dtau_001"


@pytest.fixture(scope="module")
def hip():
    from acetn import _hip
    _hip.require_gpu()
    return _hip


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def host(t):
    return t.detach().cpu().numpy()


def dsite(site):
    return (dev(site['A']), [dev(c) for c in site['C']], [dev(e) for e in site['E']])


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def oracle_bond_inputs(st, bond):
    s1, s2, k = bond
    a1 = np.transpose(st[s1]['A'], ofu.bond_permutation(k))
    a2 = np.transpose(st[s2]['A'], ofu.bond_permutation(k))
    return ofu.decompose_site_tensors(a1, a2)


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_bond_decompose_matches_lapack_qr(hip, k):
    A = oc.random_site_tensors(2, 2, 3, 2, seed=40 + k)
    A1, A2 = A[(0, 0)], A[(1, 0)]
    a1 = np.transpose(A1, ofu.bond_permutation(k))
    a2 = np.transpose(A2, ofu.bond_permutation(k))
    ref = ofu.decompose_site_tensors(a1, a2)
    out = hip.bond_decompose(dev(A1), dev(A2), k)
    for o, r in zip(out, ref):
        np.testing.assert_allclose(host(o), r, atol=1e-13)


@pytest.mark.parametrize("bond_idx", range(8))
def test_norm_tensor_matches_oracle(hip, bond_idx):
    st, _, D, chi = load_golden(CASE)
    bond = oc.bond_list(2, 2)[bond_idx]
    s1, s2, k = bond
    a1q, a1r, a2q, a2r = oracle_bond_inputs(st, bond)
    ref = ofu.build_norm_tensor(st, bond, a1q, a2q)
    n12 = host(hip.norm_tensor(dsite(st[s1]), dsite(st[s2]), k, dev(a1q), dev(a2q)))
    assert rel(n12, ref) < 1e-12
    # independent pin: <a12|N|a12> equals the trace of the bond density matrix
    a12 = ein("yup,xuq->yxpq", a1r, a2r)
    val = ein("yxYX,yxpq,YXpq->", n12, a12, a12)
    rho = om.bond_rdm(st, bond)
    assert val == pytest.approx(np.einsum("pqpq->", rho), rel=1e-11)


def test_apply_gate_matches_oracle(hip):
    st, _, D, chi = load_golden(CASE)
    bond = oc.bond_list(2, 2)[3]
    _, a1r, _, a2r = oracle_bond_inputs(st, bond)
    gate = omod.bond_gate(omod.Heisenberg(1.0), 0.01, bond)
    ref = ein("yxpq,pqrs->yxrs", ein("yup,xuq->yxpq", a1r, a2r), gate)
    out = host(hip.apply_gate(dev(a1r), dev(a2r), dev(gate)))
    np.testing.assert_allclose(out, ref, atol=1e-14 * np.abs(ref).max())


def test_positive_approx_matches_oracle(hip):
    st, _, D, chi = load_golden(CASE)
    bond = oc.bond_list(2, 2)[1]
    a1q, a1r, a2q, a2r = oracle_bond_inputs(st, bond)
    n12 = ofu.build_norm_tensor(st, bond, a1q, a2q)
    nD = n12.shape[0]
    for shift_case in (0.0, 1.0):
        # shift_case 1 makes the matrix indefinite so that the shift loop runs
        N = n12 - shift_case * 1e-3 * np.abs(n12).max() * np.eye(nD * nD).reshape(n12.shape)
        ref = ofu.positive_approx(N, 1e-12)
        nz, shifts = hip.positive_approx(dev(N), 1e-12)
        nz = host(nz)
        # nz is fixed up to an orthogonal mixing of its last index
        p_ref = ein("xyz,XYz->xyXY", ref, ref)
        p_hip = ein("xyz,XYz->xyXY", nz, nz)
        assert rel(p_hip, p_ref) < 1e-11
        assert (shifts > 0) == (shift_case > 0)


def test_gauge_fix_matches_oracle(hip):
    st, _, D, chi = load_golden(CASE)
    bond = oc.bond_list(2, 2)[5]
    a1q, a1r, a2q, a2r = oracle_bond_inputs(st, bond)
    n12 = ofu.build_norm_tensor(st, bond, a1q, a2q)
    gate = omod.bond_gate(omod.Heisenberg(1.0), 0.01, bond)
    a12g = ein("yxpq,pqrs->yxrs", ein("yup,xuq->yxpq", a1r, a2r), gate)
    nz = ofu.positive_approx(n12, 1e-12)
    rn12, ra12g, rxi, ryi = ofu.gauge_fix(nz, a12g, 1e-12)
    hn12, ha12g, hxi, hyi = [host(t) for t in hip.gauge_fix(dev(nz), dev(a12g), 1e-12)]
    # the R factors are fixed up to the signs of their rows: compare the
    # gauge-fixed problem through invariants of that sign freedom
    for r, h in ((rxi, hxi), (ryi, hyi)):
        np.testing.assert_allclose(np.abs(h), np.abs(r), atol=1e-10 * np.abs(r).max())
    sx = np.sign(np.sum(hxi * rxi, axis=0))
    sy = np.sign(np.sum(hyi * ryi, axis=0))
    np.testing.assert_allclose(hn12, ein("yxYX,y,x,Y,X->yxYX", rn12, sy, sx, sy, sx),
                               atol=1e-10 * np.abs(rn12).max())
    np.testing.assert_allclose(ha12g, ein("yxpq,y,x->yxpq", ra12g, sy, sx),
                               atol=1e-10 * np.abs(ra12g).max())


def als_problem(bond_idx=2):
    st, _, D, chi = load_golden(CASE)
    bond = oc.bond_list(2, 2)[bond_idx]
    a1q, a1r, a2q, a2r = oracle_bond_inputs(st, bond)
    n12 = ofu.build_norm_tensor(st, bond, a1q, a2q)
    gate = omod.bond_gate(omod.Heisenberg(1.0), 0.01, bond)
    a12g = ein("yxpq,pqrs->yxrs", ein("yup,xuq->yxpq", a1r, a2r), gate)
    nz = ofu.positive_approx(n12, 1e-12)
    n12, a12g, _, _ = ofu.gauge_fix(nz, a12g, 1e-12)
    return n12, a12g, a1r.shape


def test_als_plain_start_matches_oracle(hip):
    n12, a12g, shape = als_problem()
    r1, r2 = ofu.initialize_reduced_tensors(a12g, shape)
    h1, h2 = [host(t) for t in hip.als_initialize(dev(a12g), shape[1])]
    # singular-vector signs are free; the bond product is not
    np.testing.assert_allclose(ein("yup,xuq->yxpq", h1, h2), ein("yup,xuq->yxpq", r1, r2),
                               atol=1e-12 * np.abs(a12g).max())


@pytest.mark.parametrize("bond_idx,tol", [(2, 1e-15), (6, 1e-15), (0, 1e-10)])
def test_als_solve_matches_oracle(hip, bond_idx, tol):
    n12, a12g, shape = als_problem(bond_idx)
    start = ofu.initialize_reduced_tensors(a12g, shape)
    r1, r2, rs, rc = ofu.als_solve(n12, a12g, shape, 100, tol, start=start)
    h1, h2, hs, hc, flags = hip.als_solve(dev(n12), dev(a12g), dev(start[0]), dev(start[1]), 100,
                                          tol, "cholesky", 1e-12)
    print(f"bond {bond_idx}: sweeps oracle {rs} hip {hs}, cost {rc:.15e} {hc:.15e}, flags {flags}")
    assert hs == rs
    assert hc == pytest.approx(rc, rel=1e-11)
    np.testing.assert_allclose(ein("yup,xuq->yxpq", host(h1), host(h2)),
                               ein("yup,xuq->yxpq", r1, r2), atol=1e-9 * np.abs(a12g).max())


@pytest.mark.parametrize("bond_idx", [0, 5])
def test_als_pinv_matches_oracle(hip, bond_idx):
    """method='pinv' (als_solver.py:228-230): the pseudo-inverse solve of the
    symmetrised system with relative cutoff epsilon, sweep for sweep."""
    n12, a12g, shape = als_problem(bond_idx)
    start = ofu.initialize_reduced_tensors(a12g, shape)
    r1, r2, rs, rc = ofu.als_solve(n12, a12g, shape, 100, 1e-12, "pinv", 1e-12, start=start)
    h1, h2, hs, hc, flags = hip.als_solve(dev(n12), dev(a12g), dev(start[0]), dev(start[1]), 100,
                                          1e-12, "pinv", 1e-12)
    print(f"pinv bond {bond_idx}: sweeps oracle {rs} hip {hs}, cost {rc:.15e} {hc:.15e}")
    assert hs == rs
    assert hc == pytest.approx(rc, rel=1e-10)
    np.testing.assert_allclose(ein("yup,xuq->yxpq", host(h1), host(h2)),
                               ein("yup,xuq->yxpq", r1, r2), atol=1e-8 * np.abs(a12g).max())


def test_als_zero_distance_for_exact_product(hip):
    """Closed form: a12g = a1r a2r exactly (rank <= D): the optimum cost is
    -<a12g|N|a12g>."""
    n12, _, shape = als_problem()
    rng = np.random.default_rng(3)
    t1, t2 = rng.standard_normal(shape), rng.standard_normal(shape)
    a12g = ein("yup,xuq->yxpq", t1, t2)
    start = ofu.initialize_reduced_tensors(a12g, shape)
    _, _, _, cost, _ = hip.als_solve(dev(n12), dev(a12g), dev(start[0]), dev(start[1]), 100,
                                     1e-15, "cholesky", 1e-12)
    exact = -ein("yxYX,yxpq,YXpq->", n12, a12g, a12g)
    assert cost == pytest.approx(exact, rel=1e-9)


def hip_tensor_update(hip, st_dev, bond, gate, use_gauge_fix=True):
    s1, s2, k = bond
    A1, A2 = st_dev[s1][0], st_dev[s2][0]
    a1q, a1r, a2q, a2r = hip.bond_decompose(A1, A2, k)
    n12 = hip.norm_tensor(st_dev[s1], st_dev[s2], k, a1q, a2q)
    a12g = hip.apply_gate(a1r, a2r, dev(gate))
    nz, _ = hip.positive_approx(n12, 1e-12)
    n12, a12g, gx, gy = hip.gauge_fix(nz, a12g, 1e-12)
    b1, b2 = hip.als_initialize(a12g, A1.shape[0])
    b1, b2, sweeps, cost, _ = hip.als_solve(n12, a12g, b1, b2, 100, 1e-15, "cholesky", 1e-12)
    return hip.finalize_bond(b1, b2, gy, gx, a1q, a2q, k), sweeps


@pytest.mark.parametrize("bond_idx", [0, 2, 4, 7])
def test_tensor_update_matches_oracle(hip, bond_idx):
    st, _, D, chi = load_golden(CASE)
    bond = oc.bond_list(2, 2)[bond_idx]
    s1, s2, k = bond
    gate = omod.bond_gate(omod.Heisenberg(1.0), 0.01, bond)
    a1 = np.transpose(st[s1]['A'], ofu.bond_permutation(k))
    a2 = np.transpose(st[s2]['A'], ofu.bond_permutation(k))
    rec = []
    r1, r2 = ofu.tensor_update(st, bond, a1, a2, gate, ofu.FullUpdateConfig(), rec)
    r1 = np.transpose(r1, ofu.bond_permutation(4 - k))
    r2 = np.transpose(r2, ofu.bond_permutation(4 - k))
    st_dev = {s: dsite(v) for s, v in st.items()}
    (h1, h2), sweeps = hip_tensor_update(hip, st_dev, bond, gate)
    print(f"bond {bond_idx}: sweeps oracle {rec[0]['sweeps']} hip {sweeps}")
    assert sweeps == rec[0]['sweeps']
    np.testing.assert_allclose(host(h1), r1, atol=1e-9)
    np.testing.assert_allclose(host(h2), r2, atol=1e-9)


def test_fast_full_update_steps_match_oracle(hip):
    """Two evolution steps of the fast full update (8 bond updates, each
    followed by its two directional moves) on the converged fixture, HIP
    engine through the acetn API pieces vs the oracle, same sketches."""
    from acetn.evolution.fast_full_update import FastFullUpdater
    from acetn.evolution.gate import Gate
    from acetn.ipeps.bond import Bond
    from acetn.ipeps.ipeps_config import EvolutionConfig
    from acetn.model.models.heisenberg import HeisenbergModel
    from acetn.renormalization.projectors import SketchCounter
    st, e_ref, D, chi = load_golden(CASE)
    seed = 21
    ost = copy.deepcopy(st)
    ocalc = oc.ProjectorCalculator(omega_fn=SketchStream(seed))
    omover = oc.DirectionalMover(ocalc, 2, 2, chi)
    obonds = oc.bond_list(2, 2)
    ofu.evolve_full(ost, 2, 2, omod.Heisenberg(1.0), 0.01, 2, ofu.FullUpdateConfig(), omover,
                    obonds)

    class Net:
        nx = ny = 2
        dims = {'bond': D, 'chi': chi, 'phys': 2}
        dtype, device = torch.float64, torch.device("cuda")
        site_list = oc.site_list(2, 2)

        class config:
            class ctmrg:
                projectors, svd_type, svd_cutoff = "half-system", "rsvd", 1e-12
                rsvd_niter, rsvd_oversampling, precision = 2, 2, "fp64"

    net = Net()
    net.sites = {s: {'A': dev(v['A']), 'C': [dev(c) for c in v['C']],
                     'E': [dev(e) for e in v['E']]} for s, v in st.items()}
    net.__class__.__getitem__ = lambda self, s: self.sites[tuple(s)]
    net.sketch = SketchCounter(seed)
    net.bond_list = [Bond(*b) for b in oc.bond_list(2, 2)]

    class MCfg:
        dtype, device, dim, name, params = torch.float64, "cpu", 2, "heisenberg", {"J": 1.0}
    model = HeisenbergModel(MCfg())
    gate = Gate(model, 0.01, net.bond_list, net.site_list)
    upd = FastFullUpdater(net, gate, EvolutionConfig())
    for _ in range(2):
        for b in net.bond_list:
            upd.bond_update(b)
        net.bond_list.reverse()
    hst = {s: {'A': host(v['A']), 'C': [host(c) for c in v['C']],
               'E': [host(e) for e in v['E']]} for s, v in net.sites.items()}
    # Two full-update steps on this converged state amplify rounding: relative
    # perturbations of 1e-15 in the corners move the oracle's own site tensors
    # by up to 7e-8 and its energy by 2.2e-10 (relative), so the tensors are
    # compared at 5e-7 and the energy at 1e-9.
    for s in hst:
        np.testing.assert_allclose(hst[s]['A'], ost[s]['A'], atol=5e-7)
    e_o = om.measure(ost, 2, 2, omod.Heisenberg(1.0))['Energy']
    e_h = om.measure(hst, 2, 2, omod.Heisenberg(1.0))['Energy']
    print(f"energy after 2 FU steps: hip {e_h:.15f} oracle {e_o:.15f}")
    assert e_h == pytest.approx(e_o, rel=1e-9)


@pytest.mark.parametrize("k", [1, 2])
def test_simple_update_bond_matches_oracle(hip, k):
    A = oc.random_site_tensors(2, 2, 4, 2, seed=60 + k)
    bonds = oc.bond_list(2, 2)
    bond = [b for b in bonds if b[2] == k][0]
    rng = np.random.default_rng(k)
    lam = ofu.LambdaTensors(bonds, 4)
    for b in bonds:
        lam.data[b] = np.sort(rng.uniform(0.1, 1.0, 4))[::-1].copy()
    gate = omod.bond_gate(omod.Heisenberg(1.0), 0.05, bond)
    st = {s: {'A': a.copy()} for s, a in A.items()}
    ofu.simple_update_bond(st, bond, gate, lam)
    s1, s2, _ = bond

    def legs(site, excl):
        return [None if d == excl else dev(lam.get(site, d)) for d in range(4)]
    # the oracle has already replaced this bond's lambda; the legs it used
    # are those of the other bonds, unchanged
    h1, h2, hl = hip.simple_update_bond(dev(A[s1]), dev(A[s2]), k, legs(s1, k),
                                        legs(s2, (k + 2) % 4), dev(gate))
    np.testing.assert_allclose(host(hl), lam.data[bond], atol=1e-12)
    np.testing.assert_allclose(host(h1), st[s1]['A'], atol=1e-11)
    np.testing.assert_allclose(host(h2), st[s2]['A'], atol=1e-11)


def test_ipeps_evolve_api_matches_oracle(hip):
    """The acetn API end to end: Ipeps(config), evolve() (CTMRG, fast full
    update with its moves, CTMRG) and measure() against the oracle on the
    converged fixture with the same sketches.  (From a random state a few
    CTMRG iterations are too far from the fixed point for a 1e-10 bound:
    there the energy moves by ~1e-4 when only the sketch seed changes.)"""
    from acetn.ipeps import Ipeps
    st, e_ref, D, chi = load_golden(CASE)
    seed = 17
    config = {'dtype': 'float64', 'device': 'cuda',
              'TN': {'nx': 2, 'ny': 2, 'dims': {'phys': 2, 'bond': D, 'chi': chi}},
              'model': {'name': 'heisenberg', 'params': {'J': 1.0}},
              'ctmrg': {'steps': 2, 'sketch_seed': seed, 'disable_progressbar': True},
              'evolution': {'disable_progressbar': True}}
    ipeps = Ipeps(config)
    for s in ipeps.site_list:
        ipeps[s]['A'] = dev(st[s]['A'])
        ipeps[s]['C'] = [dev(c) for c in st[s]['C']]
        ipeps[s]['E'] = [dev(e) for e in st[s]['E']]
    ipeps.site_states_initialized = True
    ipeps.evolve(0.01, 1)
    e_h = float(ipeps.measure()['Energy'])

    ost = copy.deepcopy(st)
    calc = oc.ProjectorCalculator(omega_fn=SketchStream(seed))
    mover = oc.DirectionalMover(calc, 2, 2, chi)
    oc.ctmrg(ost, 2, 2, chi, 2, calc)
    ofu.evolve_full(ost, 2, 2, omod.Heisenberg(1.0), 0.01, 1, ofu.FullUpdateConfig(), mover,
                    oc.bond_list(2, 2))
    oc.ctmrg(ost, 2, 2, chi, 2, calc)
    e_o = om.measure(ost, 2, 2, omod.Heisenberg(1.0))['Energy']
    print(f"Ipeps.evolve: hip {e_h:.15f} oracle {e_o:.15f} fixture {e_ref:.15f}")
    assert e_h == pytest.approx(e_o, rel=1e-10)
    assert e_h == pytest.approx(e_ref, rel=1e-4)   # reference criterion, test_ground_states.py:33-35


@pytest.mark.parametrize("method", ["cholesky", "pinv"])
def test_als_batch_equals_single_solves(hip, method):
    """acetn_hip_als_solve_batch (one workgroup per problem) against one
    acetn_hip_als_initialize + acetn_hip_als_solve per problem: the same
    kernels on the same data, so bit-equal results; with given starts too."""
    probs = [als_problem(b) for b in range(8)]
    shape = probs[0][2]
    n12 = torch.stack([dev(p[0]) for p in probs])
    a12g = torch.stack([dev(p[1]) for p in probs])
    b1, b2, sw, cost, fl = hip.als_solve_batch(n12, a12g, shape[1], 100, 1e-12, method, 1e-12)
    torch.cuda.synchronize()
    for i in range(len(probs)):
        s1, s2 = hip.als_initialize(a12g[i], shape[1])
        h1, h2, hs, hc, hf = hip.als_solve(n12[i], a12g[i], s1, s2, 100, 1e-12, method, 1e-12)
        assert int(sw[i]) == hs and float(cost[i]) == hc and int(fl[i]) == hf, i
        assert torch.equal(b1[i], h1) and torch.equal(b2[i], h2), i
    # from given starts (the plain starts, perturbed)
    g = torch.Generator(device="cuda").manual_seed(4)
    st = [hip.als_initialize(a12g[i], shape[1]) for i in range(len(probs))]
    s1 = torch.stack([s[0] for s in st])
    s2 = torch.stack([s[1] for s in st])
    s1 = s1 + 1e-3 * torch.randn(s1.shape, generator=g, device="cuda", dtype=torch.float64)
    c1, c2, sw2, cost2, _ = hip.als_solve_batch(n12, a12g, shape[1], 100, 1e-12, method, 1e-12,
                                                starts=(s1, s2))
    for i in range(len(probs)):
        h1, h2, hs, hc, _ = hip.als_solve(n12[i], a12g[i], s1[i], s2[i], 100, 1e-12, method, 1e-12)
        assert int(sw2[i]) == hs and float(cost2[i]) == hc, i
        assert torch.equal(c1[i], h1) and torch.equal(c2[i], h2), i
    print(f"{method}: sweeps {sw.tolist()} (perturbed starts {sw2.tolist()})")
