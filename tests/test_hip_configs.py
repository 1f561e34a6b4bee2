"""Parity of the HIP path at the sizes of the benchmark configurations, against
the oracle on the same Philox sketches (FP64) and against the FP64 oracle's
energy (mixed precision):

  * D=6, chi=144 (the headline size): one left move;
  * D=4, chi=64: 20 CTMRG iterations;
  * the 12-instance transverse-field Ising sweep (D in {3..6}, chi = 4 D^2)
    through ctmrg_iteration_many;
  * D=6, chi=72: one full-update bond with its two directional moves;
  * D=2, chi=20: evolve(0.01, 100) and CTMRG to convergence, end to end;
  * D=8, chi=256 (oversampling 16): two iterations from a simple-update state,
    FP64 against mixed precision (the oracle is too slow at this size; the
    checks are the energy and the projector spectra).
"""
import copy

import numpy as np
import pytest
import torch

from oracle import ctmrg as oc
from oracle import full_update as ofu
from oracle import measure as om
from oracle import model as omod
from oracle.rng import SketchStream
from test_hip_ctmrg import Cfg, DevState, MixedCfg, assert_state_close, dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from acetn import _hip
    _hip.require_gpu()
    return _hip


def heis_energy(st, nx=2, ny=2):
    return om.measure(st, nx, ny, omod.Heisenberg(1.0))['Energy']


def device_heis_energy(hip, net):
    """Energy per site of the J=1 Heisenberg bond operator from the HIP bond
    reduced density matrices of the 8 bonds of the 2x2 cell (for sizes where
    the oracle's contraction does not fit the host)."""
    from acetn.model.pauli_matrix import pauli_matrices
    X, Y, Z, _ = pauli_matrices(torch.float64)
    h = torch.as_tensor((0.25 * (X * X + Y * Y + Z * Z)).mat, dtype=torch.float64)
    h = h.cpu().numpy().reshape(2, 2, 2, 2)
    bonds = [((x, y), ((x + 1) % 2, y), 2) for x in range(2) for y in range(2)] + \
            [((x, y), (x, (y + 1) % 2), 1) for y in range(2) for x in range(2)]
    e = 0.0
    for s1, s2, k in bonds:
        a, b = net[s1], net[s2]
        rho = hip.bond_rdm((a['A'], a['C'], a['E']), (b['A'], b['C'], b['E']), k).cpu().numpy()
        e += float(np.einsum("PQpq,pqPQ->", rho, h) / np.einsum("pqpq->", rho))
    return e / 4.0


def saturated_state(D, chi, seed, bias, precision_cfg=Cfg, iters=1):
    """Seeded random site tensors, C/E from A, then ``iters`` device CTMRG
    iterations so that every boundary dimension is chi."""
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover
    st0 = oc.make_state(oc.random_site_tensors(2, 2, D, 2, seed=seed, bias=bias))
    hst = DevState(st0, 2, 2, D, chi, seed + 1000)
    mover = DirectionalMover(precision_cfg(), hst.sketch)
    for _ in range(iters):
        ctmrg_iteration(mover, hst)
    return hst.to_numpy()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_headline_left_move_matches_oracle(hip, precision):
    """D=6, chi=144 (N = 5184, q = 146): one left move (two projector pairs,
    two renormalisations) from a chi-saturated environment."""
    from acetn.renormalization.directional_mover import DirectionalMover
    D, chi, seed = 6, 144, 811
    st = saturated_state(D, chi, seed=31, bias=0.5)
    ost = copy.deepcopy(st)
    oc.DirectionalMover(oc.ProjectorCalculator(omega_fn=SketchStream(seed)), 2, 2, chi).left_move(ost, 0)
    hst = DevState(st, 2, 2, D, chi, seed)
    DirectionalMover(Cfg() if precision == "fp64" else MixedCfg(), hst.sketch).left_move(hst, 0)
    h = hst.to_numpy()
    worst = max(np.abs(h[s][key][k] - ost[s][key][k]).max() / np.abs(ost[s][key][k]).max()
                for s in ost for key in ('C', 'E') for k in range(4))
    print(f"D=6 chi=144 left move ({precision}): max relative deviation {worst:.2e}")
    if precision == "fp64":
        assert_state_close(h, ost, rtol=1e-9)
        return
    # Mixed precision moves nearly degenerate singular vectors of the
    # projectors within their subspace, a gauge change of the new boundary
    # tensors: compare what the gauge leaves alone -- the normalised corner
    # spectra and the energy of the moved environment (device RDMs, as the
    # oracle's contraction at this size is slow on the host).  The projected
    # matrix comes from FP32-class (bf16x6) products, so a singular pair
    # s_k is resolved to about 1e-7 s_0 / s_k: with s_min / s_0 ~ 3e-4 here
    # the tail of the kept spectrum, and with it the new corner, moves by up
    # to a few 1e-4 of s_0.
    for s in ost:
        for k in range(4):
            sh = np.linalg.svd(h[s]['C'][k], compute_uv=False)
            so = np.linalg.svd(ost[s]['C'][k], compute_uv=False)
            np.testing.assert_allclose(sh / sh[0], so / so[0], rtol=0, atol=1e-3, err_msg=f"{s} C{k}")
    e_m = device_heis_energy(hip, hst)
    e_o = device_heis_energy(hip, DevState(ost, 2, 2, D, chi, seed))
    print(f"D=6 chi=144 left move (mixed): E/site {e_m:.12f} vs FP64 oracle {e_o:.12f}")
    assert abs(e_m - e_o) <= 1e-6


def test_d4_chi64_twenty_iterations(hip):
    """D=4, chi=64 (N = 1024): 20 CTMRG iterations from seeded N(0,1) site
    tensors; FP64 energy against the oracle, mixed-precision energy within
    1e-6 of the FP64 oracle."""
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover
    D, chi, seed = 4, 64, 4064
    st0 = oc.make_state(oc.random_site_tensors(2, 2, D, 2, seed=64))
    ost = copy.deepcopy(st0)
    omover = oc.DirectionalMover(oc.ProjectorCalculator(omega_fn=SketchStream(seed)), 2, 2, chi)
    h64 = DevState(st0, 2, 2, D, chi, seed)
    m64 = DirectionalMover(Cfg(), h64.sketch)
    hmx = DevState(st0, 2, 2, D, chi, seed)
    mmx = DirectionalMover(MixedCfg(), hmx.sketch)
    for it in range(20):
        oc.ctmrg_iteration(omover, ost)
        ctmrg_iteration(m64, h64)
        ctmrg_iteration(mmx, hmx)
    e_o, e_h, e_m = heis_energy(ost), heis_energy(h64.to_numpy()), heis_energy(hmx.to_numpy())
    print(f"D=4 chi=64, 20 iterations: oracle {e_o:.15f} hip {e_h:.15f} mixed {e_m:.15f} "
          f"(rel {abs(e_h - e_o) / abs(e_o):.1e}, mixed dE {abs(e_m - e_o):.1e})")
    assert e_h == pytest.approx(e_o, rel=1e-10)
    assert abs(e_m - e_o) <= 1e-6


def sweep_instances(seed=3):
    """The 12 instances of the parameter sweep: D drawn from {3,4,5,6} and h
    from [2.5, 3.5] with a seeded generator, chi = 4 D^2, site tensors N(0,1)
    plus 0.5 on the spin-up component."""
    g = torch.Generator().manual_seed(seed)
    Ds = [int(x) for x in torch.randint(3, 7, (12,), generator=g)]
    hs = [float(x) for x in 2.5 + torch.rand(12, generator=g, dtype=torch.float64)]
    return [(D, 4 * D * D, seed * 100 + i, h) for i, (D, h) in enumerate(zip(Ds, hs))]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sweep_instances_together_match_alone_and_oracle(hip, precision):
    """All 12 instances through ctmrg_iteration_many (stage s of every
    instance in one round) equal each instance run alone, bit for bit; in
    FP64 the D=3 and D=4 instances also match the oracle."""
    from acetn.renormalization.ctmrg import ctmrg_iteration, ctmrg_iteration_many
    from acetn.renormalization.directional_mover import DirectionalMover
    cfg = Cfg if precision == "fp64" else MixedCfg
    insts = sweep_instances()
    assert sorted(set(D for D, _, _, _ in insts)) == [3, 4, 5, 6]
    states = {i: oc.make_state(oc.random_site_tensors(2, 2, D, 2, seed=sd, bias=0.5))
              for i, (D, chi, sd, h) in enumerate(insts)}
    pairs, alone = [], []
    for i, (D, chi, sd, h) in enumerate(insts):
        a = DevState(states[i], 2, 2, D, chi, sd)
        b = DevState(states[i], 2, 2, D, chi, sd)
        pairs.append((DirectionalMover(cfg(), a.sketch), a))
        alone.append((DirectionalMover(cfg(), b.sketch), b))
    ctmrg_iteration_many(pairs)
    for m, net in alone:
        ctmrg_iteration(m, net)
    torch.cuda.synchronize()
    for (_, a), (_, b) in zip(pairs, alone):
        for s in a.site_list:
            for key in ('C', 'E'):
                for k in range(4):
                    assert torch.equal(a[s][key][k], b[s][key][k])
    if precision != "fp64":
        return
    checked = 0
    for i, (D, chi, sd, h) in enumerate(insts):
        if D > 4:
            continue
        ost = copy.deepcopy(states[i])
        oc.ctmrg_iteration(oc.DirectionalMover(oc.ProjectorCalculator(omega_fn=SketchStream(sd)),
                                               2, 2, chi), ost)
        assert_state_close(pairs[i][1].to_numpy(), ost, rtol=1e-9)
        e_o = om.measure(ost, 2, 2, omod.Ising(jz=1.0, hx=h))['Energy']
        e_h = om.measure(pairs[i][1].to_numpy(), 2, 2, omod.Ising(jz=1.0, hx=h))['Energy']
        assert e_h == pytest.approx(e_o, rel=1e-10)
        checked += 1
    assert checked >= 2


def test_d6_chi72_full_update_bond_matches_oracle(hip):
    """D=6, chi=72 (nD = 12): one fast-full-update bond (norm tensor,
    positive approximation, gauge fix, ALS, finalize) with its two
    directional moves, on a chi-saturated environment."""
    from acetn.evolution.fast_full_update import FastFullUpdater
    from acetn.evolution.gate import Gate
    from acetn.ipeps.bond import Bond
    from acetn.ipeps.ipeps_config import EvolutionConfig
    from acetn.model.models.heisenberg import HeisenbergModel
    from acetn.renormalization.projectors import SketchCounter
    D, chi, seed = 6, 72, 672
    st = saturated_state(D, chi, seed=72, bias=0.0, iters=2)
    bond = oc.bond_list(2, 2)[0]
    ost = copy.deepcopy(st)
    omover = oc.DirectionalMover(oc.ProjectorCalculator(omega_fn=SketchStream(seed)), 2, 2, chi)
    rec = []
    ofu.bond_update_full(ost, bond, omod.bond_gate(omod.Heisenberg(1.0), 0.01, bond),
                         ofu.FullUpdateConfig(), omover, rec)

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
    gate = Gate(HeisenbergModel(MCfg()), 0.01, net.bond_list, net.site_list)
    upd = FastFullUpdater(net, gate, EvolutionConfig())
    upd.bond_update(net.bond_list[0])
    torch.cuda.synchronize()
    h = {s: {'A': v['A'].cpu().numpy(), 'C': [c.cpu().numpy() for c in v['C']],
             'E': [e.cpu().numpy() for e in v['E']]} for s, v in net.sites.items()}
    dA = max(np.abs(h[s]['A'] - ost[s]['A']).max() / np.abs(ost[s]['A']).max() for s in ost)
    print(f"D=6 chi=72 bond update: max relative deviation of the site tensors {dA:.1e}")
    assert dA < 1e-8
    assert_state_close(h, ost, rtol=1e-8)
    e_o, e_h = heis_energy(ost), heis_energy(h)
    print(f"D=6 chi=72 bond update: ALS sweeps {rec[0]['sweeps']}, E hip {e_h:.15f} oracle {e_o:.15f}")
    assert e_h == pytest.approx(e_o, rel=1e-10)


def _oracle_ctmrg_to_convergence(st, mover, conv_tol, max_steps):
    """ctmrg with the early stop of acetn.renormalization.ctmrg.ctmrg: the
    normalised corner spectra change by less than conv_tol."""
    def spectra(s):
        out = []
        for site in oc.site_list(2, 2):
            for c in s[site]['C']:
                sv = np.linalg.svd(c, compute_uv=False)
                out.append(sv / sv[0])
        return out

    def change(a, b):
        m = 0.0
        for x, y in zip(a, b):
            n = max(len(x), len(y))
            xx, yy = np.zeros(n), np.zeros(n)
            xx[:len(x)] = x
            yy[:len(y)] = y
            m = max(m, float(np.abs(xx - yy).max()))
        return m

    prev = spectra(st)
    done = 0
    for _ in range(max_steps):
        oc.ctmrg_iteration(mover, st)
        done += 1
        cur = spectra(st)
        if change(prev, cur) < conv_tol:
            break
        prev = cur
    return done


def test_d2_chi20_evolve_and_converge_end_to_end(hip):
    """The CPU-runnable configuration end to end through the acetn API:
    Heisenberg J=1 on a 2x2 cell, D=2, chi=20, seeded N(0,1) site tensors of
    unit norm, evolve(0.01, 100) (CTMRG before and after, as Ipeps.evolve
    does) with CTMRG run to a corner-spectrum change below 1e-10 (at most 50
    iterations), then the energy per site; the oracle runs the same protocol
    on the same sketches."""
    from acetn.ipeps import Ipeps
    D, chi, seed = 2, 20, 2020
    A = oc.random_site_tensors(2, 2, D, 2, seed=2)
    st0 = oc.make_state(A)
    config = {'dtype': 'float64', 'device': 'cuda',
              'TN': {'nx': 2, 'ny': 2, 'dims': {'phys': 2, 'bond': D, 'chi': chi}},
              'model': {'name': 'heisenberg', 'params': {'J': 1.0}},
              'ctmrg': {'steps': 50, 'conv_tol': 1e-10, 'sketch_seed': seed,
                        'disable_progressbar': True},
              'evolution': {'disable_progressbar': True}}
    ipeps = Ipeps(config)
    for s in ipeps.site_list:
        ipeps[s]['A'] = dev(st0[s]['A'])
        ipeps[s]['C'] = [dev(c) for c in st0[s]['C']]
        ipeps[s]['E'] = [dev(e) for e in st0[s]['E']]
    ipeps.site_states_initialized = True
    ipeps.evolve(0.01, 100)
    e_h = float(ipeps.measure()['Energy'])

    def oracle_run(st):
        st = copy.deepcopy(st)
        calc = oc.ProjectorCalculator(omega_fn=SketchStream(seed))
        mover = oc.DirectionalMover(calc, 2, 2, chi)
        n1 = _oracle_ctmrg_to_convergence(st, mover, 1e-10, 50)
        ofu.evolve_full(st, 2, 2, omod.Heisenberg(1.0), 0.01, 100, ofu.FullUpdateConfig(), mover,
                        oc.bond_list(2, 2))
        n2 = _oracle_ctmrg_to_convergence(st, mover, 1e-10, 50)
        return heis_energy(st), n1, n2

    e_o, n1, n2 = oracle_run(st0)
    # the oracle's own sensitivity: the same protocol from site tensors moved
    # by one part in 1e15 (a rounding-level change of the input)
    rng = np.random.default_rng(5)
    pert = copy.deepcopy(st0)
    for s in pert:
        pert[s]['A'] = pert[s]['A'] * (1.0 + 1e-15 * rng.standard_normal(pert[s]['A'].shape))
    e_p, _, _ = oracle_run(pert)
    sens = abs(e_p - e_o) / abs(e_o)
    rel = abs(e_h - e_o) / abs(e_o)
    print(f"D=2 chi=20 evolve(0.01, 100): hip {e_h:.15f} oracle {e_o:.15f} (rel {rel:.1e}); "
          f"oracle CTMRG iterations {n1}, {n2}; oracle moved by {sens:.1e} (relative) when "
          f"its input moved by 1e-15")
    assert e_o < 0.0
    # 1e-10, or the spread the oracle itself shows under rounding-level input
    # changes when that is larger (100 full-update steps and CTMRG runs that
    # stop at the 50-iteration limit amplify rounding)
    assert rel <= max(1e-10, 10.0 * sens)


def test_d8_chi256_two_iterations_fp64_vs_mixed(hip):
    """D=8, chi=256 (N = 16384), oversampling 16 (q = 272): site tensors from
    50 device simple-update steps (dtau = 0.05) on seeded N(0,1) tensors, then
    two CTMRG iterations in FP64 and in mixed precision from the same state.
    The oracle is not run at this size: the checks are the mixed-vs-FP64
    energy and the projector spectra (positive, descending, the kept ones
    above the cutoff)."""
    from acetn.evolution.evolve import evolve
    from acetn.ipeps import Ipeps
    from acetn.model.model_factory import model_factory
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover
    D, chi, seed = 8, 256, 8256
    config = {'dtype': 'float64', 'device': 'cuda',
              'TN': {'nx': 2, 'ny': 2, 'dims': {'phys': 2, 'bond': D, 'chi': chi}},
              'model': {'name': 'heisenberg', 'params': {'J': 1.0}},
              'ctmrg': {'steps': 2, 'sketch_seed': seed, 'rsvd_oversampling': 16,
                        'disable_progressbar': True},
              'evolution': {'update_type': 'simple', 'disable_progressbar': True}}
    ipeps = Ipeps(config)
    A = oc.random_site_tensors(2, 2, D, 2, seed=8)
    for s in ipeps.site_list:
        ipeps[s]['A'] = dev(A[s])
    ipeps.site_states_initialized = True
    evolve(ipeps, 0.05, 50, model_factory.create(ipeps.config.model), ipeps.config.evolution)
    st = {}
    for s in ipeps.site_list:
        a = ipeps[s]['A'].cpu().numpy()
        st[s] = a / np.linalg.norm(a)
    st = oc.make_state(st)

    class C5Cfg(Cfg):
        rsvd_oversampling = 16

    class C5Mixed(C5Cfg):
        precision = "mixed"

    energies = {}
    for name, cfg in (("fp64", C5Cfg), ("mixed", C5Mixed)):
        hst = DevState(st, 2, 2, D, chi, seed)
        mover = DirectionalMover(cfg(), hst.sketch)
        for _ in range(2):
            ctmrg_iteration(mover, hst)
        spec = mover.projector_calculator.last_spectrum.cpu().numpy()
        q = chi + 16
        assert spec.shape[0] >= q
        s = spec[:q]
        assert np.all(s > 0) and np.all(np.diff(s) <= 1e-12 * s[0]), name
        assert s[chi - 1] / s[0] > 1e-12
        for site in hst.site_list:
            assert hst[site]['C'][0].shape == (chi, chi)
        energies[name] = device_heis_energy(hip, hst)
        print(f"D=8 chi=256 {name}: E/site {energies[name]:.12f}, s_min/s0 {s[-1] / s[0]:.2e}")
    assert -0.70 < energies["fp64"] < -0.30
    assert abs(energies["mixed"] - energies["fp64"]) <= 1e-6
