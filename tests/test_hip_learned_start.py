"""The learned ALS start: the perceptron's device kernels against the
oracle (oracle/mlp.py, bf16 rounding emulated at the same points), training
on the device, and the proposer inside the full-update ALS with the plain
start as fallback and judge."""
import copy

import numpy as np
import pytest
import torch

from conftest import load_golden
from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from acetn import _hip
    _hip.require_gpu()
    return _hip


def dev32(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float32, device="cuda")


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("fin,h,fout,B", [(70, 64, 33, 20), (300, 128, 17, 67)])
def test_mlp_forward_and_gradients_match_oracle(hip, fin, h, fout, B):
    rng = np.random.default_rng(fin)
    n = hip.mlp_param_count(fin, h, fout)
    p = (rng.standard_normal(n) * 0.2).astype(np.float32)
    X = rng.standard_normal((B, fin)).astype(np.float32)
    T = rng.standard_normal((B, fout)).astype(np.float32)
    Yh = hip.mlp_forward(dev32(p), fin, h, fout, dev32(X)).cpu().numpy()
    Yo, _, _ = omlp.forward(p, fin, h, fout, X)
    assert rel(Yh, Yo) < 1e-4
    grad = torch.zeros(n, dtype=torch.float32, device="cuda")
    loss, _ = hip.mlp_backward(dev32(p), fin, h, fout, dev32(X), dev32(T), grad)
    lo, go = omlp.loss_and_grad(p, fin, h, fout, X, T)
    assert loss == pytest.approx(lo, rel=1e-4)
    gh = grad.cpu().numpy().astype(np.float64)
    # per parameter block: W1, b1, W2, b2, W3, b3
    cuts = np.cumsum([0, fin * h, h, h * h, h, h * fout, fout])
    for a, b in zip(cuts[:-1], cuts[1:]):
        assert rel(gh[a:b], go[a:b]) < 2e-3, (a, b, rel(gh[a:b], go[a:b]))


def test_mlp_training_fits_a_linear_map(hip):
    from acetn.evolution.learned_start import MLP, train
    rng = np.random.default_rng(2)
    fin, fout, n = 48, 10, 512
    M = rng.standard_normal((fin, fout)) / np.sqrt(fin)
    X = rng.standard_normal((n, fin))
    mlp = MLP(fin, fout, hidden=128, seed=1)
    losses = train(mlp, dev32(X), dev32(X @ M), epochs=60, batch=64, lr=2e-3, weight_decay=0.0)
    print("losses", losses[0], losses[-1])
    assert losses[-1] < 0.05 * losses[0]


def test_learned_start_keeps_the_solution_and_reports_sweeps(hip):
    """ALS problems recorded from the fast full update of the converged D=3
    fixture (Heisenberg, dtau=0.01); the perceptron is trained on the first
    part and evaluated on the rest from both starts to rel. cost change
    < 1e-10.  The final costs agree and the chosen start never needs more
    sweeps on average than the plain one; the gain is printed."""
    from acetn.evolution.learned_start import (MLP, MLPProposer, dataset, evaluate,
                                               feature_sizes, generate_problems, train)
    from acetn.ipeps import Ipeps
    from acetn.model.model_factory import model_factory
    st, _, D, chi = load_golden("heisenberg_dims_3_16_dtau_001")
    config = {'dtype': 'float64', 'device': 'cuda',
              'TN': {'nx': 2, 'ny': 2, 'dims': {'phys': 2, 'bond': D, 'chi': chi}},
              'model': {'name': 'heisenberg', 'params': {'J': 1.0}},
              'ctmrg': {'steps': 2, 'sketch_seed': 3, 'disable_progressbar': True},
              'evolution': {'disable_progressbar': True}}
    ipeps = Ipeps(config)
    for s in ipeps.site_list:
        ipeps[s]['A'] = torch.as_tensor(st[s]['A'], device="cuda")
        ipeps[s]['C'] = [torch.as_tensor(c, device="cuda") for c in st[s]['C']]
        ipeps[s]['E'] = [torch.as_tensor(e, device="cuda") for e in st[s]['E']]
    model = model_factory.create(ipeps.config.model)
    probs = generate_problems(ipeps, model, 0.01, 24, ipeps.config.evolution)
    train_p, test_p = probs[:160], probs[160:]
    fin, fout = feature_sizes(2 * D, 2)
    X, T = dataset(train_p)
    mlp = MLP(fin, fout, hidden=512, seed=0)
    losses = train(mlp, X, T, epochs=60, batch=32, lr=1e-3)
    res = evaluate(MLPProposer(mlp, ipeps.config.evolution), test_p, ipeps.config.evolution)
    print(f"loss {losses[0]:.3e} -> {losses[-1]:.3e}; {res}")
    assert res["max_rel_cost_diff"] < 1e-8
    assert res["mean_sweeps_learned"] <= res["mean_sweeps_plain"] + 0.25


def test_c4_plain_solves_sit_at_the_three_sweep_floor(hip):
    """Why the learned start cannot cut sweeps at the C4 configuration
    (Heisenberg, D=6, chi=72, dtau=0.01, tol 1e-10): the stopping rule
    `error < tol and i > 1` (als_solver.py:79) makes 3 sweeps the minimum,
    the plain start already reaches it on these problems -- from the first
    sweep of a random start on -- and so does a start at the converged
    solution itself.  (The 20k-problem run of start_workload gives the same
    histogram: all 20000 plain solves take 3 sweeps.)"""
    from acetn.evolution import start_workload as sw
    from acetn.ipeps.ipeps_config import EvolutionConfig
    cfg = EvolutionConfig()
    data = sw.generate(6, 72, 0.01, [7], 2)
    _, _, sweeps, cost, flags = hip.als_solve_batch(data['n12'], data['a12g'], 6, cfg.als_niter,
                                                    1e-10, cfg.als_method, cfg.als_epsilon)
    _, _, s_opt, c_opt, _ = hip.als_solve_batch(data['n12'], data['a12g'], 6, cfg.als_niter, 1e-10,
                                                cfg.als_method, cfg.als_epsilon,
                                                starts=(data['a1r'], data['a2r']))
    torch.cuda.synchronize()
    print(f"{sweeps.numel()} problems: plain sweeps {sw.histogram(sweeps)}, "
          f"from the solution {sw.histogram(s_opt)}")
    assert sweeps.numel() == 16 and not flags.any()
    assert bool((sweeps == 3).all()) and bool((s_opt == 3).all())
    rel = ((cost - c_opt).abs() / c_opt.abs()).max()
    assert float(rel) < 1e-9


def test_als_start_learned_through_the_config(hip, tmp_path):
    """evolution.als_start='learned' with evolution.als_start_model: evolve()
    loads the saved network and every bond's ALS goes through the proposer
    (the plain start judging it); the energy agrees with the plain run to the
    ALS tolerance."""
    from acetn.evolution.evolve import evolve
    from acetn.evolution.learned_start import (MLP, dataset, feature_sizes, generate_problems,
                                               train)
    from acetn.ipeps import Ipeps
    from acetn.model.model_factory import model_factory
    st, _, D, chi = load_golden("heisenberg_dims_3_16_dtau_001")

    def net(start, path=None):
        ev = {'disable_progressbar': True}
        if start == "learned":
            ev.update(als_start="learned", als_start_model=str(path))
        cfg = {'dtype': 'float64', 'device': 'cuda',
               'TN': {'nx': 2, 'ny': 2, 'dims': {'phys': 2, 'bond': D, 'chi': chi}},
               'model': {'name': 'heisenberg', 'params': {'J': 1.0}},
               'ctmrg': {'steps': 2, 'sketch_seed': 3, 'disable_progressbar': True},
               'evolution': ev}
        ipeps = Ipeps(cfg)
        for s in ipeps.site_list:
            ipeps[s]['A'] = torch.as_tensor(st[s]['A'], device="cuda")
            ipeps[s]['C'] = [torch.as_tensor(c, device="cuda") for c in st[s]['C']]
            ipeps[s]['E'] = [torch.as_tensor(e, device="cuda") for e in st[s]['E']]
        ipeps.site_states_initialized = True
        return ipeps

    train_net = net("plain")
    model = model_factory.create(train_net.config.model)
    probs = generate_problems(train_net, model, 0.01, 4, train_net.config.evolution)
    fin, fout = feature_sizes(2 * D, 2)
    mlp = MLP(fin, fout, hidden=128, seed=0)
    X, T = dataset(probs)
    train(mlp, X, T, epochs=10, batch=16, lr=1e-3)
    path = tmp_path / "start.pt"
    mlp.save(str(path))
    energies = {}
    for start in ("plain", "learned"):
        ipeps = net(start, path)
        upd = evolve(ipeps, 0.01, 2, model, ipeps.config.evolution)
        if start == "learned":
            p = upd.proposer
            assert p is not None and p.used + p.rejected == 2 * len(ipeps.bond_list)
            print(f"proposals used {p.used}, rejected {p.rejected}")
        else:
            assert upd.proposer is None
        ipeps.renormalize()
        energies[start] = float(ipeps.measure()['Energy'])
    print(energies)
    assert energies["learned"] == pytest.approx(energies["plain"], rel=1e-7)
    bad = net("learned", tmp_path / "missing.pt")
    bad.config.evolution.als_start_model = None
    with pytest.raises(ValueError, match="als_start_model"):
        evolve(bad, 0.01, 1, model, bad.config.evolution)
