"""The learned-start workload at scale: many full-update ALS problems from
seeded random starts, the plain start's sweep histogram, training, and the
comparison of both starts with batched device solves.

Problems come from the method itself: for each seed, iid N(0,1) site
tensors (unit Frobenius norm) on the 2x2 cell, a CTMRG environment, then
fast-full-update sweeps (fast_full_update.py:29-101, gate exp(-dtau H)),
recording every bond's gauge-fixed ALS problem (n12, a12g) and its
converged solution from the first sweep on, so the early, far-from-converged
steps are in the set.  Solves run batched (acetn_hip_als_solve_batch, one
workgroup per problem) to relative cost change < tol from both starts; the
learned start replaces the plain start (als_solver.py:117-146) for a problem
only when its cost does not exceed the plain start's, exactly as
MLPProposer.choose does one problem at a time.

Run as ``python -m acetn.evolution.start_workload --problems 20000`` (from
the package directory on sys.path); writes a JSON report and, with --save,
the trained network (plain tensors, loadable with weights_only=True).
"""
import argparse
import json
import math
import sys
import time

import torch

from .. import _hip
from .learned_start import MLP, dataset, feature_sizes, generate_problems, train


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def random_network(D, chi, seed, ctmrg_steps, model_name="heisenberg", params=None):
    from ..ipeps import Ipeps
    config = {'dtype': 'float64', 'device': 'cuda',
              'TN': {'nx': 2, 'ny': 2, 'dims': {'phys': 2, 'bond': D, 'chi': chi}},
              'model': {'name': model_name, 'params': params or {'J': 1.0}},
              'ctmrg': {'steps': ctmrg_steps, 'sketch_seed': seed, 'disable_progressbar': True},
              'evolution': {'disable_progressbar': True}}
    ipeps = Ipeps(config)
    g = torch.Generator().manual_seed(seed)
    for s in ipeps.site_list:
        a = torch.randn(D, D, D, D, 2, generator=g, dtype=torch.float64)
        ipeps[s]['A'] = a / a.norm()
        ipeps[s].initialize_corner_tensors()
        ipeps[s].initialize_edge_tensors()
    ipeps.site_states_initialized = True
    ipeps.renormalize()
    return ipeps


def generate(D, chi, dtau, seeds, sweeps_per_seed, ctmrg_steps=3):
    """Stacked problems of all seeds: dict of device tensors n12, a12g,
    a1r, a2r (the method's converged solutions), and per problem its seed
    index and sweep number."""
    from ..model.model_factory import model_factory
    parts, seed_of, step_of = [], [], []
    for i, seed in enumerate(seeds):
        t0 = time.perf_counter()
        ipeps = random_network(D, chi, seed, ctmrg_steps)
        model = model_factory.create(ipeps.config.model)
        probs = generate_problems(ipeps, model, dtau, sweeps_per_seed, ipeps.config.evolution)
        nb = len(ipeps.bond_list)
        parts += probs
        seed_of += [i] * len(probs)
        step_of += [j // nb for j in range(len(probs))]
        torch.cuda.synchronize()
        _log(f"seed {seed}: {len(probs)} problems in {time.perf_counter() - t0:.1f} s "
             f"({len(parts)} total)")
    out = {k: torch.stack([p[k] for p in parts]) for k in ('n12', 'a12g', 'a1r', 'a2r')}
    out['seed_index'] = torch.tensor(seed_of)
    out['sweep'] = torch.tensor(step_of)
    return out


def histogram(sweeps):
    vals, counts = torch.unique(sweeps.cpu(), return_counts=True)
    return {int(v): int(c) for v, c in zip(vals, counts)}


def proposals(mlp, n12, a12g, p1, p2, chunk=2048):
    """a12p + ||a12g|| MLP(features) for every problem (device), a12p the
    rank-D truncation of a12g from the plain start (p1, p2)."""
    from .learned_start import bond_product, features, rescaled
    out = torch.empty_like(a12g)
    for i in range(0, a12g.shape[0], chunk):
        js = range(i, min(i + chunk, a12g.shape[0]))
        a12p = [rescaled(bond_product(p1[j], p2[j]), a12g[j]) for j in js]
        X = torch.stack([features(n12[j], a12g[j], a12p[k]) for k, j in enumerate(js)])
        y = mlp.forward(X).to(torch.float64)
        a = a12g[i:i + chunk]
        norms = a.reshape(a.shape[0], -1).norm(dim=1).reshape(-1, 1, 1, 1, 1)
        out[i:i + chunk] = torch.stack(a12p) + norms * y.reshape(a.shape)
    return out


def compare_starts(mlp, n12, a12g, D, config, tol):
    """Both starts on every problem, batched.  Returns per-problem sweeps and
    final costs of the plain and of the chosen start, and the accept mask."""
    niter, meth, eps = config.als_niter, config.als_method, config.als_epsilon
    # plain start and its cost (zero sweeps)
    p1, p2, _, c_plain0, _ = _hip.als_solve_batch(n12, a12g, D, 0, tol, meth, eps)
    # the proposal's factors (plain start of the proposed product) and their cost
    prop = proposals(mlp, n12, a12g, p1, p2)
    l1, l2, _, _, _ = _hip.als_solve_batch(n12, prop, D, 0, tol, meth, eps)
    _, _, _, c_learn0, _ = _hip.als_solve_batch(n12, a12g, D, 0, tol, meth, eps, starts=(l1, l2))
    use = c_learn0 <= c_plain0
    m = use.reshape(-1, 1, 1, 1)
    s1, s2 = torch.where(m, l1, p1), torch.where(m, l2, p2)
    _, _, sw_p, cost_p, fl_p = _hip.als_solve_batch(n12, a12g, D, niter, tol, meth, eps,
                                                    starts=(p1, p2))
    _, _, sw_l, cost_l, fl_l = _hip.als_solve_batch(n12, a12g, D, niter, tol, meth, eps,
                                                    starts=(s1, s2))
    return {"sweeps_plain": sw_p, "sweeps_learned": sw_l, "cost_plain": cost_p,
            "cost_learned": cost_l, "used": use, "flags": fl_p | fl_l}


def summarize(cmp):
    sp, sl = cmp["sweeps_plain"].double(), cmp["sweeps_learned"].double()
    rel = ((cmp["cost_learned"] - cmp["cost_plain"]).abs() / cmp["cost_plain"].abs())
    above = sp > 3
    res = {"problems": int(sp.numel()),
           "mean_sweeps_plain": float(sp.mean()), "mean_sweeps_learned": float(sl.mean()),
           "sweep_reduction": 1.0 - float(sl.mean()) / float(sp.mean()),
           "histogram_plain": histogram(cmp["sweeps_plain"]),
           "histogram_learned": histogram(cmp["sweeps_learned"]),
           "proposals_used": int(cmp["used"].sum()),
           "max_rel_final_cost_diff": float(rel.max()),
           "problems_above_floor": int(above.sum())}
    if int(above.sum()):
        res["above_floor"] = {"mean_sweeps_plain": float(sp[above].mean()),
                              "mean_sweeps_learned": float(sl[above].mean()),
                              "sweep_reduction": 1.0 - float(sl[above].mean()) /
                              float(sp[above].mean())}
    return res


def run(args):
    from ..ipeps.ipeps_config import EvolutionConfig
    cfg = EvolutionConfig()
    D, chi = args.D, args.chi
    nb = 8  # bonds of the 2x2 cell
    sweeps = max(1, math.ceil(args.problems / (nb * len(args.seeds))))
    t0 = time.perf_counter()
    data = generate(D, chi, args.dtau, args.seeds, sweeps, args.ctmrg_steps)
    t_gen = time.perf_counter() - t0
    N = data['a12g'].shape[0]
    # plain-start sweeps of every problem, batched, and the batched throughput
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, _, sw, _, fl = _hip.als_solve_batch(data['n12'], data['a12g'], D, cfg.als_niter, args.tol,
                                           cfg.als_method, cfg.als_epsilon)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    hist_all = histogram(sw)
    by_sweep = {}
    for s in sorted(set(data['sweep'].tolist()))[:5]:
        by_sweep[int(s)] = histogram(sw[(data['sweep'] == s).to(sw.device)])
    _log(f"plain-start sweeps over {N} problems: {hist_all}; first sweeps: {by_sweep}")
    # train on all seeds but the held-out ones
    test_mask = data['seed_index'] >= len(args.seeds) - args.test_seeds
    tr = (~test_mask).nonzero().flatten()
    te = test_mask.nonzero().flatten()
    recs = [{k: data[k][int(i)] for k in ('n12', 'a12g', 'a1r', 'a2r')} for i in tr]
    X, T = dataset(recs)
    # ill-conditioned norm tensors give optima far from a12g (the corrections
    # of a few problems exceed the others by orders of magnitude); such
    # outliers would dominate the squared loss, so rows whose target norm
    # exceeds 10x the median are left out of training
    tn = T.double().norm(dim=1)
    keep = tn <= 10.0 * tn.median()
    X, T = X[keep], T[keep]
    n_out = int((~keep).sum())
    fin, fout = feature_sizes(data['a12g'].shape[1], data['a12g'].shape[3])
    mlp = MLP(fin, fout, hidden=args.hidden, seed=0)
    t0 = time.perf_counter()
    losses = train(mlp, X, T, epochs=args.epochs, batch=args.batch, lr=args.lr)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del X, T
    _log(f"trained on {len(tr)} problems in {t_train:.1f} s, loss {losses[0]:.3e} -> "
         f"{losses[-1]:.3e}")
    ted = te.to(data['n12'].device)
    cmp = compare_starts(mlp, data['n12'][ted], data['a12g'][ted], D, cfg, args.tol)
    res = summarize(cmp)
    if args.save:
        mlp.save(args.save)
    report = {"workload": f"Heisenberg J=1, 2x2 cell, d=2, D={D}, chi={chi}, dtau={args.dtau}; "
                          f"{len(args.seeds)} seeded N(0,1) starts x {sweeps} fast-full-update "
                          f"sweeps x 8 bonds; CTMRG {args.ctmrg_steps} iterations per start",
              "problems": N, "tol": args.tol, "generate_s": t_gen,
              "plain_histogram_all": hist_all, "plain_histogram_first_sweeps": by_sweep,
              "plain_flags_nonzero": int((fl != 0).sum()),
              "batched_als_solves_per_s": N / t_batch,
              "train": {"problems": len(tr) - n_out, "outliers_left_out": n_out,
                        "epochs": args.epochs, "hidden": args.hidden,
                        "seconds": t_train, "loss_first_last": [losses[0], losses[-1]]},
              "test": {"seeds_held_out": args.test_seeds, **res}}
    return report


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--problems", type=int, default=20000)
    ap.add_argument("--D", type=int, default=6)
    ap.add_argument("--chi", type=int, default=72)
    ap.add_argument("--dtau", type=float, default=0.01)
    ap.add_argument("--seeds", type=int, nargs="+", default=list(range(101, 111)))
    ap.add_argument("--test-seeds", type=int, default=2)
    ap.add_argument("--ctmrg-steps", type=int, default=3)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--hidden", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--save", default=None, help="write the trained network here")
    ap.add_argument("--out", default=None, help="write the JSON report here")
    args = ap.parse_args(argv)
    _hip.require_gpu()
    rep = run(args)
    text = json.dumps(rep)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text + "\n")
    print(text, flush=True)


if __name__ == "__main__":
    main()
