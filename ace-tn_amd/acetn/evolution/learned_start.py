"""Learned ALS starting guess (no counterpart in the reference).

A 3-layer perceptron (hidden width 512 by default, bf16 matrix-core
products with FP32 accumulation, hand-written forward / backward / Adam
kernels in libacetn_hip) maps one ALS problem of the full update -- the
gauge-fixed norm tensor n12[y][x][Y][X] and the gated bond tensor
a12g[y][x][p][q] -- to a correction of a12p, the rank-D truncation of a12g
(the plain start's bond product a1r a2r, als_solver.py:117-146, which divides
the singular values by the largest, times that largest value):

    a12_proposed = a12p + ||a12g|| * MLP(features)

where the training target is (a12* - a12p) / ||a12g|| with a12* = a1r* a2r*
the converged ALS solution (bond products are bond-gauge invariant, so the
target is well defined; the network only has to learn how the norm tensor's
metric moves the optimum away from the Frobenius truncation).  The proposal
is factored like the plain start (truncated SVD, acetn_hip_als_initialize);
it replaces the plain start only when its cost does not exceed the plain
start's (both evaluated with acetn_hip_als_solve at zero sweeps), so the
plain ALS stays the fallback and the judge.

Features (FP32, one row per problem): a12g/||a12g||, a12p/||a12g||,
n12g/||n12g|| with n12g = sum_{y,x} n12[y][x][Y][X] a12g[y][x][p][q] (the
ALS right-hand side), and n12/||n12||.
"""
import math

import torch

from .. import _hip


def feature_sizes(nD, d):
    nE = nD * nD * d * d
    return 3 * nE + nD ** 4, nE


def features(n12, a12g, a12p):
    """Feature row (FP32, device) of one ALS problem; a12p is the plain
    start's bond product."""
    nD, _, d, _ = a12g.shape
    a = a12g.reshape(nD * nD, d * d)
    n = n12.reshape(nD * nD, nD * nD)
    g = _hip.gemm(n, a, ta=True)                      # n12g[(Y,X)][(p,q)]
    an = a.norm()
    parts = [a / an, a12p / an, g / g.norm(), n / n.norm()]
    return torch.cat([p.reshape(-1) for p in parts]).to(torch.float32)


def rescaled(a12p, a12g):
    """The plain start's product scaled to a12g: the start divides the
    singular values by the largest one (als_solver.py:140), so its product is
    the rank-D truncation of a12g over s_0; alpha = <a12p, a12g> / |a12p|^2
    restores the truncation itself."""
    return a12p * ((a12p * a12g).sum() / (a12p * a12p).sum())


def plain_product(a12g):
    """The rank-D truncation of a12g from the plain start (device)."""
    nD, _, d, _ = a12g.shape
    p1, p2 = _hip.als_initialize(a12g, nD // d)
    return rescaled(bond_product(p1, p2), a12g)


def bond_product(a1r, a2r):
    """a12[y][x][p][q] = sum_u a1r[y][u][p] a2r[x][u][q] on the device."""
    nD, D, d = a1r.shape
    A = a1r.permute(0, 2, 1).reshape(nD * d, D)
    B = a2r.permute(0, 2, 1).reshape(nD * d, D)
    P = _hip.gemm(A, B, tb=True)                      # [(y,p)][(x,q)]
    return P.reshape(nD, d, nD, d).permute(0, 2, 1, 3).contiguous()


def target(a12g, a12p, a1r, a2r):
    return ((bond_product(a1r, a2r) - a12p) / a12g.norm()).reshape(-1).to(torch.float32)


class MLP:
    """Parameters and Adam state of the perceptron, all on the device."""

    def __init__(self, fin, fout, hidden=512, seed=0, device="cuda"):
        self.fin, self.fout, self.hidden = fin, fout, hidden
        n = _hip.mlp_param_count(fin, hidden, fout)
        g = torch.Generator().manual_seed(seed)
        p = torch.zeros(n, dtype=torch.float32)
        off = 0
        for rows, cols in ((fin, hidden), (hidden, hidden), (hidden, fout)):
            w = torch.randn(rows, cols, generator=g, dtype=torch.float32) * math.sqrt(2.0 / rows)
            p[off:off + rows * cols] = w.reshape(-1)
            off += rows * cols + cols                 # bias stays zero
        self.params = p.to(device)
        if fout > 0:                                  # start from the plain method (zero output)
            self.params[off - fout - hidden * fout:off - fout].mul_(0.01)
        self.grad = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0

    def forward(self, X):
        return _hip.mlp_forward(self.params, self.fin, self.hidden, self.fout, X)

    def save(self, path):
        """Plain tensors only (loadable with torch.load(weights_only=True))."""
        torch.save({"format": "acetn-als-start-mlp/1", "fin": self.fin, "hidden": self.hidden,
                    "fout": self.fout, "params": self.params.detach().cpu()}, path)

    @classmethod
    def load(cls, path, device="cuda"):
        st = torch.load(path, weights_only=True, map_location="cpu")
        if st.get("format") != "acetn-als-start-mlp/1":
            raise ValueError(f"{path}: not a saved ALS-start network")
        mlp = cls(int(st["fin"]), int(st["fout"]), hidden=int(st["hidden"]), device=device)
        if mlp.params.numel() != st["params"].numel():
            raise ValueError(f"{path}: parameter count does not match the layer sizes")
        mlp.params.copy_(st["params"].to(device=device, dtype=torch.float32))
        return mlp

    def train_step(self, X, T, lr, weight_decay=0.0):
        loss, _ = _hip.mlp_backward(self.params, self.fin, self.hidden, self.fout, X, T, self.grad)
        self.step += 1
        _hip.mlp_adam_step(self.params, self.grad, self.m, self.v, self.step, lr,
                           weight_decay=weight_decay)
        return loss


def train(mlp, X, T, epochs=40, batch=64, lr=1e-3, weight_decay=1e-5, seed=0):
    """Mini-batch Adam on the device; returns the per-epoch mean losses."""
    g = torch.Generator().manual_seed(seed)
    n = X.shape[0]
    losses = []
    for _ in range(epochs):
        perm = torch.randperm(n, generator=g).to(X.device)
        tot, cnt = 0.0, 0
        for i in range(0, n, batch):
            idx = perm[i:i + batch]
            tot += mlp.train_step(X[idx], T[idx], lr, weight_decay) * len(idx)
            cnt += len(idx)
        losses.append(tot / cnt)
    return losses


def start_cost(n12, a12g, a1r, a2r, config):
    """ALS cost of a start (acetn_hip_als_solve with zero sweeps)."""
    _, _, _, cost, _ = _hip.als_solve(n12, a12g, a1r, a2r, 0, config.als_tol, config.als_method,
                                      config.als_epsilon)
    return cost


class MLPProposer:
    """ALSSolver hook: choose(n12, a12g, a1r, a2r) -> (a1r, a2r, 'learned'|'plain')."""

    def __init__(self, mlp, config):
        self.mlp = mlp
        self.config = config
        self.used = 0
        self.rejected = 0

    def propose(self, n12, a12g, a12p):
        y = self.mlp.forward(features(n12, a12g, a12p).unsqueeze(0))[0]
        return a12p + a12g.norm() * y.to(torch.float64).reshape(a12g.shape)

    def choose(self, n12, a12g, a1r, a2r):
        """(a1r, a2r) is the plain start."""
        bD = a1r.shape[1]
        a12p = rescaled(bond_product(a1r, a2r), a12g)
        p1, p2 = _hip.als_initialize(self.propose(n12, a12g, a12p).contiguous(), bD)
        if start_cost(n12, a12g, p1, p2, self.config) <= start_cost(n12, a12g, a1r, a2r,
                                                                    self.config):
            self.used += 1
            return p1, p2, "learned"
        self.rejected += 1
        return a1r, a2r, "plain"


def solve_from(n12, a12g, a1r, a2r, config, tol):
    """(sweeps, cost) of the device ALS from a given start."""
    _, _, sweeps, cost, _ = _hip.als_solve(n12, a12g, a1r, a2r, config.als_niter, tol,
                                           config.als_method, config.als_epsilon)
    return sweeps, cost


def evaluate(proposer, problems, config, tol=1e-10):
    """Mean ALS sweeps from the plain and from the chosen start, and the
    largest relative difference of the final costs."""
    sp, sl, dc = [], [], 0.0
    for pr in problems:
        n12, a12g = pr['n12'], pr['a12g']
        bD = pr['a1r'].shape[1]
        a1r, a2r = _hip.als_initialize(a12g, bD)
        s0, c0 = solve_from(n12, a12g, a1r, a2r, config, tol)
        b1, b2, _ = proposer.choose(n12, a12g, a1r, a2r)
        s1, c1 = solve_from(n12, a12g, b1, b2, config, tol)
        sp.append(s0)
        sl.append(s1)
        dc = max(dc, abs(c1 - c0) / abs(c0))
    return {"mean_sweeps_plain": sum(sp) / len(sp), "mean_sweeps_learned": sum(sl) / len(sl),
            "max_rel_cost_diff": dc, "problems": len(problems),
            "proposals_used": proposer.used, "proposals_rejected": proposer.rejected}


def dataset(problems):
    """Stack (features, targets) of recorded ALS problems."""
    X, T = [], []
    for p in problems:
        a12p = plain_product(p['a12g'])
        X.append(features(p['n12'], p['a12g'], a12p))
        T.append(target(p['a12g'], a12p, p['a1r'], p['a2r']))
    return torch.stack(X), torch.stack(T)


def generate_problems(ipeps, model, dtau, steps, config):
    """Run ``steps`` fast-full-update sweeps of ``ipeps`` and record every
    bond's ALS problem and converged solution (n12, a12g, a1r, a2r, sweeps,
    cost), the training tuples of the proposer."""
    from .fast_full_update import FastFullUpdater
    from .gate import Gate
    gate = Gate(model, dtau, ipeps.bond_list, ipeps.site_list, device=ipeps.device)
    upd = FastFullUpdater(ipeps, gate, config)
    upd.record = []
    for _ in range(steps):
        for bond in ipeps.bond_list:
            upd.bond_update(bond)
        ipeps.bond_list.reverse()
    return upd.record
