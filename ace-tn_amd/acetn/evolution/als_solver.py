"""ALS bond solver (acetn/evolution/als_solver.py:6-270) on the device.

solve() runs the plain start (SVD of a12g, als_solver.py:117-146) and then
the whole sweep loop with its stopping rule in one persistent HIP workgroup
(acetn_hip_als_solve), in place of solve_torch / solve_cutensor.  With
``als_start='learned'`` the MLP proposal is tried and kept only when its
cost does not exceed the plain start's.  The reduced tensors may have any
reduced dimension nD (ar_shape = (nD, bD, pD)); host tensors are moved to
the GPU and the results returned on the caller's device."""
import torch

from .. import _hip


def _home(t):
    return dict(device=t.device, dtype=t.dtype)


class ALSSolver:
    def __init__(self, n12, a12g, ar_shape, config, proposer=None):
        self.niter = config.als_niter
        self.tol = config.als_tol
        self.method = config.als_method
        self.epsilon = config.als_epsilon
        self.backend = getattr(config, "backend", "hip")
        self.n12 = n12
        self.a12g = a12g
        self.ar_shape = tuple(ar_shape)
        self.proposer = proposer
        self.sweeps = None
        self.cost = None
        self.flags = 0
        self.start = "plain"

    def initialize_tensors(self):
        """als_solver.py:112-115: the plain start (a1r, a2r) and
        n12g = "yxYX,yxpq->YXpq" (n12, a12g)."""
        home = _home(self.a12g)
        n12, a12g = _hip.to_engine(self.n12), _hip.to_engine(self.a12g)
        a1r, a2r = self.initialize_reduced_tensors(a12g, self.ar_shape)
        nD, _, d, _ = a12g.shape
        n12g = _hip.gemm(n12.reshape(nD * nD, nD * nD), a12g.reshape(nD * nD, d * d), ta=True)
        return a1r.to(**home), a2r.to(**home), n12g.reshape(nD, nD, d, d).to(**home)

    @staticmethod
    def initialize_reduced_tensors(a12g, ar_shape):
        """als_solver.py:117-146: truncated SVD of a12g as (nD pD, nD pD),
        split with sqrt(S / S0)."""
        nD, bD, pD = ar_shape
        home = _home(a12g)
        a12g = _hip.to_engine(a12g)
        _hip._expect("a12g", a12g, (nD, nD, pD, pD))
        a1r, a2r = _hip.als_initialize(a12g, bD)
        return a1r.to(**home), a2r.to(**home)

    def solve(self):
        """als_solver.py:34-53."""
        home = _home(self.a12g)
        n12, a12g = _hip.to_engine(self.n12), _hip.to_engine(self.a12g)
        nD, bD, pD = self.ar_shape
        _hip._expect("a12g", a12g, (nD, nD, pD, pD))
        a1r, a2r = _hip.als_initialize(a12g, bD)
        if self.proposer is not None:
            a1r, a2r, self.start = self.proposer.choose(n12, a12g, a1r, a2r)
        a1r, a2r, self.sweeps, self.cost, self.flags = _hip.als_solve(
            n12, a12g, a1r, a2r, self.niter, self.tol, self.method, self.epsilon)
        if self.flags & 2:
            raise RuntimeError("ALS: singular normal equations")
        return a1r.to(**home), a2r.to(**home)

    @staticmethod
    def calculate_cost(a1r, a2r, a12g, n12):
        """als_solver.py:251-261: <a12n|N|a12n> - 2 <a12n|N|a12g> with
        a12n = "yup,xuq->yxpq" (a1r, a2r), as a 0-d tensor (the ALS kernel's
        cost at zero sweeps)."""
        home = _home(a1r)
        a1r, a2r, a12g, n12 = (_hip.to_engine(t) for t in (a1r, a2r, a12g, n12))
        _, _, _, cost, _ = _hip.als_solve(n12, a12g, a1r, a2r, 0, 0.0, "cholesky", 0.0)
        return torch.tensor(cost, **home)
