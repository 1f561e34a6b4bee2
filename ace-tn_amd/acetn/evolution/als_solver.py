"""ALS bond solver (acetn/evolution/als_solver.py:6-270) on the device.

initialize_tensors() is the plain start (SVD of a12g, als_solver.py:117-146);
solve() runs the whole sweep loop with its stopping rule in one persistent
HIP workgroup (acetn_hip_als_solve), in place of solve_torch /
solve_cutensor.  With ``als_start='learned'`` the MLP proposal is tried
and kept only when its cost does not exceed the plain start's."""
from .. import _hip


class ALSSolver:
    def __init__(self, n12, a12g, ar_shape, config, proposer=None):
        self.niter = config.als_niter
        self.tol = config.als_tol
        self.method = config.als_method
        self.epsilon = config.als_epsilon
        self.backend = getattr(config, "backend", "hip")
        self.n12 = n12
        self.a12g = a12g
        self.ar_shape = ar_shape
        self.proposer = proposer
        self.sweeps = None
        self.cost = None
        self.flags = 0
        self.start = "plain"

    def initialize_tensors(self):
        """als_solver.py:112-146: (a1r, a2r) of the plain start."""
        nD, bD, pD = self.ar_shape
        return _hip.als_initialize(self.a12g, bD)

    def solve(self):
        """als_solver.py:34-53."""
        a1r, a2r = self.initialize_tensors()
        if self.proposer is not None:
            a1r, a2r, self.start = self.proposer.choose(self.n12, self.a12g, a1r, a2r)
        a1r, a2r, self.sweeps, self.cost, self.flags = _hip.als_solve(
            self.n12, self.a12g, a1r, a2r, self.niter, self.tol, self.method, self.epsilon)
        if self.flags & 2:
            raise RuntimeError("ALS: singular normal equations")
        return a1r, a2r
