"""Full update of one bond (acetn/evolution/full_update.py:8-343).

Every step runs on the GPU: QR split of the site tensors, norm tensor from
the CTMRG environment, gate application, positive approximation (FP64
Jacobi eigh with the reference's shift loop), gauge fix (Householder R
factors, SVD pseudo-inverses), ALS, and the QR-QR-SVD rebalancing with
recomposition; the Python side only sequences the calls."""
from .. import _hip
from .als_solver import ALSSolver
from .tensor_update import TensorUpdater


class FullUpdater(TensorUpdater):
    def __init__(self, ipeps, gate, config, proposer=None):
        super().__init__(ipeps, gate)
        self.config = config
        self.use_gauge_fix = config.use_gauge_fix
        self.gauge_fix_atol = config.gauge_fix_atol
        self.positive_approx_cutoff = config.positive_approx_cutoff
        self.backend = config.backend
        self.proposer = proposer
        self.last = {}
        self.record = None  # optional list collecting (n12, a12g, a1r, a2r, sweeps, cost)

    def tensor_update(self, a1, a2, bond):
        """full_update.py:34-63 (a1, a2 in the site frame)."""
        s1, s2, k = bond
        a1q, a1r, a2q, a2r = _hip.bond_decompose(a1, a2, k)
        n12 = build_norm_tensor(self.ipeps, bond, a1q, a2q)
        gate = self.gate[bond]
        b1, b2, gy, gx = self.update_reduced_tensors(a1r, a2r, n12, gate)
        return _hip.finalize_bond(b1, b2, gy, gx, a1q, a2q, k)

    def update_reduced_tensors(self, a1r, a2r, n12, gate):
        """full_update.py:65-95; returns the ALS solution and the gauge
        inverses that finalize_reduced_tensors applies."""
        a12g = _hip.apply_gate(a1r, a2r, gate)
        n12, a12g, gx, gy = self.precondition_norm_tensor(n12, a12g)
        solver = ALSSolver(n12, a12g, tuple(a1r.shape), self.config, self.proposer)
        b1, b2 = solver.solve()
        self.last = dict(sweeps=solver.sweeps, cost=solver.cost, start=solver.start,
                         flags=solver.flags)
        if self.record is not None:
            self.record.append(dict(n12=n12, a12g=a12g, a1r=b1, a2r=b2, sweeps=solver.sweeps,
                                    cost=solver.cost))
        return b1, b2, gy, gx

    def precondition_norm_tensor(self, n12, a12g):
        """full_update.py:97-121: (n12, a12g, nzxr_inv, nzyr_inv)."""
        nD, _, d, _ = a12g.shape
        D = nD // d
        nz, _ = _hip.positive_approx(n12, D, d, self.positive_approx_cutoff)
        if self.use_gauge_fix:
            return _hip.gauge_fix(nz, a12g, D, d, self.gauge_fix_atol)
        nzm = nz.reshape(nD * nD, nD * nD)
        n12 = _hip.gemm(nzm, nzm, tb=True).reshape(nD, nD, nD, nD)
        return n12, a12g, None, None


def build_norm_tensor(ipeps, bond, a1q, a2q):
    """full_update.py:163-206."""
    s1, s2, k = bond
    st1, st2 = ipeps[s1], ipeps[s2]
    return _hip.norm_tensor((st1['A'], st1['C'], st1['E']), (st2['A'], st2['C'], st2['E']), k,
                            a1q, a2q)
