"""Full update of one bond (acetn/evolution/full_update.py:8-343).

Every step runs on the GPU: QR split of the site tensors, norm tensor from
the CTMRG environment, gate application, positive approximation (FP64
Jacobi eigh with the reference's shift loop), gauge fix (Householder R
factors, SVD pseudo-inverses), ALS, and the QR-QR-SVD rebalancing with
recomposition; the Python side only sequences the calls.

The methods keep the reference's names and return values.  update() (the
evolution's path) folds the rotation into and out of the bond frame, the
rebalancing and the recomposition into the decomposition and finalisation
kernels; tensor_update() takes bond-frame tensors as the reference's does.
Host tensors handed to the public functions are moved to the GPU and the
results returned on the caller's device."""
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
        self.nzxr_inv = None
        self.nzyr_inv = None
        self.last = {}
        self.record = None  # optional list collecting (n12, a12g, a1r, a2r, sweeps, cost)

    def tensor_update(self, a1, a2, bond):
        """full_update.py:34-63: a1, a2 in the bond frame of ``bond`` (as
        TensorUpdater.update hands them over) -> the updated pair in the bond
        frame, each of unit norm."""
        home = dict(device=a1.device, dtype=a1.dtype)
        a1, a2 = _hip.to_engine(a1, self.ipeps), _hip.to_engine(a2, self.ipeps)
        out = self._update(a1, a2, bond, 0)
        return tuple(t.to(**home) for t in out)

    def site_frame_update(self, a1, a2, bond):
        """tensor_update.py:18-38 for site-frame tensors, the rotation by the
        bond's k folded into the kernels."""
        _, _, k = bond
        return self._update(a1, a2, bond, k)

    def _update(self, a1, a2, bond, k):
        a1q, a1r, a2q, a2r = _hip.bond_decompose(a1, a2, k)
        n12 = build_norm_tensor(self.ipeps, bond, a1q, a2q)
        b1, b2 = self.solve_reduced_tensors(a1r, a2r, n12, self.gate[bond])
        gy, gx = (self.nzyr_inv, self.nzxr_inv) if self.use_gauge_fix else (None, None)
        return _hip.finalize_bond(b1, b2, gy, gx, a1q, a2q, k)

    def update_reduced_tensors(self, a1r, a2r, n12, gate):
        """full_update.py:65-95: gate, preconditioning, ALS and
        finalize_reduced_tensors -> the updated (a1r, a2r)."""
        b1, b2 = self.solve_reduced_tensors(a1r, a2r, n12, gate)
        return self.finalize_reduced_tensors(b1, b2)

    def solve_reduced_tensors(self, a1r, a2r, n12, gate):
        """update_reduced_tensors up to the ALS solution (full_update.py:80-93);
        the gauge inverses stay on the updater for the finalisation."""
        a12g = _hip.apply_gate(a1r, a2r, gate)
        n12, a12g = self.precondition_norm_tensor(n12, a12g)
        solver = ALSSolver(n12, a12g, tuple(a1r.shape), self.config, self.proposer)
        b1, b2 = solver.solve()
        self.last = dict(sweeps=solver.sweeps, cost=solver.cost, start=solver.start,
                         flags=solver.flags)
        if self.record is not None:
            self.record.append(dict(n12=n12, a12g=a12g, a1r=b1, a2r=b2, sweeps=solver.sweeps,
                                    cost=solver.cost))
        return b1, b2

    def precondition_norm_tensor(self, n12, a12g):
        """full_update.py:97-121: (n12, a12g); with the gauge fix the R-factor
        inverses are kept as self.nzxr_inv / self.nzyr_inv."""
        nz, _ = _hip.positive_approx(n12, self.positive_approx_cutoff)
        if self.use_gauge_fix:
            n12, a12g, self.nzxr_inv, self.nzyr_inv = _hip.gauge_fix(nz, a12g,
                                                                     self.gauge_fix_atol)
            return n12, a12g
        nD = nz.shape[0]
        nzm = nz.reshape(nD * nD, nD * nD)
        return _hip.gemm(nzm, nzm, tb=True).reshape(nD, nD, nD, nD), a12g

    def finalize_reduced_tensors(self, a1r, a2r):
        """full_update.py:123-161."""
        gy, gx = (self.nzyr_inv, self.nzxr_inv) if self.use_gauge_fix else (None, None)
        return _hip.finalize_reduced(a1r, a2r, gy, gx)


def build_norm_tensor(ipeps, bond, a1q, a2q):
    """full_update.py:163-206."""
    s1, s2, k = bond
    st1, st2 = ipeps[s1], ipeps[s2]
    return _hip.norm_tensor((st1['A'], st1['C'], st1['E']), (st2['A'], st2['C'], st2['E']), k,
                            a1q, a2q)


def positive_approx(n12, cutoff=1e-12):
    """full_update.py:262-297: nz (nD, nD, nD^2), the square root of the
    shifted norm tensor."""
    nz, _ = _hip.positive_approx(_hip.to_engine(n12), cutoff)
    return nz.to(device=n12.device, dtype=n12.dtype)


def gauge_fix(nz, a12g, atol=1e-12):
    """full_update.py:300-343: (n12, a12g, nzxr_inv, nzyr_inv)."""
    out = _hip.gauge_fix(_hip.to_engine(nz), _hip.to_engine(a12g), atol)
    return tuple(t.to(device=nz.device, dtype=nz.dtype) for t in out)
