"""Simple update (acetn/evolution/simple_update.py:8-76): one HIP call per
bond (lambdas absorbed, QR split, gate, SVD truncation, lambdas stripped)."""
import time

from .. import _hip
from .lambda_tensors import LambdaTensors
from .tensor_update import TensorUpdater


class SimpleUpdater(TensorUpdater):
    def __init__(self, ipeps, gate, config):
        super().__init__(ipeps, gate)
        self.config = config
        self.bD = ipeps.dims['bond']
        self.lambda_tensors = LambdaTensors(ipeps.bond_list, self.bD, ipeps.dtype, ipeps.device)

    def tensor_update(self, a1, a2, bond):
        raise NotImplementedError("SimpleUpdater overrides bond_update")

    def bond_update(self, bond):
        """simple_update.py:28-76: returns (ctm_time, upd_time)."""
        start = time.time()
        s1, s2, k = bond
        a1, a2, lam = _hip.simple_update_bond(
            self.ipeps[s1]['A'], self.ipeps[s2]['A'], k,
            self.lambda_tensors.legs(s1, k), self.lambda_tensors.legs(s2, (k + 2) % 4),
            self.gate[bond], LambdaTensors.CUTOFF)
        self.lambda_tensors[bond] = lam
        self.ipeps[s1]['A'] = a1
        self.ipeps[s2]['A'] = a2
        return 0, time.time() - start

    def finalize(self):
        pass
