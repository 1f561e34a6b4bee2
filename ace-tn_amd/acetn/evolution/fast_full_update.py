"""Fast full update (acetn/evolution/fast_full_update.py:7-129): after each
bond update the two directional moves that absorb it into the boundary
tensors (single-GPU form of absorb_bond, :70-101)."""
import time

import torch

from ..renormalization.directional_mover import DirectionalMover
from .full_update import FullUpdater


class FastFullUpdater(FullUpdater):
    def __init__(self, ipeps, gate, config, proposer=None):
        super().__init__(ipeps, gate, config, proposer)
        self.mover = DirectionalMover(ipeps.config.ctmrg, ipeps.sketch)

    def bond_update(self, bond):
        """fast_full_update.py:29-68: returns (ctm_time, upd_time)."""
        s1, s2, _ = bond
        t0 = time.perf_counter()
        a1, a2 = self.update(bond)
        self.ipeps[s1]['A'] = a1
        self.ipeps[s2]['A'] = a2
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        self.absorb_bond(bond)
        torch.cuda.synchronize()
        return time.perf_counter() - t1, t1 - t0

    def absorb_bond(self, bond):
        """fast_full_update.py:70-101."""
        s1, s2, k = bond
        if k == 0:
            self.mover.right_move(self.ipeps, s1[0])
            self.mover.left_move(self.ipeps, s2[0])
        elif k == 1:
            self.mover.down_move(self.ipeps, s1[1])
            self.mover.up_move(self.ipeps, s2[1])
        elif k == 2:
            self.mover.left_move(self.ipeps, s1[0])
            self.mover.right_move(self.ipeps, s2[0])
        else:
            self.mover.up_move(self.ipeps, s1[1])
            self.mover.down_move(self.ipeps, s2[1])

    def finalize(self):
        pass
