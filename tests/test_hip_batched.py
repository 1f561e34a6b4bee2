"""Independent iPEPS instances of uneven (D, chi) advanced together
(ctmrg_iteration_many: stage s of every instance in one round, projector
pairs spread over the stream slots with a workspace each).  Each instance
must come out bit-for-bit as when it runs alone, and match the oracle."""
import copy

import numpy as np
import pytest
import torch

from oracle import ctmrg as oc
from oracle.rng import SketchStream
from test_hip_ctmrg import Cfg, DevState, assert_state_close

pytestmark = pytest.mark.gpu

# (D, chi, seed): uneven shapes, as in a parameter sweep
INSTANCES = [(2, 8, 11), (3, 12, 12), (2, 12, 13), (3, 9, 14), (4, 16, 15)]


@pytest.fixture(scope="module")
def hip():
    from acetn import _hip
    _hip.require_gpu()
    return _hip


def _states():
    return [oc.make_state(oc.random_site_tensors(2, 2, D, 2, seed=300 + s))
            for D, chi, s in INSTANCES]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_instances_equal_single_runs(hip, precision):
    from acetn.renormalization.ctmrg import ctmrg_iteration, ctmrg_iteration_many
    from acetn.renormalization.directional_mover import DirectionalMover
    cfg = Cfg()
    cfg.precision = precision
    sts = _states()
    alone, together = [], []
    for (D, chi, seed), st in zip(INSTANCES, sts):
        a = DevState(st, 2, 2, D, chi, seed)
        b = DevState(st, 2, 2, D, chi, seed)
        alone.append((DirectionalMover(cfg, a.sketch), a))
        together.append((DirectionalMover(cfg, b.sketch), b))
    for it in range(3):
        for m, net in alone:
            ctmrg_iteration(m, net)
        ctmrg_iteration_many(together)
        torch.cuda.synchronize()
        for (_, a), (_, b), inst in zip(alone, together, INSTANCES):
            na, nb = a.to_numpy(), b.to_numpy()
            for s in na:
                for key in ('C', 'E'):
                    for k in range(4):
                        x, y = na[s][key][k], nb[s][key][k]
                        assert x.shape == y.shape, (inst, it, s, key, k)
                        assert np.array_equal(x, y), (inst, it, s, key, k,
                                                      float(np.abs(x - y).max()))


def test_batched_instances_match_oracle(hip):
    from acetn.renormalization.ctmrg import ctmrg_iteration_many
    from acetn.renormalization.directional_mover import DirectionalMover
    sts = _states()
    nets, movers, orc = [], [], []
    for (D, chi, seed), st in zip(INSTANCES, sts):
        net = DevState(st, 2, 2, D, chi, seed)
        nets.append(net)
        movers.append(DirectionalMover(Cfg(), net.sketch))
        ost = copy.deepcopy(st)
        om = oc.DirectionalMover(oc.ProjectorCalculator(omega_fn=SketchStream(seed)), 2, 2, chi)
        orc.append((om, ost))
    for it in range(2):
        ctmrg_iteration_many(list(zip(movers, nets)))
        for om, ost in orc:
            oc.ctmrg_iteration(om, ost)
    for net, (_, ost) in zip(nets, orc):
        assert_state_close(net.to_numpy(), ost, rtol=1e-9)
