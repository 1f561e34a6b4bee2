"""Record the reference's own trajectories: what only a run of the reference
fixes.

The reference's CTMRG draws every range-finder sketch with torch.randn from
torch's global generator (acetn/linalg/fused_matmul_svd_lowrank.py:32), one
call per projector pair in call order.  The oracle restates the algorithm
with the sketch as an input; these vectors pin that restatement against the
reference run end to end on the same sketches (tests/test_oracle.py replays
them with torch.randn in the same order):

  * ctmrg_*: Heisenberg J=1, 2x2 cell, seeded N(0,1) site tensors of unit
    norm, the reference's own C/E initialisation (site_tensor.py:152-202),
    then Ipeps.renormalize() (ctmrg.py:4-31) for ``steps`` iterations with
    torch.manual_seed(seed) right before it: initial and final C/E, energy;
  * evolve_*: the same start, then Ipeps.evolve(dtau, n) (CTMRG, fast full
    update with its moves, CTMRG; ipeps.py:107-123): final A/C/E, energy;
    evolve_d2_chi20_100 is the D=2, chi=20, evolve(0.01, 100) protocol.

The reference is imported from the checkout in this container.  It imports
``toml`` for its log messages only (acetn/utils/logger.py:3), which this
image lacks: a stub with dumps() stands in for it.  Nothing else is replaced.

Usage:  python tests/golden/make_reference_trajectory.py /root/reference
Writes tests/golden/reference_trajectory.npz.
"""
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SITES = [(0, 0), (0, 1), (1, 0), (1, 1)]

CASES = {
    # name: (D, chi, A seed, torch seed, ctmrg steps, evolve (dtau, steps) or None)
    "ctmrg_d2_chi8": (2, 8, 11, 1234, 6, None),
    "ctmrg_d3_chi12": (3, 12, 12, 4321, 4, None),
    "evolve_d2_chi8": (2, 8, 13, 777, 3, (0.01, 4)),
    # the CPU-runnable benchmark protocol: D=2, chi=20, evolve(0.01, 100) with
    # 50 CTMRG iterations before and after (the reference's ctmrg has no
    # convergence test, so a fixed count)
    "evolve_d2_chi20_100": (2, 20, 2, 2020, 50, (0.01, 100)),
}


def random_site_tensors(D, d, seed):
    """Seeded iid N(0,1) site tensors of unit Frobenius norm, in the x-major
    site order (the same draw as oracle.ctmrg.random_site_tensors)."""
    sys.path.insert(0, ROOT)
    from oracle.ctmrg import random_site_tensors as rst
    return rst(2, 2, D, d, seed=seed)


def main(ref_root):
    stub = types.ModuleType("toml")
    stub.dumps = lambda d: repr(d)
    sys.modules.setdefault("toml", stub)
    sys.path.insert(0, ref_root)
    import torch
    from acetn.ipeps import Ipeps

    out = {}
    for name, (D, chi, aseed, tseed, steps, evo) in CASES.items():
        A = random_site_tensors(D, 2, aseed)
        config = {'dtype': 'float64', 'device': 'cpu',
                  'TN': {'nx': 2, 'ny': 2, 'dims': {'phys': 2, 'bond': D, 'chi': chi}},
                  'model': {'name': 'heisenberg', 'params': {'J': 1.0}},
                  'ctmrg': {'steps': steps, 'projectors': 'half-system', 'svd_type': 'rsvd',
                            'disable_progressbar': True},
                  'evolution': {'backend': 'torch', 'disable_progressbar': True}}
        ipeps = Ipeps(config)
        for s in SITES:
            ipeps[s]['A'] = torch.as_tensor(A[s])
            ipeps[s].initialize_corner_tensors()
            ipeps[s].initialize_edge_tensors()
        ipeps.site_states_initialized = True
        for s in SITES:
            for k in range(4):
                out[f"{name}/init/{s[0]}{s[1]}_C{k}"] = ipeps[s]['C'][k].numpy().copy()
                out[f"{name}/init/{s[0]}{s[1]}_E{k}"] = ipeps[s]['E'][k].numpy().copy()
            out[f"{name}/init/{s[0]}{s[1]}_A"] = A[s]
        torch.manual_seed(tseed)
        if evo is None:
            ipeps.renormalize()
        else:
            ipeps.evolve(*evo)
        e = float(ipeps.measure()['Energy'])
        for s in SITES:
            out[f"{name}/final/{s[0]}{s[1]}_A"] = ipeps[s]['A'].numpy().copy()
            for k in range(4):
                out[f"{name}/final/{s[0]}{s[1]}_C{k}"] = ipeps[s]['C'][k].numpy().copy()
                out[f"{name}/final/{s[0]}{s[1]}_E{k}"] = ipeps[s]['E'][k].numpy().copy()
        out[f"{name}/energy"] = np.array(e)
        out[f"{name}/meta"] = np.array([D, chi, aseed, tseed, steps] +
                                       ([evo[0], evo[1]] if evo else [0.0, 0]), dtype=np.float64)
        print(name, e, flush=True)
    np.savez_compressed(os.path.join(HERE, "reference_trajectory.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
