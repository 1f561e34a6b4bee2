"""Imaginary-time evolution driver (acetn/evolution/evolve.py:10-59)."""
import time

from tqdm import tqdm

from .fast_full_update import FastFullUpdater
from .gate import Gate
from .simple_update import SimpleUpdater


def create_tensor_updater(ipeps, gate, config, proposer=None):
    """evolve.py:10-18."""
    if config.update_type == "full":
        return FastFullUpdater(ipeps, gate, config, proposer)
    if config.update_type == "simple":
        return SimpleUpdater(ipeps, gate, config)
    raise ValueError(f"Unknown update_type: {config.update_type}")


def load_proposer(ipeps, config):
    """The learned ALS start of ``config.als_start='learned'``: the network
    saved at ``config.als_start_model`` (MLP.save), checked against the
    bond problem's feature sizes."""
    from .learned_start import MLP, MLPProposer, feature_sizes
    path = getattr(config, "als_start_model", None)
    if not path:
        raise ValueError("evolution.als_start='learned' needs evolution.als_start_model: the "
                         "path of a network saved by MLP.save (python -m "
                         "acetn.evolution.start_workload --save PATH trains one)")
    mlp = MLP.load(path, device=ipeps.device)
    D, d = ipeps.dims['bond'], ipeps.dims['phys']
    fin, fout = feature_sizes(D * d, d)
    if (mlp.fin, mlp.fout) != (fin, fout):
        raise ValueError(f"{path}: the network was trained for other bond sizes "
                         f"(features {mlp.fin}, outputs {mlp.fout}; this network needs "
                         f"{fin}, {fout})")
    return MLPProposer(mlp, config)


def evolve(ipeps, dtau, steps, model, config, proposer=None):
    """evolve.py:21-59: ``steps`` sweeps over the bond list, reversed after
    every sweep.  Returns the tensor updater (its statistics stay readable)."""
    gate = Gate(model, dtau, ipeps.bond_list, ipeps.site_list, device=ipeps.device)
    if proposer is None and getattr(config, "als_start", "plain") == "learned":
        proposer = load_proposer(ipeps, config)
    tensor_updater = create_tensor_updater(ipeps, gate, config, proposer)
    loop_start = time.time()
    for _ in tqdm(range(1, steps + 1), desc="Updating tensors",
                  disable=config.disable_progressbar):
        for bond in ipeps.bond_list:
            tensor_updater.bond_update(bond)
        ipeps.bond_list.reverse()
    tensor_updater.finalize()
    tqdm.write("Finished imaginary-time evolution")
    tqdm.write(f"Total runtime: {(time.time() - loop_start):.6f} seconds")
    return tensor_updater
