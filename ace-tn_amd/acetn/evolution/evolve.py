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


def evolve(ipeps, dtau, steps, model, config, proposer=None):
    """evolve.py:21-59: ``steps`` sweeps over the bond list, reversed after
    every sweep.  Returns the tensor updater (its statistics stay readable)."""
    gate = Gate(model, dtau, ipeps.bond_list, ipeps.site_list, device=ipeps.device)
    if proposer is None and getattr(config, "als_start", "plain") == "learned":
        raise NotImplementedError("evolution.als_start='learned' needs a trained proposer: "
                                  "pass one to evolve(..., proposer=...)")
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
