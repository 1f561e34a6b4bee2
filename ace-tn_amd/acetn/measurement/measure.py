"""Energy and observables (acetn/measurement/measure.py:5-198).  The RDMs
come from the HIP engine; the averaging over the unit cell and the traces
with d x d operators are host-side bookkeeping, as in the reference."""
import torch
from tqdm import tqdm

from .rdm import RDM


def _mat(op):
    op = op.mat if hasattr(op, 'mat') else op
    return torch.as_tensor(op, dtype=torch.float64).cpu()


def measure_one_site(rdm, operator):
    """measure.py:173-184."""
    return torch.einsum("Pp,pP->", rdm, operator).real


def measure_two_site(rdm, operator):
    """measure.py:187-198."""
    return torch.einsum("PQpq,pqPQ->", rdm, operator).real


def measure(ipeps, model):
    """measure.py:5-28.  Returns a dict of 0-d tensors (Energy and the
    model's observables averaged over the unit cell)."""
    out = {'Energy': torch.zeros((), dtype=torch.float64)}
    rdm = RDM(ipeps)
    one_keys = list(model.one_site_observables((0, 0)).keys())
    if one_keys:
        tqdm.write("Calculated one-site measurements:")
        tqdm.write("site\t" + "\t".join(k[:10].ljust(10) for k in one_keys))
    for key in one_keys:
        out[key] = torch.zeros((), dtype=torch.float64)
    for site in ipeps.site_list:
        rho = rdm[site].cpu()
        nrm = torch.einsum("pp->", rho).real
        h = model.one_site_hamiltonian(site)
        if h is not None:
            out['Energy'] += measure_one_site(rho, _mat(h)) / nrm
        line = f"{site}"
        for key, op in model.one_site_observables(site).items():
            v = measure_one_site(rho, _mat(op)) / nrm
            out[key] += v
            line += f"\t{v.item():.8f}"
        if one_keys:
            tqdm.write(line)
    for key in one_keys:
        out[key] /= len(ipeps.site_list)
    two_keys = list(model.two_site_observables(ipeps.bond_list[0]).keys())
    for key in two_keys:
        out[key] = torch.zeros((), dtype=torch.float64)
    d = model.dim
    for bond in ipeps.bond_list:
        rho = rdm[bond].cpu()
        nrm = torch.einsum("pqpq->", rho).real
        h = _mat(model.two_site_hamiltonian(bond)).reshape(d, d, d, d)
        out['Energy'] += measure_two_site(rho, h) / nrm
        for key, op in model.two_site_observables(bond).items():
            out[key] += measure_two_site(rho, _mat(op).reshape(d, d, d, d)) / nrm
    for key in two_keys:
        out[key] /= len(ipeps.bond_list)
    out['Energy'] /= len(ipeps.site_list)
    tqdm.write("Measurement averages:")
    tqdm.write("\n".join(f"{k}: {v}" for k, v in out.items()))
    return out
