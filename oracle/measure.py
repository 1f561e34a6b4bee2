"""Reduced density matrices and energy (oracle side).

Restates acetn/measurement/rdm.py and acetn/measurement/measure.py.
"""
import numpy as np

from .ctmrg import bond_permute, ein


def site_rdm(site):
    """rdm.py:35-67."""
    c1, c2, c3, c4 = site['C']
    e1, e2, e3, e4 = site['E']
    a1 = site['A']
    t1 = ein("ab,bclL->aclL", c4, e4)
    t1 = ein("aclL,eadD->clLedD", t1, e3)
    t1 = ein("clLedD,LURDP->cledURP", t1, a1.conj())
    t2 = ein("ab,bcuU->acuU", c1, e1)
    t3 = ein("ab,carR->bcrR", c3, e2)
    t3 = ein("ec,bcrR->ebrR", c2, t3)
    t3 = ein("ebrR,aeuU->brRauU", t3, t2)
    t3 = ein("erRcuU,cledURP->ruldP", t3, t1)
    return ein("ruldP,lurdp->Pp", t3, a1)


def bond_env(st, bond):
    """The environment pieces gathered by rdm.py:82-96 (and by
    full_update.py:189-203 for the norm tensor)."""
    s1, s2, k = bond
    a, b = st[s1], st[s2]
    return dict(
        c12=a['C'][(k + 1) % 4], e12=a['E'][(k + 1) % 4], e11=a['E'][k % 4],
        c13=a['C'][(k + 2) % 4], e13=a['E'][(k + 2) % 4],
        c21=b['C'][k % 4], e21=b['E'][k % 4], e24=b['E'][(k + 3) % 4],
        c24=b['C'][(k + 3) % 4], e23=b['E'][(k + 2) % 4])


def bond_rdm_core(env, a1, a2):
    """rdm.py:102-120 (unblocked form; the blocked form :123-154 computes the
    same tensor block by block)."""
    t = ein("ab,bcrR->acrR", env['c12'], env['e12'])
    t = ein("acrR,eauU->crReuU", t, env['e11'])
    t = ein("crReuU,LURDP->creuLDP", t, a1.conj())
    t = ein("creuLDP,lurdp->ceLDPldp", t, a1)
    r1 = ein("ab,bfdD->afdD", env['c13'], env['e13'])
    r1 = ein("afdD,acLDPldp->fcLPlp", r1, t)
    t = ein("ab,bcuU->acuU", env['c21'], env['e21'])
    t = ein("acuU,ealL->cuUelL", t, env['e24'])
    t = ein("cuUelL,LURDQ->cuelRDQ", t, a2.conj())
    t = ein("cuelRDQ,lurdq->ceRDQrdq", t, a2)
    r2 = ein("ae,fadD->efdD", env['c24'], env['e23'])
    r2 = ein("efdD,ceRDQrdq->fcRQrq", r2, t)
    return ein("fcRPrp,fcRQrq->PQpq", r1, r2)


def bond_rdm(st, bond):
    """rdm.py:69-99."""
    s1, s2, k = bond
    return bond_rdm_core(bond_env(st, bond), bond_permute(st[s1]['A'], k),
                         bond_permute(st[s2]['A'], k))


def measure(st, nx, ny, model, bonds=None, sites=None):
    """measure.py:5-28 + helpers :30-198.  Returns a dict of floats."""
    from .ctmrg import bond_list, site_list
    sites = site_list(nx, ny) if sites is None else sites
    bonds = bond_list(nx, ny) if bonds is None else bonds
    out = {'Energy': 0.0}
    obs_keys = list(model.one_site_observables((0, 0)).keys())
    for key in obs_keys:
        out[key] = 0.0
    for s in sites:
        rho = site_rdm(st[s])
        nrm = np.trace(rho).real
        h = model.one_site_hamiltonian(s)
        if h is not None:
            out['Energy'] += ein("Pp,pP->", rho, h).real / nrm
        for key, op in model.one_site_observables(s).items():
            out[key] += ein("Pp,pP->", rho, op).real / nrm
    for key in obs_keys:
        out[key] /= len(sites)
    d = model.dim
    for b in bonds:
        rho = bond_rdm(st, b)
        nrm = ein("pqpq->", rho).real
        h = model.two_site_hamiltonian(b).reshape(d, d, d, d)
        out['Energy'] += ein("PQpq,pqPQ->", rho, h).real / nrm
    out['Energy'] /= len(sites)
    return out


def bond_energies(st, nx, ny, model):
    """Per-bond <h_b> (the nearest-neighbour correlators the energy sums)."""
    from .ctmrg import bond_list
    d = model.dim
    res = []
    for b in bond_list(nx, ny):
        rho = bond_rdm(st, b)
        nrm = ein("pqpq->", rho).real
        res.append(ein("PQpq,pqPQ->", rho, model.two_site_hamiltonian(b).reshape(d, d, d, d)).real / nrm)
    return res
