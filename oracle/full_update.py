"""Full update, ALS bond solver and simple update (oracle side, FP64).

Restates acetn/evolution/{tensor_update,full_update,als_solver,
fast_full_update,evolve,simple_update,lambda_tensors}.py.

Gauge convention shared with the device engine: after
``finalize_reduced_tensors`` the sign of every retained bond index u is chosen
so that the weighted sum of a1r[:, u, :] (weights oracle.ctmrg.gauge_weights
over the flattened (y, p) entries) is positive; a2r follows.  The
reference leaves this sign to LAPACK; it is a bond gauge and leaves every
observable unchanged.
"""
import numpy as np
import scipy.linalg

from . import linalg

from .ctmrg import ein, bond_permute, DirectionalMover
from .measure import bond_env


def bond_permutation(k):
    return [(i + k) % 4 for i in range(4)] + [4]


def decompose_site_tensors(a1, a2):
    """tensor_update.py:52-68."""
    bD, pD = a1.shape[3:]
    nD = min(bD ** 3, pD * bD)
    q1, r1 = linalg.qr(ein("lurdp->rdulp", a1).reshape(bD ** 3, pD * bD))
    q2, r2 = linalg.qr(ein("lurdp->dlurp", a2).reshape(bD ** 3, pD * bD))
    return (q1.reshape(bD, bD, bD, nD), r1.reshape(nD, bD, pD),
            q2.reshape(bD, bD, bD, nD), r2.reshape(nD, bD, pD))


def recompose_site_tensors(a1q, a1r, a2q, a2r):
    """tensor_update.py:70-75."""
    return ein('rdux,xlp->lurdp', a1q, a1r), ein('dlux,xrp->lurdp', a2q, a2r)


def norm_tensor_core(env, a1q, a2q):
    """full_update.py:209-227."""
    t = ein("ab,bcrR->acrR", env['c12'], env['e12'])
    t = ein("acrR,eauU->crReuU", t, env['e11'])
    t = ein("crReuU,RDUY->creuDY", t, a1q.conj())
    t = ein("creuDY,rduy->ceDYdy", t, a1q)
    n1 = ein("ab,bfdD->afdD", env['c13'], env['e13'])
    n1 = ein("afdD,aeDYdy->feYy", n1, t)
    t = ein("ab,bcuU->acuU", env['c21'], env['e21'])
    t = ein("acuU,ealL->cuUelL", t, env['e24'])
    t = ein("cuUelL,DLUX->cuelXD", t, a2q.conj())
    t = ein("cuelXD,dlux->ceXDxd", t, a2q)
    n2 = ein("ab,fadD->bfdD", env['c24'], env['e23'])
    n2 = ein("bfdD,cbXDxd->fcXx", n2, t)
    return ein("fcYy,fcXx->yxYX", n1, n2)


def build_norm_tensor(st, bond, a1q, a2q):
    """full_update.py:163-206."""
    return norm_tensor_core(bond_env(st, bond), a1q, a2q)


def positive_approx(n12, cutoff=1e-12):
    """full_update.py:262-297 (LAPACK eigh; the SVD fallback of :282-287 is
    kept for a LinAlgError)."""
    nD = n12.shape[0]
    N = n12.reshape(nD ** 2, nD ** 2).copy()

    def _eig(M):
        try:
            return linalg.eigh(M)
        except np.linalg.LinAlgError:
            u, w, _ = linalg.svd(M)
            return w[::-1].copy(), u[:, ::-1].copy()

    nw, nz = _eig(N)
    while nw[0] < cutoff:
        N += 2 * max(cutoff, abs(nw[0])) * np.eye(nD ** 2)
        nw, nz = _eig(N)
    return nz.reshape(nD, nD, nD ** 2) * np.sqrt(nw)


def pinv_atol(a, atol):
    """torch.linalg.pinv(a, atol=atol) with rtol=0 (singular values <= atol
    are dropped)."""
    u, s, vh = linalg.svd(a)
    keep = s > atol
    sinv = np.zeros_like(s)
    sinv[keep] = 1.0 / s[keep]
    return (vh.conj().T * sinv) @ u.conj().T


def gauge_fix(nz, a12g, atol=1e-12):
    """full_update.py:300-343."""
    nD = a12g.shape[0]
    _, nzyr = linalg.qr(ein("yxz->zxy", nz).reshape(nD ** 3, nD))
    _, nzxr = linalg.qr(ein("yxz->zyx", nz).reshape(nD ** 3, nD))
    nzyr_inv = pinv_atol(nzyr, atol)
    nzxr_inv = pinv_atol(nzxr, atol)
    nz = ein("yxz,xw->yzw", nz, nzxr_inv)
    nz = ein("yzw,yv->zvw", nz, nzyr_inv)
    n12 = ein("zvw,zVW->vwVW", nz, nz.conj())
    a12g = ein("zx,yxpq->yzpq", nzxr, a12g)
    a12g = ein("wy,yzpq->wzpq", nzyr, a12g)
    return n12, a12g, nzxr_inv, nzyr_inv


# --------------------------------------------------------------------------
# ALS: acetn/evolution/als_solver.py
# --------------------------------------------------------------------------
def initialize_reduced_tensors(a12g, ar_shape):
    """als_solver.py:117-146 (the 'plain start')."""
    nD, bD, pD = ar_shape
    m = ein("yxpq->ypxq", a12g).reshape(nD * pD, nD * pD)
    U, S, Vh = linalg.svd(m)
    V = Vh.conj().T
    S = np.sqrt(S[:bD] / S[0])
    U = U[:, :bD].reshape(nD, pD, bD)
    V = V[:, :bD].reshape(nD, pD, bD)
    return ein("ypu,u->yup", U, S), ein("xqv,v->xvq", V, S)


def build_s1(n12g, a2r):
    return ein('YXpQ,XUQ->YUp', n12g, a2r.conj())


def build_s2(n12g, a1r):
    return ein('YXPq,YVP->XVq', n12g, a1r.conj())


def build_r1(n12, a2r):
    R = ein('yxYX,xuq->yYXuq', n12, a2r)
    return ein('yYXuQ,XUQ->YUyu', R, a2r.conj())


def build_r2(n12, a1r):
    R = ein('yxYX,yvp->xYXvp', n12, a1r)
    return ein('xYXvP,YVP->XVxv', R, a1r.conj())


def calculate_cost(a1r, a2r, a12g, n12):
    """als_solver.py:250-260."""
    a12n = ein("yup,xuq->yxpq", a1r, a2r)
    d2 = ein("YXpq,YXpq->", ein("yxYX,yxpq->YXpq", n12, a12n), a12n.conj())
    d3 = ein("YXpq,YXpq->", ein("yxYX,yxpq->YXpq", n12, a12g), a12n.conj())
    return float(d2.real - 2 * d3.real)


def solve_ar(R, S, method="cholesky", epsilon=1e-12):
    """als_solver.py:194-230."""
    nD, bD, pD = S.shape
    S = S.reshape(nD * bD, pD)
    R = R.reshape(nD * bD, nD * bD)
    R = 0.5 * (R + R.conj().T)
    if method == "cholesky":
        R = R + epsilon * np.abs(R).max() * np.eye(R.shape[0])
        try:
            L = np.linalg.cholesky(R)
            Yv = scipy.linalg.solve_triangular(L, S, lower=True)
            ar = scipy.linalg.solve_triangular(L.conj().T, Yv, lower=False)
        except np.linalg.LinAlgError:
            ar = np.linalg.solve(R, S)
    elif method == "pinv":
        u, s, vh = linalg.svd(R)
        keep = s > epsilon * s[0]
        sinv = np.where(keep, 1.0 / np.where(keep, s, 1.0), 0.0)
        ar = (vh.conj().T * sinv) @ (u.conj().T @ S)
    else:
        raise ValueError(method)
    return ar.reshape(nD, bD, pD)


def als_solve(n12, a12g, ar_shape, niter=100, tol=1e-15, method="cholesky",
              epsilon=1e-12, start=None):
    """ALSSolver.solve + solve_torch, als_solver.py:34-82.  Returns
    (a1r, a2r, sweeps, final_cost).  ``start`` optionally replaces the plain
    SVD start with (a1r, a2r)."""
    a1r, a2r = initialize_reduced_tensors(a12g, ar_shape) if start is None else start
    n12g = ein("yxYX,yxpq->YXpq", n12, a12g)
    d1 = abs(calculate_cost(a1r, a2r, a12g, n12))
    sweeps = 0
    d2 = d1
    for i in range(niter):
        a1r = solve_ar(build_r1(n12, a2r), build_s1(n12g, a2r), method, epsilon)
        a2r = solve_ar(build_r2(n12, a1r), build_s2(n12g, a1r), method, epsilon)
        sweeps = i + 1
        d2 = calculate_cost(a1r, a2r, a12g, n12)
        error = abs(d2 - d1) / abs(d1)
        if error < tol and i > 1:
            break
        d1 = d2
    return a1r, a2r, sweeps, d2


def fix_bond_gauge(a1r, a2r):
    """Sign convention shared with the device engine (module docstring):
    sign(sum_i w_i a1r[i, u]) > 0 over the flattened (y, p) entries of each
    bond index u, with the weights of oracle.ctmrg.gauge_weights."""
    from .ctmrg import gauge_weights
    if not linalg.state.gauges:
        return a1r, a2r
    nD, bD, pD = a1r.shape
    f = np.transpose(a1r, (1, 0, 2)).reshape(bD, -1)
    s = f @ gauge_weights(f.shape[1])
    sgn = np.where(s < 0, -1.0, 1.0)
    return a1r * sgn[None, :, None], a2r * sgn[None, :, None]


def finalize_reduced_tensors(a1r, a2r, nzyr_inv=None, nzxr_inv=None):
    """full_update.py:123-161."""
    if nzyr_inv is not None:
        a1r = ein("yz,zup->yup", nzyr_inv, a1r)
        a2r = ein("xw,wvq->xvq", nzxr_inv, a2r)
    nD, bD, pD = a1r.shape
    q1, r1 = linalg.qr(ein("yup->ypu", a1r).reshape(nD * pD, bD))
    q2, r2 = linalg.qr(ein("xvq->xqv", a2r).reshape(nD * pD, bD))
    U, s, Vh = linalg.svd(ein("au,bu->ab", r1, r2))
    s = np.sqrt(s[:bD] / np.linalg.norm(s))
    r1 = ein('ab,b->ab', U[:, :bD], s)
    r2 = ein('ba,b->ba', Vh[:bD, :], s)
    a1r = ein("ypa,au->yup", q1.reshape(nD, pD, bD), r1)
    a2r = ein("xqb,vb->xvq", q2.reshape(nD, pD, bD), r2)
    return fix_bond_gauge(a1r, a2r)


class FullUpdateConfig:
    def __init__(self, use_gauge_fix=True, gauge_fix_atol=1e-12, positive_approx_cutoff=1e-12,
                 als_niter=100, als_tol=1e-15, als_method="cholesky", als_epsilon=1e-12):
        self.use_gauge_fix = use_gauge_fix
        self.gauge_fix_atol = gauge_fix_atol
        self.positive_approx_cutoff = positive_approx_cutoff
        self.als_niter = als_niter
        self.als_tol = als_tol
        self.als_method = als_method
        self.als_epsilon = als_epsilon


def tensor_update(st, bond, a1, a2, gate, cfg, record=None):
    """FullUpdater.tensor_update + update_reduced_tensors +
    precondition_norm_tensor (full_update.py:34-121).  a1, a2 are already in
    the bond frame."""
    a1q, a1r, a2q, a2r = decompose_site_tensors(a1, a2)
    n12 = build_norm_tensor(st, bond, a1q, a2q)
    a12g = ein("yup,xuq->yxpq", a1r, a2r)
    a12g = ein("yxpq,pqrs->yxrs", a12g, gate)
    nz = positive_approx(n12, cfg.positive_approx_cutoff)
    if cfg.use_gauge_fix:
        n12, a12g, nzxr_inv, nzyr_inv = gauge_fix(nz, a12g, cfg.gauge_fix_atol)
    else:
        n12 = ein("xyz,XYz->xyXY", nz, nz.conj())
        nzxr_inv = nzyr_inv = None
    b1, b2, sweeps, cost = als_solve(n12, a12g, a1r.shape, cfg.als_niter, cfg.als_tol,
                                     cfg.als_method, cfg.als_epsilon)
    if record is not None:
        record.append(dict(n12=n12, a12g=a12g, a1r=b1, a2r=b2, sweeps=sweeps, cost=cost))
    b1, b2 = finalize_reduced_tensors(b1, b2, nzyr_inv, nzxr_inv)
    n1, n2 = recompose_site_tensors(a1q, b1, a2q, b2)
    return n1 / np.linalg.norm(n1), n2 / np.linalg.norm(n2)


def bond_update_full(st, bond, gate, cfg, mover, record=None):
    """TensorUpdater.update (tensor_update.py:18-38) + FastFullUpdater
    bond_update/absorb_bond (fast_full_update.py:29-101)."""
    s1, s2, k = bond
    a1 = np.transpose(st[s1]['A'], bond_permutation(k))
    a2 = np.transpose(st[s2]['A'], bond_permutation(k))
    a1, a2 = tensor_update(st, bond, a1, a2, gate, cfg, record)
    st[s1]['A'] = np.ascontiguousarray(np.transpose(a1, bond_permutation(4 - k)))
    st[s2]['A'] = np.ascontiguousarray(np.transpose(a2, bond_permutation(4 - k)))
    if k == 0:
        mover.right_move(st, s1[0]); mover.left_move(st, s2[0])
    elif k == 1:
        mover.down_move(st, s1[1]); mover.up_move(st, s2[1])
    elif k == 2:
        mover.left_move(st, s1[0]); mover.right_move(st, s2[0])
    else:
        mover.up_move(st, s1[1]); mover.down_move(st, s2[1])


def evolve_full(st, nx, ny, model, dtau, steps, cfg, mover, bonds, record=None):
    """acetn/evolution/evolve.py:21-59 for update_type='full'.  ``bonds`` is
    the (mutable) bond list; it is reversed after every step as in :44."""
    from .model import bond_gate
    gates = {b: bond_gate(model, dtau, b) for b in bonds}
    for _ in range(steps):
        for b in bonds:
            bond_update_full(st, b, gates[b], cfg, mover, record)
        bonds.reverse()
    return st


# --------------------------------------------------------------------------
# simple update: acetn/evolution/simple_update.py, lambda_tensors.py
# --------------------------------------------------------------------------
class LambdaTensors:
    CUTOFF = 1e-12

    def __init__(self, bonds, bD):
        self.data = {}
        self.site_bonds = {}
        for b in bonds:
            self.data[b] = np.ones(bD)
            self.site_bonds.setdefault(b[0], {})[b[2]] = b
            self.site_bonds.setdefault(b[1], {})[(b[2] + 2) % 4] = b

    def get(self, site, direction):
        return self.data[self.site_bonds[site][direction]]

    def _weights(self, site, exclude_dir, invert):
        w = []
        for d in range(4):
            v = self.get(site, d)
            if d == exclude_dir:
                w.append(np.ones_like(v))
            else:
                w.append(1.0 / np.maximum(v, self.CUTOFF) if invert else v)
        return w

    def absorb(self, a, site, exclude_dir):
        return ein("l,u,r,d,lurdp->lurdp", *self._weights(site, exclude_dir, False), a)

    def strip(self, a, site, exclude_dir):
        return ein("l,u,r,d,lurdp->lurdp", *self._weights(site, exclude_dir, True), a)


def simple_update_bond(st, bond, gate, lam):
    """SimpleUpdater.bond_update, simple_update.py:28-76."""
    s1, s2, k = bond
    a1 = lam.absorb(st[s1]['A'].copy(), s1, k)
    a2 = lam.absorb(st[s2]['A'].copy(), s2, (k + 2) % 4)
    a1 = np.transpose(a1, bond_permutation(k))
    a2 = np.transpose(a2, bond_permutation(k))
    a1q, a1r, a2q, a2r = decompose_site_tensors(a1, a2)
    nD, bD, pD = a1r.shape
    a12 = ein("yup,xuq->ypxq", a1r, a2r)
    a12g = ein("ypxq,pqrs->yrxs", a12, gate)
    U, S, Vh = linalg.svd(a12g.reshape(nD * pD, nD * pD))
    S = S[:bD]
    lt = np.sqrt(S / np.linalg.norm(S))
    U = U[:, :bD].reshape(nD, pD, bD)
    Vh = Vh[:bD, :].reshape(bD, nD, pD)
    a1r = ein("ypa,a->yap", U, lt)
    a2r = ein("axq,a->xaq", Vh, lt)
    a1r, a2r = fix_bond_gauge(a1r, a2r)   # bond-index sign gauge shared with the device
    a1, a2 = recompose_site_tensors(a1q, a1r, a2q, a2r)
    a1 = np.transpose(a1, bond_permutation(4 - k))
    a2 = np.transpose(a2, bond_permutation(4 - k))
    a1 = lam.strip(a1, s1, k)
    a2 = lam.strip(a2, s2, (k + 2) % 4)
    lam.data[bond] = lt
    st[s1]['A'] = np.ascontiguousarray(a1)
    st[s2]['A'] = np.ascontiguousarray(a2)


def evolve_simple(site_tensors, nx, ny, model, dtau, steps, bonds):
    """evolve() with update_type='simple' on bare site tensors."""
    from .model import bond_gate
    st = {s: {'A': a.copy()} for s, a in site_tensors.items()}
    bD = next(iter(site_tensors.values())).shape[0]
    lam = LambdaTensors(bonds, bD)
    gates = {b: bond_gate(model, dtau, b) for b in bonds}
    for _ in range(steps):
        for b in bonds:
            simple_update_bond(st, b, gates[b], lam)
        bonds.reverse()
    return {s: v['A'] for s, v in st.items()}
