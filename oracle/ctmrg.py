"""CTMRG restatement (oracle side, FP64 numpy).

State layout (same as the reference's SiteTensor, acetn/ipeps/site_tensor.py):
    state[(x, y)] = {'A': (D,D,D,D,d) indices (l,u,r,d,p),
                     'C': [C0..C3] each (chi, chi),
                     'E': [E0..E3] each (chi, chi, D_ket, D_bra)}
Direction k: 0=left, 1=up, 2=right, 3=down (acetn/ipeps/bond.py:11).

Projector gauge: the reference leaves the sign of each singular pair to
LAPACK.  Both this oracle and the device engine fix it the same way (the
weighted column sum sum_i w_i P1[i, z] with w_i = 1 + frac(i*phi) is made
positive, P2 follows); this is a pure boundary gauge choice and leaves every
observable unchanged.
"""
import numpy as np

from . import linalg
from .rng import SketchStream

ein = lambda s, *ops: np.einsum(s, *ops, optimize=True)


# --------------------------------------------------------------------------
# lattice bookkeeping: acetn/ipeps/tensor_network.py:153-193
# --------------------------------------------------------------------------
def site_list(nx, ny):
    return [(x, y) for x in range(nx) for y in range(ny)]


def bond_list(nx, ny):
    """Horizontal bonds (k=2) for every (x,y), then vertical (k=1)."""
    bonds = []
    for x in range(nx):
        for y in range(ny):
            bonds.append(((x, y), ((x + 1) % nx, y), 2))
    for y in range(ny):
        for x in range(nx):
            bonds.append(((x, y), (x, (y + 1) % ny), 1))
    return bonds


def bond_permute(a, k):
    """acetn/ipeps/site_tensor.py:138-150."""
    return np.transpose(a, [(i + k) % 4 for i in range(4)] + [4])


def init_corners_edges(a):
    """Initial C/E from A: site_tensor.py:152-202."""
    D = a.shape[0]
    C, E = [], []
    for k in range(4):
        ak = bond_permute(a, k)
        ck = ein("lurdp,luRDp->dDrR", ak, ak.conj()).reshape(D * D, D * D)
        C.append(ck / np.linalg.norm(ck))
        ek = ein("lurdp,LuRDp->lLrRdD", ak, ak.conj()).reshape(D * D, D * D, D, D)
        E.append(ek / np.linalg.norm(ek))
    return C, E


def make_state(site_tensors):
    """site_tensors: dict site -> A.  C/E from init_corners_edges."""
    st = {}
    for s, a in site_tensors.items():
        C, E = init_corners_edges(a)
        st[s] = {'A': a.copy(), 'C': C, 'E': E}
    return st


def random_site_tensors(nx, ny, D, d, seed, bias=0.0):
    """Seeded iid N(0,1) site tensors (plus an optional bias on the
    [0,0,0,0,0] spin-up slot, the slot site_tensor.py:134-135 fills), each
    normalised to unit Frobenius norm, drawn in site_list order."""
    rng = np.random.default_rng(seed)
    out = {}
    for s in site_list(nx, ny):
        a = rng.standard_normal((D, D, D, D, d))
        a[0, 0, 0, 0, 0] += bias
        out[s] = a / np.linalg.norm(a)
    return out


# --------------------------------------------------------------------------
# quarter tensors and projectors: acetn/renormalization/projectors.py
# --------------------------------------------------------------------------
def make_quarter_tensor(site, k):
    """projectors.py:35-60.  Returns Q as an (N x N) matrix and the 6-index
    shape (c, r, R, e, d, D).  The max-abs scaling of :59 is kept."""
    ak = bond_permute(site['A'], k)
    ck = site['C'][k % 4]
    ek1 = site['E'][(3 + k) % 4]
    ek2 = site['E'][k % 4]
    qk = ein("ab,bcuU->acuU", ck, ek2)
    qk = ein("acuU,ealL->cuUelL", qk, ek1)
    qk = ein("cuUelL,LURDP->cuelRDP", qk, ak.conj())
    qk = ein("lurdp,cuelRDp->crRedD", ak, qk)
    qD = qk.shape
    qk = qk.reshape(qD[0] * qD[1] * qD[2], qD[3] * qD[4] * qD[5])
    qk = qk / np.abs(qk).max()
    return qk, qD


def fused_matmul_svd_lowrank(A, B, q, niter, omega_fn):
    """acetn/linalg/fused_matmul_svd_lowrank.py:4-52 with the sketch supplied
    by ``omega_fn(n, q)``."""
    m, n = A.shape[0], B.shape[1]
    q = min(q, m, n)
    Om = omega_fn(n, q)
    Y = A @ (B @ Om)
    for _ in range(niter):
        Y, _ = linalg.qr(Y)
        Y = A @ (B @ (B.conj().T @ (A.conj().T @ Y)))
    Qm, _ = linalg.qr(Y)
    Bp = (Qm.conj().T @ A) @ B
    Ub, S, Vh = linalg.svd(Bp, full_matrices=False)
    return Qm @ Ub, S, Vh.conj().T


def fused_3matmul_svd_lowrank(A, B, C, D, q, niter, omega_fn):
    """acetn/linalg/fused_3matmul_svd_lowrank.py:4-56."""
    m, n = A.shape[0], D.shape[1]
    q = min(q, m, n)
    Om = omega_fn(n, q)
    H = lambda X: X.conj().T
    Y = A @ (B @ (C @ (D @ Om)))
    for _ in range(niter):
        Y, _ = linalg.qr(Y)
        Y = H(D) @ (H(C) @ (H(B) @ (H(A) @ Y)))
        Y, _ = linalg.qr(Y)
        Y = A @ (B @ (C @ (D @ Y)))
    Qm, _ = linalg.qr(Y)
    Bp = (((H(Qm) @ A) @ B) @ C) @ D
    Ub, S, Vh = linalg.svd(Bp, full_matrices=False)
    return Qm @ Ub, S, H(Vh)


def svd_lowrank(A, q, niter, omega_fn):
    """acetn/linalg/svd_lowrank.py:4-44."""
    m, n = A.shape
    q = min(q, m, n)
    Om = omega_fn(n, q)
    Y = A @ Om
    for _ in range(niter):
        Y, _ = linalg.qr(Y)
        Y = A @ (A.conj().T @ Y)
    Qm, _ = linalg.qr(Y)
    Ub, S, Vh = linalg.svd(Qm.conj().T @ A, full_matrices=False)
    return Qm @ Ub, S, Vh.conj().T


def gauge_weights(n):
    """w_i = 1 + frac(i * 0.6180339887498949): fixed, positive and asymmetric
    under any index swap, so the weighted sum below has no symmetry-forced
    zeros (the ket/bra symmetry of the environment makes max-|entry| rules
    tie exactly)."""
    return 1.0 + np.mod(np.arange(n, dtype=np.float64) * 0.6180339887498949, 1.0)


def fix_projector_gauge(P1, P2):
    """Sign convention shared with the device engine: for every retained
    singular index z, sign(sum_i w_i P1[i, z]) is made positive and P2[:, z]
    follows."""
    if not linalg.state.gauges:
        return P1, P2
    n = P1.shape[-1]
    f1 = P1.reshape(-1, n)
    s = gauge_weights(f1.shape[0]) @ f1
    sgn = np.where(s < 0, -1.0, 1.0)
    return P1 * sgn, P2 * sgn


class ProjectorCalculator:
    """acetn/renormalization/projectors.py:6-236."""

    def __init__(self, projectors="half-system", svd_type="rsvd", svd_cutoff=1e-12,
                 rsvd_niter=2, rsvd_oversampling=2, omega_fn=None):
        if projectors not in ("half-system", "full-system", None):
            raise ValueError(f"Invalid ctmrg projector type: {projectors} provided.")
        self.projectors = projectors or "full-system"
        self.svd_type = svd_type
        self.svd_cutoff = svd_cutoff
        self.rsvd_niter = rsvd_niter
        self.rsvd_oversampling = rsvd_oversampling
        self.omega_fn = omega_fn if omega_fn is not None else SketchStream(0)
        self.last_spectrum = None

    def _truncate(self, U, s, V, chi):
        """projectors.py:163-167 (and :128-132)."""
        s = s / s[0]
        cD_new = int(min(chi, int((s > self.svd_cutoff).sum())))
        self.last_spectrum = s.copy()
        self.last_cD = cD_new
        U = U[:, :cD_new] * (1.0 / np.sqrt(s[:cD_new]))
        V = V[:, :cD_new] * (1.0 / np.sqrt(s[:cD_new]))
        return U, V, cD_new

    def calculate(self, st, sites, k, chi):
        if any(s not in st for s in sites):
            raise ValueError(f"Site tensor not defined at sites {sites}.")
        if self.projectors == "half-system":
            return self.half_system(st, sites, k, chi)
        return self.full_system(st, sites, k, chi)

    def half_system(self, st, sites, k, chi):
        """projectors.py:138-174 (rsvd) and :62-82 + :114-136 (full-rank)."""
        s1, s4 = sites[0], sites[3]
        Q1, q1D = make_quarter_tensor(st[s1], k)
        Q4, q4D = make_quarter_tensor(st[s4], k + 3)
        if self.svd_type == "full-rank":
            R = Q1 @ Q4
            R = R / np.abs(R).max()
            Uf, s, Vh = linalg.svd(R)
            U, V = Uf, Vh.conj().T
        else:
            U, s, V = fused_matmul_svd_lowrank(Q1, Q4, chi + self.rsvd_oversampling,
                                               self.rsvd_niter, self.omega_fn)
        U, V, cD_new = self._truncate(U, s, V, chi)
        Q1v = Q1.reshape(Q1.shape[0], *q1D[3:])
        Q4v = Q4.reshape(*q4D[:3], Q4.shape[1])
        P1 = ein("xedD,xz->edDz", Q1v, U.conj())
        P2 = ein("cuUy,yz->cuUz", Q4v, V)
        return fix_projector_gauge(P1, P2)

    def full_system(self, st, sites, k, chi):
        """projectors.py:176-217 (rsvd) and :84-112 (full-rank)."""
        s1, s2, s3, s4 = sites
        Q1, q1D = make_quarter_tensor(st[s1], k)
        Q2, _ = make_quarter_tensor(st[s2], k + 1)
        Q3, _ = make_quarter_tensor(st[s3], k + 2)
        Q4, q4D = make_quarter_tensor(st[s4], k + 3)
        if self.svd_type == "full-rank":
            R1 = Q2 @ Q1
            R2 = Q4 @ Q3
            R1 = R1 / np.abs(R1).max()
            R2 = R2 / np.abs(R2).max()
            F = R1 @ R2
            F = F / np.abs(F).max()
            Uf, s, Vh = linalg.svd(F)
            U, V, cD_new = self._truncate(Uf, s, Vh.conj().T, chi)
            P1 = ein("xedD,xz->edDz", R1.reshape(R1.shape[0], *q1D[3:]), U.conj())
            P2 = ein("cuUy,yz->cuUz", R2.reshape(*q4D[:3], R2.shape[1]), V)
            return fix_projector_gauge(P1, P2)
        U, s, V = fused_3matmul_svd_lowrank(Q2, Q1, Q4, Q3, chi + self.rsvd_oversampling,
                                            self.rsvd_niter, self.omega_fn)
        U, V, cD_new = self._truncate(U, s, V, chi)
        P1 = (Q1.conj().T @ (Q2.conj().T @ U.conj())).reshape(*q1D[3:], cD_new)
        P2 = (Q4 @ (Q3 @ V)).reshape(*q4D[:3], cD_new)
        return fix_projector_gauge(P1, P2)


# --------------------------------------------------------------------------
# renormalisation: acetn/renormalization/directional_mover.py:273-366
# --------------------------------------------------------------------------
def renormalize_cj1(ci, ei, proj):
    cj = ein("ablL,bc->alLc", ei, ci)
    cj = ein("alLc,clLx->ax", cj, proj)
    return cj / np.linalg.norm(cj)


def renormalize_cj2(ci, ei, proj):
    cj = ein("ab,bcrR->acrR", ci, ei)
    cj = ein("arRx,acrR->xc", proj, cj)
    return cj / np.linalg.norm(cj)


def renormalize_ej(ei, ai, proj2, proj1):
    ej = ein("ablL,buUx->alLuUx", ei, proj1)
    ej = ein("LURDP,alLuUx->RDPalux", ai.conj(), ej)
    ej = ein("lurdp,RDpalux->rdRDax", ai, ej)
    ej = ein("rdRDax,adDy->yxrR", ej, proj2)
    return ej / np.linalg.norm(ej)


def renormalize_boundary(st, proj1, proj2, s1, s2, i, j, k):
    """directional_mover.py:273-303."""
    site1 = st[s1]
    new_c1 = renormalize_cj1(site1['C'][(3 + k) % 4], site1['E'][(2 + k) % 4], proj1[i])
    new_c2 = renormalize_cj2(site1['C'][k], site1['E'][k], proj2[j])
    new_e = renormalize_ej(site1['E'][(3 + k) % 4], bond_permute(site1['A'], k),
                           proj2[i], proj1[j])
    st[s2]['C'][(3 + k) % 4] = new_c1
    st[s2]['C'][k] = new_c2
    st[s2]['E'][(3 + k) % 4] = new_e


def plaquette(nx, ny, xi, yi, k):
    """Site quadruples of calculate_{left,up,right,down}_projectors,
    directional_mover.py:99-181."""
    if k == 0:
        xj, yj = (xi + 1) % nx, (yi - 1 + ny) % ny
        return [(xi, yi), (xj, yi), (xj, yj), (xi, yj)]
    if k == 2:
        xj, yj = (xi - 1 + nx) % nx, (yi + 1) % ny
        return [(xi, yi), (xj, yi), (xj, yj), (xi, yj)]
    if k == 1:
        xj, yj = (xi - 1 + nx) % nx, (yi - 1 + ny) % ny
        return [(xi, yi), (xi, yj), (xj, yj), (xj, yi)]
    xj, yj = (xi + 1) % nx, (yi + 1) % ny
    return [(xi, yi), (xi, yj), (xj, yj), (xj, yi)]


def move_targets(nx, ny, line, k, idx):
    """(s1, s2, i, j) of the renormalisation loop of each move,
    directional_mover.py:23-97.  ``line`` is the column (k=0,2) or the row
    (k=1,3); ``idx`` runs over the other coordinate."""
    if k == 0:
        return (line, idx), ((line + 1) % nx, idx), idx, (idx + 1) % ny
    if k == 2:
        return (line, idx), ((line - 1 + nx) % nx, idx), idx, (idx - 1 + ny) % ny
    if k == 1:
        return (idx, line), (idx, (line - 1 + ny) % ny), idx, (idx + 1) % nx
    return (idx, line), (idx, (line + 1) % ny), idx, (idx - 1 + nx) % nx


class DirectionalMover:
    """directional_mover.py:5-97."""

    def __init__(self, calc, nx, ny, chi):
        self.calc, self.nx, self.ny, self.chi = calc, nx, ny, chi

    def move(self, st, line, k):
        n_other = self.ny if k in (0, 2) else self.nx
        proj1, proj2 = {}, {}
        for idx in range(n_other):
            xi, yi = (line, idx) if k in (0, 2) else (idx, line)
            sites = plaquette(self.nx, self.ny, xi, yi, k)
            proj1[idx], proj2[idx] = self.calc.calculate(st, sites, k, self.chi)
        for idx in range(n_other):
            s1, s2, i, j = move_targets(self.nx, self.ny, line, k, idx)
            renormalize_boundary(st, proj1, proj2, s1, s2, i, j, k)

    def left_move(self, st, xi):
        self.move(st, xi, 0)

    def right_move(self, st, xi):
        self.move(st, xi, 2)

    def up_move(self, st, yi):
        self.move(st, yi, 1)

    def down_move(self, st, yi):
        self.move(st, yi, 3)


def ctmrg_iteration(mover, st):
    """One pass of the loop body of acetn/renormalization/ctmrg.py:26-31."""
    nx, ny = mover.nx, mover.ny
    for xi in range(nx):
        mover.left_move(st, xi)
        mover.right_move(st, (nx - xi + 1) % nx)
    for yi in range(ny):
        mover.up_move(st, (ny - yi + 1) % ny)
        mover.down_move(st, yi)


def ctmrg(st, nx, ny, chi, steps, calc):
    """acetn/renormalization/ctmrg.py:4-31 (non-distributed branch)."""
    mover = DirectionalMover(calc, nx, ny, chi)
    for _ in range(steps):
        ctmrg_iteration(mover, st)
    return st


def corner_spectra(st):
    """Normalised singular values of every corner; used by the
    convergence rule (max change < tol) that drives CTMRG 'to convergence'
    (the reference has none: it runs a fixed number of steps)."""
    out = []
    for s in sorted(st):
        for c in st[s]['C']:
            sv = np.linalg.svd(c, compute_uv=False)
            out.append(sv / sv[0])
    return out


def spectra_change(a, b):
    m = 0.0
    for x, y in zip(a, b):
        n = max(len(x), len(y))
        xx = np.zeros(n); yy = np.zeros(n)
        xx[:len(x)] = x; yy[:len(y)] = y
        m = max(m, float(np.abs(xx - yy).max()))
    return m
