"""Pins for the CPU oracle (no GPU).  Each check is chosen so that a plausible
restatement mistake (wrong index, dropped conj/term, transposed operand)
fails: golden fixtures from ace-tn's own tests, published known-answer
vectors, closed forms and brute-force identities."""
import copy

import numpy as np
import pytest
import scipy.linalg

from conftest import GOLDEN_CASES, load_golden
from oracle import ctmrg as oc
from oracle import full_update as ofu
from oracle import measure as om
from oracle import model as omod
from oracle.rng import SketchStream, gaussian_sketch, philox4x32_10


def make_model(name, params):
    return omod.Heisenberg(**params) if name == "heisenberg" else omod.Ising(**params)


# ---------------------------------------------------------------- rng
def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10
    kat = [((0, 0, 0, 0, 0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 6, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344, 0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for inp, out in kat:
        assert tuple(int(v) for v in philox4x32_10(*inp)) == out


def test_sketch_is_standard_normal():
    z = gaussian_sketch(2000, 64, seed=11, stream=4)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    assert abs((z ** 4).mean() - 3) < 0.05
    # distinct streams are not the same draw
    assert not np.allclose(z, gaussian_sketch(2000, 64, seed=11, stream=5))


# ---------------------------------------------------------------- models
def test_pauli_real_y_phase_and_spectra():
    h = omod.Heisenberg(J=1.0).two_site_hamiltonian(None)
    ev = np.sort(np.linalg.eigvalsh(h))
    np.testing.assert_allclose(ev, [-0.75, 0.25, 0.25, 0.25], atol=1e-15)
    g = omod.bond_gate(omod.Ising(jz=1.0, hx=0.0), 0.1, ((0, 0), (1, 0), 2)).reshape(4, 4)
    np.testing.assert_allclose(np.diag(g), np.exp([0.1, -0.1, -0.1, 0.1]), rtol=1e-14)
    h2 = omod.Ising(jz=0.0, hx=1.0)
    g2 = omod.bond_gate(h2, 0.2, ((0, 0), (1, 0), 2)).reshape(4, 4)
    ref = scipy.linalg.expm(0.2 * 0.25 * (np.kron(omod.X, np.eye(2)) + np.kron(np.eye(2), omod.X)))
    np.testing.assert_allclose(g2, ref, rtol=1e-14)


# ---------------------------------------------------------------- golden
@pytest.mark.parametrize("case", list(GOLDEN_CASES))
def test_measure_matches_golden_energy(case):
    st, e_ref, D, chi = load_golden(case)
    name, params = GOLDEN_CASES[case]
    e = om.measure(st, 2, 2, make_model(name, params))['Energy']
    assert e == pytest.approx(e_ref, rel=1e-10)


@pytest.mark.parametrize("case", list(GOLDEN_CASES))
@pytest.mark.parametrize("proj,svd", [("half-system", "rsvd"), ("half-system", "full-rank"),
                                      ("full-system", "full-rank")])
def test_ctmrg_keeps_converged_fixture(case, proj, svd):
    st, e_ref, D, chi = load_golden(case)
    name, params = GOLDEN_CASES[case]
    calc = oc.ProjectorCalculator(proj, svd, omega_fn=SketchStream(1))
    oc.ctmrg(st, 2, 2, chi, 2, calc)
    e = om.measure(st, 2, 2, make_model(name, params))['Energy']
    # the fixtures were converged with half-system projectors; the full-system
    # truncation is a different (equally valid) fixed point at the 1e-7 level
    assert e == pytest.approx(e_ref, rel=2e-8 if proj == "half-system" else 1e-6)


# ---------------------------------------------------------------- closed forms
def product_state(D, d, states):
    out = {}
    for s, v in states.items():
        a = np.zeros((D, D, D, D, d))
        a[0, 0, 0, 0, :] = v
        out[s] = a
    return out


def test_product_state_energies():
    sites = oc.site_list(2, 2)
    neel = product_state(2, 2, {s: ([1, 0] if sum(s) % 2 == 0 else [0, 1]) for s in sites})
    st = oc.make_state(neel)
    calc = oc.ProjectorCalculator(omega_fn=SketchStream(2))
    oc.ctmrg(st, 2, 2, 8, 3, calc)
    assert calc.last_cD == 1
    assert om.measure(st, 2, 2, omod.Heisenberg(J=1.3))['Energy'] == pytest.approx(-0.5 * 1.3, rel=1e-12)
    up = oc.make_state(product_state(2, 2, {s: [1, 0] for s in sites}))
    oc.ctmrg(up, 2, 2, 8, 2, calc)
    assert om.measure(up, 2, 2, omod.Ising(jz=0.7, hx=0.0))['Energy'] == pytest.approx(-2 * 0.7, rel=1e-12)


def test_projectors_are_identity_without_truncation():
    # chi >= N: Q1 (P2 P1^T) Q4 == s0 * Q1 Q4 exactly (rsvd with q >= N is
    # exact; s0 enters through the s/s[0] scaling of projectors.py:163)
    A = oc.random_site_tensors(2, 2, 2, 2, seed=3)
    st = oc.make_state(A)
    sites = oc.plaquette(2, 2, 0, 0, 0)
    for svd in ("rsvd", "full-rank"):
        calc = oc.ProjectorCalculator("half-system", svd, svd_cutoff=0.0, omega_fn=SketchStream(3))
        P1, P2 = calc.calculate(st, sites, 0, chi=64)
        Q1, _ = oc.make_quarter_tensor(st[sites[0]], 0)
        Q4, _ = oc.make_quarter_tensor(st[sites[3]], 3)
        n = P1.shape[-1]
        X = P2.reshape(-1, n) @ P1.reshape(-1, n).T
        R = Q1 @ Q4
        s0 = np.linalg.svd(R, compute_uv=False)[0]
        np.testing.assert_allclose(Q1 @ X @ Q4, s0 * R, atol=1e-10 * s0 * np.abs(R).max())


def test_rsvd_exact_low_rank_recovery():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((120, 9)) @ rng.standard_normal((9, 80))
    B = rng.standard_normal((80, 100))
    U, s, V = oc.fused_matmul_svd_lowrank(A, B, 12, 2, SketchStream(7))
    s_ref = np.linalg.svd(A @ B, compute_uv=False)
    np.testing.assert_allclose(s[:9], s_ref[:9], rtol=1e-10)
    np.testing.assert_allclose((U[:, :9] * s[:9]) @ V[:, :9].T, A @ B, atol=1e-9 * s_ref[0])
    np.testing.assert_allclose(U.T @ U, np.eye(U.shape[1]), atol=1e-10)
    C = rng.standard_normal((100, 70))
    Dm = rng.standard_normal((70, 90))
    U3, s3, V3 = oc.fused_3matmul_svd_lowrank(A, B, C, Dm, 12, 2, SketchStream(8))
    np.testing.assert_allclose(s3[:9], np.linalg.svd(A @ B @ C @ Dm, compute_uv=False)[:9], rtol=1e-10)


# ---------------------------------------------------------------- full update
def test_norm_tensor_trace_identity():
    # <a12|N|a12> with a12 = a1r.a2r equals tr(rho_bond) for the same A's
    st, _, D, chi = load_golden("heisenberg_dims_3_16_dtau_001")
    for bond in oc.bond_list(2, 2)[::3]:
        s1, s2, k = bond
        a1 = np.transpose(st[s1]['A'], ofu.bond_permutation(k))
        a2 = np.transpose(st[s2]['A'], ofu.bond_permutation(k))
        a1q, a1r, a2q, a2r = ofu.decompose_site_tensors(a1, a2)
        n12 = ofu.build_norm_tensor(st, bond, a1q, a2q)
        a12 = np.einsum("yup,xuq->yxpq", a1r, a2r)
        lhs = np.einsum("yxYX,yxpq,YXpq->", n12, a12, a12)
        rho = om.bond_rdm(st, bond)
        assert lhs == pytest.approx(np.einsum("pqpq->", rho), rel=1e-11)


def random_als_problem(nD, bD, pD, seed, exact=False):
    rng = np.random.default_rng(seed)
    a1r = rng.random((nD, bD, pD))
    a2r = rng.random((nD, bD, pD))
    a12g = np.einsum("yup,xuq->yxpq", a1r, a2r)
    if not exact:
        a12g = a12g + 1e-4 * rng.random(a12g.shape)
    nz = rng.standard_normal((nD, nD, nD * nD))
    n12 = np.einsum("xyz,XYz->xyXY", nz, nz)
    return n12, a12g


def test_als_reaches_zero_distance_for_rank_bd_target():
    n12, a12g = random_als_problem(6, 3, 2, seed=1, exact=True)
    a1r, a2r, sweeps, cost = ofu.als_solve(n12, a12g, (6, 3, 2), niter=200, tol=1e-15)
    target = -np.einsum("yxYX,yxpq,YXpq->", n12, a12g, a12g)
    assert cost == pytest.approx(target, rel=1e-9)
    assert sweeps >= 3


def test_als_cost_monotone_and_solver_agrees():
    n12, a12g = random_als_problem(8, 5, 3, seed=2)
    nz = ofu.positive_approx(n12.copy(), 1e-12)
    n12g, a12gg, *_ = ofu.gauge_fix(nz, a12g, 1e-12)
    a1r, a2r = ofu.initialize_reduced_tensors(a12gg, (8, 5, 3))
    costs = [ofu.calculate_cost(a1r, a2r, a12gg, n12g)]
    ng = np.einsum("yxYX,yxpq->YXpq", n12g, a12gg)
    for _ in range(10):
        R = ofu.build_r1(n12g, a2r)
        S = ofu.build_s1(ng, a2r)
        a1r = ofu.solve_ar(R, S)
        Rm = 0.5 * (R.reshape(40, 40) + R.reshape(40, 40).T)
        Rm = Rm + 1e-12 * np.abs(Rm).max() * np.eye(40)
        np.testing.assert_allclose(a1r.reshape(40, 3), np.linalg.solve(Rm, S.reshape(40, 3)),
                                   rtol=1e-8, atol=1e-12)
        a2r = ofu.solve_ar(ofu.build_r2(n12g, a1r), ofu.build_s2(ng, a1r))
        costs.append(ofu.calculate_cost(a1r, a2r, a12gg, n12g))
    assert all(c1 <= c0 + 1e-12 * abs(c0) for c0, c1 in zip(costs, costs[1:]))


def test_gauge_fix_keeps_the_als_cost():
    """gauge_fix (full_update.py:300-343) maps the bond problem to a frame in
    which n12 is closer to the identity, and finalize_reduced_tensors
    (full_update.py:123-161) maps the solution back with nzyr_inv/nzxr_inv.
    The distance is frame-independent: for any reduced pair (b1, b2) of the
    gauged problem, cost(b1, b2; gauged) equals cost(Yinv.b1, Xinv.b2;
    original).  A swapped X/Y, a transposed R factor or a back-map on the
    wrong index breaks the equality."""
    nD, bD, pD = 6, 3, 2
    n12, a12g = random_als_problem(nD, bD, pD, seed=21)
    nz = ofu.positive_approx(n12.copy(), 1e-12)
    n12 = np.einsum("xyz,XYz->xyXY", nz, nz)  # the PSD problem gauge_fix starts from
    n12g, a12gg, xinv, yinv = ofu.gauge_fix(nz, a12g, 1e-12)
    rng = np.random.default_rng(22)
    for _ in range(3):
        b1 = rng.standard_normal((nD, bD, pD))
        b2 = rng.standard_normal((nD, bD, pD))
        c_g = ofu.calculate_cost(b1, b2, a12gg, n12g)
        c_o = ofu.calculate_cost(np.einsum("yz,zup->yup", yinv, b1),
                                 np.einsum("xw,wvq->xvq", xinv, b2), a12g, n12)
        assert c_g == pytest.approx(c_o, rel=1e-10, abs=1e-10 * abs(c_o))
    # the solution found in the gauged frame is the original frame's optimum
    # as well: ALS from the same start in both frames reaches the same cost
    b1, b2, _, cost_g = ofu.als_solve(n12g, a12gg, (nD, bD, pD), niter=60, tol=1e-15)
    c_back = ofu.calculate_cost(np.einsum("yz,zup->yup", yinv, b1),
                                np.einsum("xw,wvq->xvq", xinv, b2), a12g, n12)
    assert c_back == pytest.approx(cost_g, rel=1e-9)
    _, _, _, cost_o = ofu.als_solve(n12, a12g, (nD, bD, pD), niter=400, tol=1e-15)
    assert cost_g <= cost_o + 1e-6 * abs(cost_o)


@pytest.mark.parametrize("gauged", [False, True])
def test_finalize_keeps_the_bond_product(gauged):
    """finalize_reduced_tensors (full_update.py:123-161) redistributes the
    bond: the new pair's product over the bond index is the old product
    P (after the optional back-map) divided by its Frobenius norm, and the
    two halves carry equal Gram matrices diag(s/|s|).  A missing square root
    on s, an unnormalised s or a transposed V factor breaks one of these."""
    nD, bD, pD = 5, 3, 2
    rng = np.random.default_rng(31 + gauged)
    a1r = rng.standard_normal((nD, bD, pD))
    a2r = rng.standard_normal((nD, bD, pD))
    if gauged:
        yinv = rng.standard_normal((nD, nD))
        xinv = rng.standard_normal((nD, nD))
        P = np.einsum("yup,xuq->ypxq", np.einsum("yz,zup->yup", yinv, a1r),
                      np.einsum("xw,wvq->xvq", xinv, a2r))
        b1, b2 = ofu.finalize_reduced_tensors(a1r, a2r, yinv, xinv)
    else:
        P = np.einsum("yup,xuq->ypxq", a1r, a2r)
        b1, b2 = ofu.finalize_reduced_tensors(a1r, a2r)
    Pn = np.einsum("yup,xuq->ypxq", b1, b2)
    np.testing.assert_allclose(Pn, P / np.linalg.norm(P), atol=1e-13)
    s = np.linalg.svd(P.reshape(nD * pD, nD * pD), compute_uv=False)[:bD]
    g1 = np.einsum("yup,yvp->uv", b1, b1)
    g2 = np.einsum("xuq,xvq->uv", b2, b2)
    np.testing.assert_allclose(g1, np.diag(s / np.linalg.norm(s)), atol=1e-13)
    np.testing.assert_allclose(g2, g1, atol=1e-13)


def test_positive_approx_is_psd_square_root():
    rng = np.random.default_rng(4)
    n12 = rng.standard_normal((5, 5, 5, 5))
    nz = ofu.positive_approx(n12.copy(), 1e-12)
    M = np.einsum("xyz,XYz->xyXY", nz, nz).reshape(25, 25)
    assert np.linalg.eigvalsh(M).min() > 0
    # unshifted case: a PSD input is reproduced
    z = rng.standard_normal((5, 5, 25))
    p = np.einsum("xyz,XYz->xyXY", z, z)
    nz = ofu.positive_approx(p.copy(), 1e-12)
    np.testing.assert_allclose(np.einsum("xyz,XYz->xyXY", nz, nz), p, atol=1e-10)


@pytest.mark.parametrize("case", list(GOLDEN_CASES))
def test_full_update_keeps_ground_state(case):
    # the reference's own criterion (test_ground_states.py:33-35) at 3 steps
    st, e_ref, D, chi = load_golden(case)
    name, params = GOLDEN_CASES[case]
    model = make_model(name, params)
    calc = oc.ProjectorCalculator(omega_fn=SketchStream(5))
    mover = oc.DirectionalMover(calc, 2, 2, chi)
    ofu.evolve_full(st, 2, 2, model, 0.01, 3, ofu.FullUpdateConfig(), mover, oc.bond_list(2, 2))
    oc.ctmrg(st, 2, 2, chi, 5, calc)
    assert om.measure(st, 2, 2, model)['Energy'] == pytest.approx(e_ref, rel=1e-4)


def test_simple_update_identity_gate_keeps_bond_product():
    A = oc.random_site_tensors(2, 2, 3, 2, seed=9)
    bonds = oc.bond_list(2, 2)
    b = bonds[0]
    st = {s: {'A': a.copy()} for s, a in A.items()}
    lam = ofu.LambdaTensors(bonds, 3)
    before = np.einsum("lUxdp,xuRDq->lUdpuRDq", st[b[0]]['A'], st[b[1]]['A'])
    a1 = np.transpose(st[b[0]]['A'], ofu.bond_permutation(2))
    a2 = np.transpose(st[b[1]]['A'], ofu.bond_permutation(2))
    _, a1r, _, a2r = ofu.decompose_site_tensors(a1, a2)
    S = np.linalg.svd(np.einsum("yup,xuq->ypxq", a1r, a2r).reshape(a1r.shape[0] * 2, -1),
                      compute_uv=False)[:3]
    ofu.simple_update_bond(st, b, np.eye(4).reshape(2, 2, 2, 2), lam)
    after = np.einsum("lUxdp,xuRDq->lUdpuRDq", st[b[0]]['A'], st[b[1]]['A'])
    np.testing.assert_allclose(after, before / np.linalg.norm(S), atol=1e-12)


def test_mlp_oracle_gradient_matches_finite_differences():
    """The perceptron oracle's backward pass (FP64, no bf16 rounding) against
    central differences of its own loss."""
    from oracle import mlp as omlp
    rng = np.random.default_rng(4)
    fin, h, fout, B = 7, 5, 3, 6
    n = fin * h + h + h * h + h + h * fout + fout
    p = rng.standard_normal(n)
    X, T = rng.standard_normal((B, fin)), rng.standard_normal((B, fout))
    _, g = omlp.loss_and_grad(p, fin, h, fout, X, T, emulate_bf16=False)
    eps = 1e-6
    for i in rng.choice(n, 25, replace=False):
        e = np.zeros(n); e[i] = eps
        lp, _ = omlp.loss_and_grad(p + e, fin, h, fout, X, T, emulate_bf16=False)
        lm, _ = omlp.loss_and_grad(p - e, fin, h, fout, X, T, emulate_bf16=False)
        assert g[i] == pytest.approx((lp - lm) / (2 * eps), rel=1e-5, abs=1e-8)


def test_bf16_rounding_known_values():
    from oracle import mlp as omlp
    # bf16 keeps 8 significant bits; ties go to even
    assert omlp.bf16(1.0) == 1.0
    assert omlp.bf16(1.0 + 2 ** -8) == 1.0            # tie -> even (down)
    assert omlp.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6   # tie -> even (up)
    assert omlp.bf16(1.0 + 2 ** -7) == 1.0 + 2 ** -7
    assert omlp.bf16(-3.0) == -3.0


# ------------------------------------------------------ reference trajectory
def load_reference_trajectory(name):
    import os
    f = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "reference_trajectory.npz"))
    D, chi, aseed, tseed, steps, dtau, nsteps = f[f"{name}/meta"]

    def state(tag):
        st = {}
        for s in [(0, 0), (0, 1), (1, 0), (1, 1)]:
            p = f"{name}/{tag}/{s[0]}{s[1]}_"
            st[s] = {'A': f[p + 'A'], 'C': [f[p + f'C{k}'] for k in range(4)],
                     'E': [f[p + f'E{k}'] for k in range(4)]}
        return st
    return (state("init"), state("final"), float(f[f"{name}/energy"]), int(D), int(chi),
            int(tseed), int(steps), float(dtau), int(nsteps))


class TorchOmega:
    """The reference's sketch: torch.randn(n, q) from torch's global generator,
    one call per projector pair in call order (fused_matmul_svd_lowrank.py:32)."""

    def __init__(self, seed):
        import torch
        torch.manual_seed(seed)

    def __call__(self, n, q):
        import torch
        return torch.randn(n, q, dtype=torch.float64).numpy()


def _corner_spectra_close(a, b, atol):
    for s in a:
        for k in range(4):
            sa = np.linalg.svd(a[s]['C'][k], compute_uv=False)
            sb = np.linalg.svd(b[s]['C'][k], compute_uv=False)
            np.testing.assert_allclose(sa / sa[0], sb / sb[0], rtol=0, atol=atol)


def _replay(name, init, chi, tseed, steps, dtau, nsteps, scale=0.0):
    """The oracle on a recorded reference run: same start (optionally moved by
    ``scale`` relative noise), torch.randn sketches in call order, torch.linalg
    factorisations and no sign gauges (the reference has none)."""
    from oracle import linalg
    A = {s: v['A'] for s, v in init.items()}
    if scale:
        rng = np.random.default_rng(5)
        A = {s: a * (1.0 + scale * rng.standard_normal(a.shape)) for s, a in A.items()}
    with linalg.reference_conventions():
        st = oc.make_state(A)
        calc = oc.ProjectorCalculator(omega_fn=TorchOmega(tseed))
        mover = oc.DirectionalMover(calc, 2, 2, chi)
        for _ in range(steps):
            oc.ctmrg_iteration(mover, st)
        if nsteps:
            ofu.evolve_full(st, 2, 2, omod.Heisenberg(1.0), dtau, nsteps, ofu.FullUpdateConfig(),
                            mover, oc.bond_list(2, 2))
            for _ in range(steps):
                oc.ctmrg_iteration(mover, st)
        return st, om.measure(st, 2, 2, omod.Heisenberg(1.0))['Energy']


@pytest.mark.parametrize("name", ["ctmrg_d2_chi8", "ctmrg_d3_chi12", "evolve_d2_chi8",
                                  "evolve_d2_chi20_100"])
def test_oracle_replays_reference_trajectory(name):
    """The reference run itself (tests/golden/make_reference_trajectory.py:
    torch.randn sketches, torch.linalg factorisations): the oracle's C/E
    initialisation equals the reference's, and the oracle on the same
    sketches with the same factorisation conventions reaches the reference's
    energy within 1e-10 relative, or within the spread the oracle shows when
    its own input moves by one part in 1e15 where that is larger (a trajectory
    that amplifies rounding: at D=2 the corners and the 100-step evolution
    move by 1e-7 to 1e-6 under such a change, so no implementation that sums
    in another order than torch can be held to 1e-10 there); where the
    trajectory does not amplify rounding the final boundary tensors agree to
    1e-10 as well."""
    init, final, e_ref, D, chi, tseed, steps, dtau, nsteps = load_reference_trajectory(name)
    st0 = oc.make_state({s: v['A'] for s, v in init.items()})
    for s in st0:
        for k in range(4):
            np.testing.assert_allclose(st0[s]['C'][k], init[s]['C'][k], rtol=0, atol=1e-15)
            np.testing.assert_allclose(st0[s]['E'][k], init[s]['E'][k], rtol=0, atol=1e-15)
    st, e = _replay(name, init, chi, tseed, steps, dtau, nsteps)
    rel = abs(e - e_ref) / abs(e_ref)
    bound = 1e-10
    if rel > bound:
        _, e_p = _replay(name, init, chi, tseed, steps, dtau, nsteps, scale=1e-15)
        bound = max(bound, 10.0 * abs(e_p - e) / abs(e))
    print(f"{name}: oracle {e:.15f} reference {e_ref:.15f} rel {rel:.1e} (bound {bound:.1e})")
    assert rel <= bound
    if bound <= 1e-10:
        # a trajectory that does not amplify rounding: the boundary tensors too
        dC = max(np.abs(st[s]['C'][k] - final[s]['C'][k]).max() for s in st for k in range(4))
        dE = max(np.abs(st[s]['E'][k] - final[s]['E'][k]).max() for s in st for k in range(4))
        assert max(dC, dE) <= 1e-10
