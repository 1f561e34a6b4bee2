"""FP64 parity of the HIP CTMRG step against the oracle (same Philox
sketches, same projector sign gauge).  Compared: projectors (up to the
max-abs scale of the quarter tensors, which the engine does not apply),
renormalised C/E after moves and iterations, and the energy."""
import copy

import numpy as np
import pytest
import torch

from conftest import GOLDEN_CASES, load_golden
from oracle import ctmrg as oc
from oracle import measure as om
from oracle import model as omod
from oracle.rng import SketchStream

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from acetn import _hip
    _hip.require_gpu()
    return _hip


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def to_dev_site(site):
    return (dev(site['A']), [dev(c) for c in site['C']], [dev(e) for e in site['E']])


class DevState:
    """Minimal ipeps-like container over device tensors for the mover."""

    def __init__(self, st, nx, ny, D, chi, seed):
        from acetn.renormalization.projectors import SketchCounter
        self.nx, self.ny = nx, ny
        self.dims = {'bond': D, 'chi': chi, 'phys': 2}
        self.sites = {s: {'A': dev(v['A']), 'C': [dev(c) for c in v['C']],
                          'E': [dev(e) for e in v['E']]} for s, v in st.items()}
        self.site_list = oc.site_list(nx, ny)
        self.sketch = SketchCounter(seed)

    def __getitem__(self, s):
        if tuple(s) not in self.sites:
            raise ValueError(f"Site tensor not defined at site {s}.")
        return self.sites[tuple(s)]

    def to_numpy(self):
        return {s: {'A': v['A'].cpu().numpy(), 'C': [c.cpu().numpy() for c in v['C']],
                    'E': [e.cpu().numpy() for e in v['E']]} for s, v in self.sites.items()}


class Cfg:
    projectors = "half-system"
    svd_type = "rsvd"
    svd_cutoff = 1e-12
    rsvd_niter = 2
    rsvd_oversampling = 2
    disable_progressbar = True
    conv_tol = None


def assert_state_close(a, b, rtol):
    for s in a:
        for k in range(4):
            for key in ('C', 'E'):
                x, y = a[s][key][k], b[s][key][k]
                assert x.shape == y.shape, (s, key, k, x.shape, y.shape)
                scale = np.abs(y).max()
                np.testing.assert_allclose(x, y, atol=rtol * scale, err_msg=f"{s} {key}{k}")


@pytest.mark.parametrize("D,chi,k", [(2, 8, 0), (2, 20, 1), (3, 16, 2), (4, 32, 3)])
def test_half_projectors_match_oracle(hip, D, chi, k):
    st = oc.make_state(oc.random_site_tensors(2, 2, D, 2, seed=100 + D))
    # two oracle iterations so that chi is saturated and C/E are generic
    calc = oc.ProjectorCalculator(omega_fn=SketchStream(9))
    oc.ctmrg(st, 2, 2, chi, 1, calc)
    sites = oc.plaquette(2, 2, 0, 1, k)
    seed, stream = 77, 5
    ocalc = oc.ProjectorCalculator(omega_fn=SketchStream(seed, stream))
    P1o, P2o = ocalc.calculate(st, sites, k, chi)
    P1h, P2h, spec = hip.half_projectors(to_dev_site(st[sites[0]]), to_dev_site(st[sites[3]]), k,
                                         chi, 2, 2, 1e-12, seed, stream)
    assert P1h.shape[1] == P1o.shape[-1] == ocalc.last_cD
    m1 = np.abs(hip.quarter_tensor(to_dev_site(st[sites[0]]), k).cpu().numpy()).max()
    m4 = np.abs(hip.quarter_tensor(to_dev_site(st[sites[3]]), k + 3).cpu().numpy()).max()
    n = P1h.shape[1]
    np.testing.assert_allclose(P1h.cpu().numpy() / m1, P1o.reshape(-1, n), rtol=0,
                               atol=1e-10 * np.abs(P1o).max())
    np.testing.assert_allclose(P2h.cpu().numpy() / m4, P2o.reshape(-1, n), rtol=0,
                               atol=1e-10 * np.abs(P2o).max())
    s = spec.cpu().numpy()
    q = len(ocalc.last_spectrum)
    np.testing.assert_allclose(s[:q] / s[0], ocalc.last_spectrum, atol=1e-13)


@pytest.mark.parametrize("D,chi", [(2, 8), (2, 20), (3, 12)])
def test_ctmrg_iterations_match_oracle(hip, D, chi):
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover
    st0 = oc.make_state(oc.random_site_tensors(2, 2, D, 2, seed=200 + chi))
    seed = 4242
    ost = copy.deepcopy(st0)
    ocalc = oc.ProjectorCalculator(omega_fn=SketchStream(seed))
    omover = oc.DirectionalMover(ocalc, 2, 2, chi)
    hst = DevState(st0, 2, 2, D, chi, seed)
    hmover = DirectionalMover(Cfg(), hst.sketch)
    for it in range(3):
        oc.ctmrg_iteration(omover, ost)
        ctmrg_iteration(hmover, hst)
        assert_state_close(hst.to_numpy(), ost, rtol=1e-9)
    e_o = om.measure(ost, 2, 2, omod.Heisenberg(1.0))['Energy']
    e_h = om.measure(hst.to_numpy(), 2, 2, omod.Heisenberg(1.0))['Energy']
    assert e_h == pytest.approx(e_o, rel=1e-11)


def product_start(D, seed):
    """The default start of an acetn network (site_tensor.py:120-136): the
    polarised product state plus 1e-2 U[0,1) noise, normalised, with C/E
    built from A.  Its half-system spectra fall over ten decades within the
    first chi values, so the power-iterated sketch is graded far beyond what
    its Gram matrix resolves."""
    rng = np.random.default_rng(seed)
    st = {}
    for s in oc.site_list(2, 2):
        a = 1e-2 * rng.random((D, D, D, D, 2))
        a[0, 0, 0, 0, 0] += 1.0
        st[s] = a / np.linalg.norm(a)
    return oc.make_state(st)


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_half_projectors_on_graded_spectrum_match_oracle(hip, k):
    """D=6 chi=36 from the default start: the kept count, all chi singular
    values (down to ~1e-10 s0) and the projector columns as the oracle's
    Householder-QR range finder gives them."""
    D, chi = 6, 36
    st = product_start(D, 31)
    sites = oc.plaquette(2, 2, 0, 1, k)
    seed, stream = 13, 2
    ocalc = oc.ProjectorCalculator(omega_fn=SketchStream(seed, stream))
    P1o, P2o = ocalc.calculate(st, sites, k, chi)
    P1h, P2h, spec = hip.half_projectors(to_dev_site(st[sites[0]]), to_dev_site(st[sites[3]]), k,
                                         chi, 2, 2, 1e-12, seed, stream)
    s = spec.cpu().numpy()
    so = ocalc.last_spectrum
    rel = np.abs(s[:chi] / s[0] - so[:chi]) / so[:chi]
    print(f"k={k}: kept hip {P1h.shape[1]} oracle {ocalc.last_cD}; s_chi/s_0 {so[chi - 1]:.3e}; "
          f"max rel diff of the kept values {rel.max():.2e}")
    assert P1h.shape[1] == ocalc.last_cD == chi
    assert so[chi - 1] < 1e-8            # the case is graded
    assert rel.max() < 1e-7
    # every projector column as the oracle's, the leading ones (well-separated
    # values) to near rounding
    m1 = np.abs(hip.quarter_tensor(to_dev_site(st[sites[0]]), k).cpu().numpy()).max()
    lead = int(np.sum(so[:chi] > 1e-5))
    P1all = P1h.cpu().numpy() / m1
    P1rall = P1o.reshape(-1, chi)
    col = np.abs(P1all - P1rall).max(axis=0) / np.abs(P1rall).max(axis=0)
    print(f"largest column difference of P1 {col.max():.1e} (relative to the column), "
          f"{col[:lead].max():.1e} over the {lead} leading columns")
    assert col[:lead].max() < 1e-10
    assert col.max() < 1e-7


def test_full_system_projectors_on_graded_spectrum_match_oracle(hip):
    """Full-system rSVD (fused_3matmul_svd_lowrank, a basis orthonormalised
    between the two halves of every power iteration) from the default start:
    the product of four corners squares the graded half-system spectrum, so
    the kept count falls below chi -- the same count, values and CTMRG step
    as the oracle's."""
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover

    class FCfg(Cfg):
        projectors = "full-system"
    D, chi = 6, 36
    st0 = product_start(D, 34)
    seed = 717
    ost = copy.deepcopy(st0)
    ocalc = oc.ProjectorCalculator(projectors="full-system", omega_fn=SketchStream(seed))
    omover = oc.DirectionalMover(ocalc, 2, 2, chi)
    hst = DevState(st0, 2, 2, D, chi, seed)
    hmover = DirectionalMover(FCfg(), hst.sketch)
    oc.ctmrg_iteration(omover, ost)
    ctmrg_iteration(hmover, hst)
    h = hst.to_numpy()
    shapes = {h[s]['C'][k].shape for s in h for k in range(4)}
    print("full-system from the default start: corner shapes", sorted(shapes))
    assert any(sh[0] < chi or sh[1] < chi for sh in shapes)     # truncated below chi
    assert_state_close(h, ost, rtol=1e-7)
    e_o = om.measure(ost, 2, 2, omod.Heisenberg(1.0))['Energy']
    e_h = om.measure(h, 2, 2, omod.Heisenberg(1.0))['Energy']
    assert e_h == pytest.approx(e_o, rel=1e-12)


def test_range_finder_fast_path_is_identical_on_ungraded_spectra(hip):
    """RF_FAST (dropped columns refilled at random) and the default recovery
    give the same bits when the spectrum is not graded (no column dropped):
    the projector calculator takes the fast path for such pairs."""
    st = oc.make_state(oc.random_site_tensors(2, 2, 3, 2, seed=41))
    oc.ctmrg(st, 2, 2, 16, 1, oc.ProjectorCalculator(omega_fn=SketchStream(9)))
    sites = oc.plaquette(2, 2, 0, 1, 0)
    dsites = [to_dev_site(st[s]) for s in sites]
    out = []
    for fast in (False, True):
        job = hip.ProjectorJob(0, dsites, 0, 16, 2, 2, 1e-12, 21, 3, rf_fast=fast)
        job.ready.synchronize()
        P1, P2 = job.finish()
        torch.cuda.synchronize()
        out.append((P1.cpu(), P2.cpu(), job.spec.cpu(), job.spec_host.clone()))
    s = out[0][3]
    assert float(s[-1]) >= 1e-3 * float(s[0])     # not graded
    for a, b in zip(out[0][:3], out[1][:3]):
        assert torch.equal(a, b)


def test_ctmrg_iterations_from_default_start_match_oracle(hip):
    """Three CTMRG iterations at D=6 chi=36 from the default start (graded
    spectra in every projector pair) against the oracle, same sketches.  For
    scale: the oracle's own C/E move by ~1e-2 of their largest elements when
    only the sketch changes, so agreement at 1e-7 needs the same range finder
    down to the smallest kept values, not just an equally good one."""
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover
    D, chi = 6, 36
    st0 = product_start(D, 32)
    seed = 515
    ost = copy.deepcopy(st0)
    omover = oc.DirectionalMover(oc.ProjectorCalculator(omega_fn=SketchStream(seed)), 2, 2, chi)
    hst = DevState(st0, 2, 2, D, chi, seed)
    hmover = DirectionalMover(Cfg(), hst.sketch)
    for it in range(3):
        oc.ctmrg_iteration(omover, ost)
        ctmrg_iteration(hmover, hst)
        h = hst.to_numpy()
        err = max(np.abs(h[s][key][kk] - ost[s][key][kk]).max() / np.abs(ost[s][key][kk]).max()
                  for s in h for key in ('C', 'E') for kk in range(4))
        print(f"iteration {it}: max C/E difference {err:.1e} of the largest element")
        assert_state_close(h, ost, rtol=1e-7)
    e_o = om.measure(ost, 2, 2, omod.Heisenberg(1.0))['Energy']
    e_h = om.measure(hst.to_numpy(), 2, 2, omod.Heisenberg(1.0))['Energy']
    assert e_h == pytest.approx(e_o, rel=1e-12)


@pytest.mark.parametrize("case", list(GOLDEN_CASES))
def test_ctmrg_on_golden_fixture(hip, case):
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover
    st, e_ref, D, chi = load_golden(case)
    name, params = GOLDEN_CASES[case]
    model = omod.Heisenberg(**params) if name == "heisenberg" else omod.Ising(**params)
    seed = 99
    ost = copy.deepcopy(st)
    omover = oc.DirectionalMover(oc.ProjectorCalculator(omega_fn=SketchStream(seed)), 2, 2, chi)
    hst = DevState(st, 2, 2, D, chi, seed)
    hmover = DirectionalMover(Cfg(), hst.sketch)
    for _ in range(2):
        oc.ctmrg_iteration(omover, ost)
        ctmrg_iteration(hmover, hst)
    e_o = om.measure(ost, 2, 2, model)['Energy']
    e_h = om.measure(hst.to_numpy(), 2, 2, model)['Energy']
    # HIP vs oracle on the same sketches
    assert e_h == pytest.approx(e_o, rel=1e-10)
    # and the converged fixture stays converged (reference tolerance class)
    assert e_h == pytest.approx(e_ref, rel=2e-8)


class MixedCfg(Cfg):
    precision = "mixed"


# energy bound of the mixed-precision engine against the FP64 oracle
MIXED_DE = 1e-6


@pytest.mark.parametrize("D,chi", [(3, 12), (4, 32)])
def test_mixed_ctmrg_energy_matches_oracle(hip, D, chi):
    """Split-bf16 contractions + FP64 QR/SVD: three iterations from a random
    state stay within |dE/site| <= 1e-6 of the FP64 oracle (same sketches).
    (A random D=2, chi=20 state is left out: there the FP64 oracle itself
    moves by ~1e-4 when only the sketch seed changes, so no arithmetic can be
    held to 1e-6 on it; the mixed engine moves it by ~1e-5.)"""
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover
    st0 = oc.make_state(oc.random_site_tensors(2, 2, D, 2, seed=300 + chi, bias=0.5))
    seed = 77
    ost = copy.deepcopy(st0)
    omover = oc.DirectionalMover(oc.ProjectorCalculator(omega_fn=SketchStream(seed)), 2, 2, chi)
    hst = DevState(st0, 2, 2, D, chi, seed)
    hmover = DirectionalMover(MixedCfg(), hst.sketch)
    for _ in range(3):
        oc.ctmrg_iteration(omover, ost)
        ctmrg_iteration(hmover, hst)
    e_o = om.measure(ost, 2, 2, omod.Heisenberg(1.0))['Energy']
    e_h = om.measure(hst.to_numpy(), 2, 2, omod.Heisenberg(1.0))['Energy']
    print(f"D={D} chi={chi}: mixed {e_h:.15f} oracle {e_o:.15f} dE {abs(e_h - e_o):.2e}")
    assert abs(e_h - e_o) <= MIXED_DE


def test_mixed_ctmrg_from_default_start(hip):
    """Mixed precision from the default start at D=6 chi=36 (graded
    spectra): three iterations keep every boundary shape of the FP64 oracle
    (same sketches) and its energy per site to |dE| <= 1e-6."""
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover
    D, chi = 6, 36
    st0 = product_start(D, 33)
    seed = 616
    ost = copy.deepcopy(st0)
    omover = oc.DirectionalMover(oc.ProjectorCalculator(omega_fn=SketchStream(seed)), 2, 2, chi)
    hst = DevState(st0, 2, 2, D, chi, seed)
    hmover = DirectionalMover(MixedCfg(), hst.sketch)
    for _ in range(3):
        oc.ctmrg_iteration(omover, ost)
        ctmrg_iteration(hmover, hst)
    h = hst.to_numpy()
    for s in h:
        for key in ('C', 'E'):
            for kk in range(4):
                assert h[s][key][kk].shape == ost[s][key][kk].shape, (s, key, kk)
    e_o = om.measure(ost, 2, 2, omod.Heisenberg(1.0))['Energy']
    e_h = om.measure(h, 2, 2, omod.Heisenberg(1.0))['Energy']
    print(f"default start D=6 chi=36: mixed {e_h:.15f} oracle {e_o:.15f} dE {abs(e_h - e_o):.2e}")
    assert abs(e_h - e_o) <= MIXED_DE


@pytest.mark.parametrize("case", list(GOLDEN_CASES))
def test_mixed_ctmrg_on_golden_fixture(hip, case):
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover
    st, e_ref, D, chi = load_golden(case)
    name, params = GOLDEN_CASES[case]
    model = omod.Heisenberg(**params) if name == "heisenberg" else omod.Ising(**params)
    hst = DevState(st, 2, 2, D, chi, 5)
    hmover = DirectionalMover(MixedCfg(), hst.sketch)
    for _ in range(2):
        ctmrg_iteration(hmover, hst)
    e_h = om.measure(hst.to_numpy(), 2, 2, model)['Energy']
    print(f"{case}: mixed {e_h:.15f} golden {e_ref:.15f} dE {abs(e_h - e_ref):.2e}")
    assert abs(e_h - e_ref) <= MIXED_DE


@pytest.mark.parametrize("projectors,svd_type,D,chi", [("full-system", "rsvd", 2, 8),
                                                        ("full-system", "rsvd", 3, 12),
                                                        ("half-system", "full-rank", 2, 8),
                                                        ("full-system", "full-rank", 2, 8)])
def test_other_projector_branches_match_oracle(hip, projectors, svd_type, D, chi):
    """Full-system rSVD (fused_3matmul_svd_lowrank) and full-rank SVD
    projectors (projectors.py:62-136, :176-217): three CTMRG iterations
    against the oracle's same branch."""
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover

    class BCfg(Cfg):
        pass
    BCfg.projectors, BCfg.svd_type = projectors, svd_type
    st0 = oc.make_state(oc.random_site_tensors(2, 2, D, 2, seed=500 + chi, bias=0.5))
    seed = 31
    ost = copy.deepcopy(st0)
    ocalc = oc.ProjectorCalculator(projectors=projectors, svd_type=svd_type,
                                   omega_fn=SketchStream(seed))
    omover = oc.DirectionalMover(ocalc, 2, 2, chi)
    hst = DevState(st0, 2, 2, D, chi, seed)
    hmover = DirectionalMover(BCfg(), hst.sketch)
    for _ in range(3):
        oc.ctmrg_iteration(omover, ost)
        ctmrg_iteration(hmover, hst)
    assert_state_close(hst.to_numpy(), ost, rtol=1e-8)
    e_o = om.measure(ost, 2, 2, omod.Heisenberg(1.0))['Energy']
    e_h = om.measure(hst.to_numpy(), 2, 2, omod.Heisenberg(1.0))['Energy']
    # The device SVD is a two-pass Gram (Jacobi) SVD: a kept singular pair
    # s_j is resolved to about eps * s_0 / s_j, where the oracle's LAPACK SVD
    # is backward stable.  The full-rank branch keeps the whole spectrum down
    # to the cutoff (s/s_0 ~ 3e-6 in these cases, ~1e-10 in the pair), the
    # rSVD branches only the leading q values.
    rel = 1e-9 if svd_type == "full-rank" else 1e-10
    print(f"{projectors}/{svd_type}: dE/E = {abs(e_h - e_o) / abs(e_o):.2e}")
    assert e_h == pytest.approx(e_o, rel=rel)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("D,chi,cdi,cdj", [(3, 12, 9, 11), (4, 64, 40, 64), (4, 64, 64, 40)])
def test_renormalize_in_two_parts_matches_one_call(hip, precision, D, chi, cdi, cdj):
    """The two-part renormalisation, pair j's part first (P1_j contracted
    first, as the reference does) or pair i's (P2_i first), the other count
    only bounded at part 1, gives the one-call result for all four k.  At
    D=4, chi=64 the last edge product at the layout bound (64 columns) is
    large enough for the split-bf16 plane path and at 40 columns it is not,
    so part 2 must keep the path (and the element type of the intermediate)
    that part 1 chose."""
    A = oc.random_site_tensors(1, 1, D, 2, seed=41)[(0, 0)]
    rng = np.random.default_rng(41)
    C = [rng.standard_normal((chi, chi)) for _ in range(4)]
    E = [rng.standard_normal((chi, chi, D, D)) for _ in range(4)]
    site = (dev(A), [dev(c) for c in C], [dev(e) for e in E])
    P1i, P2i = (dev(rng.standard_normal((chi * D * D, cdi))) for _ in range(2))
    P1j, P2j = (dev(rng.standard_normal((chi * D * D, cdj))) for _ in range(2))
    # FP64: only the split-K partition may differ; mixed: FP32 partial sums
    tol = 1e-12 if precision == "fp64" else 1e-5
    for k in range(4):
        ref = hip.renormalize_boundary(site, k, P1i, P2i, P1j, P2j, precision=precision)
        for first in ("j", "i"):
            parts = hip.RenormalizeParts(site, k, chi, first_pair=first,
                                         slot=("test", k, first, chi), precision=precision)
            if first == "j":
                parts.first(P1j, P2j)
                out = parts.second(P1i, P2i)
            else:
                parts.first(P1i, P2i)
                out = parts.second(P1j, P2j)
            torch.cuda.synchronize()
            for name, a, b in zip(("C_a", "C_b", "E_o"), ref, out):
                a, b = a.cpu().numpy(), b.cpu().numpy()
                assert a.shape == b.shape, (k, first, name)
                np.testing.assert_allclose(b, a, rtol=tol, atol=tol * np.abs(a).max(),
                                           err_msg=f"{k} {first} {name}")


@pytest.mark.parametrize("nx,ny", [(3, 2), (3, 3)])
def test_ctmrg_larger_unit_cell_matches_oracle(hip, nx, ny):
    """Unit cells with more than two sites along a line: a stage then has
    more projector pairs than stream slots, so pairs share slots (each with
    its own workspace) and every slot stream must be ordered after the
    caller's stream, where the boundary tensors were uploaded."""
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover
    D, chi, seed = 2, 8, 606
    st0 = oc.make_state(oc.random_site_tensors(nx, ny, D, 2, seed=60 + nx * ny, bias=0.3))
    ost = copy.deepcopy(st0)
    omover = oc.DirectionalMover(oc.ProjectorCalculator(omega_fn=SketchStream(seed)), nx, ny, chi)
    hst = DevState(st0, nx, ny, D, chi, seed)
    hmover = DirectionalMover(Cfg(), hst.sketch)
    for _ in range(2):
        oc.ctmrg_iteration(omover, ost)
        ctmrg_iteration(hmover, hst)
        assert_state_close(hst.to_numpy(), ost, rtol=1e-9)


@pytest.mark.parametrize("mode", ["stagger", "free"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nx,ny,D,chi", [(2, 2, 3, 12), (2, 2, 4, 32), (3, 2, 2, 8)])
def test_pipelined_iteration_equals_stage_by_stage(hip, mode, precision, nx, ny, D, chi):
    """run_chain_pair (the left/right and up/down chains pipelined, with or
    without mark events between the chains) against one stage at a time:
    bit for bit."""
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover
    st = oc.make_state(oc.random_site_tensors(nx, ny, D, 2, seed=700 + D + nx))
    cfg = Cfg() if precision == "fp64" else MixedCfg()
    a = DevState(st, nx, ny, D, chi, 31)
    b = DevState(st, nx, ny, D, chi, 31)
    ma, mb = DirectionalMover(cfg, a.sketch), DirectionalMover(cfg, b.sketch)
    for it in range(3):
        ctmrg_iteration(ma, a, pipelined=mode)
        ctmrg_iteration(mb, b, pipelined=False)
        assert a.sketch.next == b.sketch.next
        for s in a.site_list:
            for key in ('C', 'E'):
                for k in range(4):
                    x, y = a[s][key][k], b[s][key][k]
                    assert x.shape == y.shape and torch.equal(x, y), (it, s, key, k)
