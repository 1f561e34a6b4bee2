"""HIP reduced density matrices and measure() against the oracle and the
golden ground-state energies (ace-tn tests/integration/ipeps_gs/energies.csv,
checked to rel 1e-10 as test_ground_states.py:30 does)."""
import numpy as np
import pytest
import torch

from conftest import GOLDEN_CASES, load_golden
from oracle import ctmrg as oc
from oracle import measure as om

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from acetn import _hip
    _hip.require_gpu()
    return _hip


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def dsite(site):
    return (dev(site['A']), [dev(c) for c in site['C']], [dev(e) for e in site['E']])


@pytest.mark.parametrize("case", list(GOLDEN_CASES))
def test_rdms_match_oracle_on_fixture(hip, case):
    st, e_ref, D, chi = load_golden(case)
    for s in oc.site_list(2, 2):
        np.testing.assert_allclose(hip.site_rdm(dsite(st[s])).cpu().numpy(), om.site_rdm(st[s]),
                                   rtol=1e-11, atol=1e-14)
    for b in oc.bond_list(2, 2):
        ref = om.bond_rdm(st, b)
        got = hip.bond_rdm(dsite(st[b[0]]), dsite(st[b[1]]), b[2]).cpu().numpy()
        np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("D,chi", [(2, 6), (3, 10)])
def test_rdms_match_oracle_random_rectangular(hip, D, chi):
    # after one oracle CTMRG iteration the boundary dims are generic
    from oracle.rng import SketchStream
    st = oc.make_state(oc.random_site_tensors(2, 2, D, 2, seed=31 + D))
    oc.ctmrg(st, 2, 2, chi, 1, oc.ProjectorCalculator(omega_fn=SketchStream(2)))
    for b in oc.bond_list(2, 2):
        ref = om.bond_rdm(st, b)
        got = hip.bond_rdm(dsite(st[b[0]]), dsite(st[b[1]]), b[2]).cpu().numpy()
        np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())
    for s in oc.site_list(2, 2):
        np.testing.assert_allclose(hip.site_rdm(dsite(st[s])).cpu().numpy(), om.site_rdm(st[s]),
                                   rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("case", list(GOLDEN_CASES))
def test_measure_api_golden_energy(hip, case):
    from acetn.ipeps import Ipeps
    st, e_ref, D, chi = load_golden(case)
    name, params = GOLDEN_CASES[case]
    ip = Ipeps({'TN': {'nx': 2, 'ny': 2, 'dims': {'phys': 2, 'bond': D, 'chi': chi}},
                'model': {'name': name, 'params': params},
                'ctmrg': {'disable_progressbar': True}})
    ip.load_arrays(st)
    e = ip.measure()['Energy'].item()
    assert e == pytest.approx(e_ref, rel=1e-10)
