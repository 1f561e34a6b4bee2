"""Checkpoints: the reference's pickled save format is read without executing
the pickle (acetn.ipeps.checkpoint), and the engine's plain-tensor format
round-trips."""
import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN_CASES, load_golden

REF_GS = "/root/reference/tests/integration/ipeps_gs"


@pytest.mark.skipif(not os.path.isdir(REF_GS), reason="reference checkout not present")
@pytest.mark.parametrize("case", list(GOLDEN_CASES))
def test_reference_checkpoint_reads_the_fixture_tensors(case):
    """The reader gives exactly the arrays of the committed fixtures, which the
    reference's energies.csv pins (tests/test_oracle.py)."""
    from acetn.ipeps.checkpoint import is_reference_checkpoint, read_reference_checkpoint
    path = os.path.join(REF_GS, case + ".pt")
    assert is_reference_checkpoint(path)
    got = read_reference_checkpoint(path, 2, 2)
    st, _, D, chi = load_golden(case)
    for s in st:
        assert np.array_equal(got[s]['A'], st[s]['A'])
        for k in range(4):
            assert np.array_equal(got[s]['C'][k], st[s]['C'][k])
            assert np.array_equal(got[s]['E'][k], st[s]['E'][k])
    with pytest.raises(ValueError, match="expected"):
        read_reference_checkpoint(path, 3, 2)


def test_plain_checkpoint_is_not_taken_for_a_reference_one(tmp_path):
    from acetn.ipeps.checkpoint import is_reference_checkpoint
    p = os.path.join(tmp_path, "plain.pt")
    torch.save({"format": "acetn-mi355x/1", "tensors": {"0,0": {"A": torch.zeros(2)}}}, p)
    assert not is_reference_checkpoint(p)
    q = os.path.join(tmp_path, "not_a_zip.pt")
    with open(q, "wb") as f:
        f.write(b"not a zip")
    assert not is_reference_checkpoint(q)


@pytest.mark.gpu
def test_engine_checkpoint_round_trip(tmp_path):
    from acetn import _hip
    from acetn.ipeps import Ipeps
    _hip.require_gpu()
    st, e_ref, D, chi = load_golden("heisenberg_dims_3_16_dtau_001")
    config = {'dtype': 'float64', 'device': 'cuda',
              'TN': {'nx': 2, 'ny': 2, 'dims': {'phys': 2, 'bond': D, 'chi': chi}},
              'model': {'name': 'heisenberg', 'params': {'J': 1.0}},
              'ctmrg': {'steps': 1, 'sketch_seed': 1, 'disable_progressbar': True}}
    a = Ipeps(config)
    a.load_arrays(st)
    p = os.path.join(tmp_path, "ipeps.pt")
    a.save(p)
    b = Ipeps(config)
    b.load(p)
    for s in a.site_list:
        assert torch.equal(a[s]['A'], b[s]['A'])
        for k in range(4):
            assert torch.equal(a[s]['C'][k], b[s]['C'][k])
            assert torch.equal(a[s]['E'][k], b[s]['E'][k])
    assert float(b.measure()['Energy']) == pytest.approx(e_ref, rel=1e-10)
