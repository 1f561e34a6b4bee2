"""The entry point's smoke check (one CTMRG iteration at D=2, chi=8 on a
minimal network container that only supports indexing by site, against the
oracle to 1e-10 in energy) runs as part of the GPU suite."""
import os
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_smoke_entry_point():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if root not in sys.path:
        sys.path.insert(0, root)
    import __graft_entry__
    __graft_entry__.smoke()
