import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "ace-tn_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: longer CPU oracle checks")


def load_golden(case):
    import numpy as np
    f = np.load(os.path.join(ROOT, "tests", "golden", case + ".npz"))
    st = {}
    for s in [(0, 0), (0, 1), (1, 0), (1, 1)]:
        p = f"{s[0]}{s[1]}_"
        st[s] = {'A': f[p + 'A'].copy(), 'C': [f[p + f'C{k}'].copy() for k in range(4)],
                 'E': [f[p + f'E{k}'].copy() for k in range(4)]}
    return st, float(f['energy']), int(f['D']), int(f['chi'])


GOLDEN_CASES = {
    "ising_dims_2_20_dtau_001_hx_295": ("ising", dict(jz=1.0, hx=2.95)),
    "heisenberg_dims_3_16_dtau_001": ("heisenberg", dict(J=1.0)),
}
