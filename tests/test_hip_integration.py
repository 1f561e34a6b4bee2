"""The reference-side ALS adapter that INTEGRATION.md shows a maintainer
adding to acetn (acetn/evolution/_extensions/_hip.py, bound in place of
_C_cutensor.als_solve, csrc/evolution/als_solve.cpp:107-137): the code block
is taken from INTEGRATION.md as written, executed against this build's
library, and driven the way ALSSolver.solve drives the cuTENSOR extension
(als_solver.py:46-53, 84-110): the plain start, n12g, then the adapter."""
import os
import re

import numpy as np
import pytest
import torch

from conftest import ROOT
from oracle import full_update as ofu
from oracle.ctmrg import ein
from test_hip_full_update import als_problem, dev, host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def adapter():
    from acetn import _hip
    _hip.require_gpu()
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```python\n(.*?)```", text, flags=re.S)
    code = [b for b in blocks if b.startswith("# acetn/evolution/_extensions/_hip.py")]
    assert len(code) == 1, "INTEGRATION.md: the adapter block is missing"
    os.environ["ACETN_HIP_LIB"] = _hip.LIB_PATH
    ns = {"__name__": "acetn_hip_adapter"}
    exec(compile(code[0], "INTEGRATION.md:_hip.py", "exec"), ns)
    return ns["als_solve"]


@pytest.mark.parametrize("method,bond_idx", [("cholesky", 2), ("cholesky", 6), ("pinv", 0)])
def test_integration_adapter_matches_oracle(adapter, method, bond_idx):
    n12, a12g, shape = als_problem(bond_idx)
    start = ofu.initialize_reduced_tensors(a12g, shape)
    niter, tol, eps = 100, 1e-15, 1e-12
    r1, r2, sweeps, cost = ofu.als_solve(n12, a12g, shape, niter, tol, method, eps, start=start)
    n12g = ein("yxYX,yxpq->YXpq", n12, a12g)
    a1r0, a2r0 = dev(start[0]), dev(start[1])
    h1, h2 = adapter(a1r0, a2r0, dev(n12g), dev(n12), dev(a12g), niter, tol, method, eps)
    torch.cuda.synchronize()
    # results are new tensors; the inputs are untouched (the pybind contract)
    assert h1.data_ptr() != a1r0.data_ptr() and h2.data_ptr() != a2r0.data_ptr()
    np.testing.assert_array_equal(host(a1r0), start[0])
    assert h1.shape == a1r0.shape and h1.dtype == torch.float64
    c = ofu.calculate_cost(host(h1), host(h2), a12g, n12)
    print(f"adapter {method} bond {bond_idx}: oracle sweeps {sweeps} cost {cost:.15e} "
          f"adapter cost {c:.15e}")
    assert c == pytest.approx(cost, rel=1e-10)
    np.testing.assert_allclose(ein("yup,xuq->yxpq", host(h1), host(h2)),
                               ein("yup,xuq->yxpq", r1, r2), atol=1e-8 * np.abs(a12g).max())


def test_integration_adapter_reports_errors(adapter):
    """Errors of the library become RuntimeError with its message (the
    extension's TORCH_CHECK behaviour): an unknown method is rejected before
    the call, a bond too large for the device solver by the library."""
    n12, a12g, shape = als_problem(2)
    start = ofu.initialize_reduced_tensors(a12g, shape)
    with pytest.raises(KeyError):
        adapter(dev(start[0]), dev(start[1]), None, dev(n12), dev(a12g), 10, 1e-12, "qr", 1e-12)
    D, d = 9, 2  # the one-workgroup solver holds systems up to about nD D = 140
    nD = D * d
    rng = np.random.default_rng(5)
    a1r = dev(rng.standard_normal((nD, D, d)))
    with pytest.raises(RuntimeError, match="too large for one workgroup"):
        adapter(a1r, a1r.clone(), None, dev(rng.standard_normal((nD,) * 4)),
                dev(rng.standard_normal((nD, nD, d, d))), 10, 1e-12, "cholesky", 1e-12)
    nD = 33  # beyond the reduced dimensions the library accepts
    a1r = dev(rng.standard_normal((nD, 2, d)))
    with pytest.raises(RuntimeError, match="unsupported reduced tensor shape"):
        adapter(a1r, a1r.clone(), None, dev(np.zeros((nD,) * 4)),
                dev(np.zeros((nD, nD, d, d))), 10, 1e-12, "cholesky", 1e-12)
