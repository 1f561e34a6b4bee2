"""The ALS entry point that allocates nothing (acetn_hip_als_solve_noalloc,
fixed device scratch) against the workspace entry point acetn_hip_als_solve
on the same problems: two calls in a row, the second smaller, both equal to
the workspace call (same kernel, same inputs); a bond too large for the
fixed scratch is refused with an error."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from acetn import _hip
    _hip.require_gpu()
    return _hip


def problem(D, d, seed):
    """A positive definite norm tensor n12 (nD, nD, nD, nD) and a gated
    pair a12g (nD, nD, d, d), nD = D d, seeded."""
    rng = np.random.default_rng(seed)
    nD = D * d
    X = rng.standard_normal((nD * nD, nD * nD))
    n12 = (X @ X.T / (nD * nD) + 0.1 * np.eye(nD * nD)).reshape(nD, nD, nD, nD)
    a12g = rng.standard_normal((nD, nD, d, d))
    t = lambda a: torch.as_tensor(a, dtype=torch.float64, device="cuda").contiguous()
    return t(n12), t(a12g)


def test_noalloc_matches_workspace_entry_point(hip):
    for D, d, seed in [(6, 2, 1), (3, 2, 2)]:   # the second call smaller than the first
        print(f"als_solve_noalloc D={D} d={d} seed={seed}", flush=True)
        n12, a12g = problem(D, d, seed)
        a1, a2 = hip.als_initialize(a12g, D)
        r1, r2, rs, rc, rf = hip.als_solve(n12, a12g, a1, a2, 100, 1e-12, "cholesky", 1e-12)
        b1, b2 = a1.clone().contiguous(), a2.clone().contiguous()
        s, c, f = hip.als_solve_noalloc(n12, a12g, b1, b2, 100, 1e-12, "cholesky", 1e-12)
        torch.cuda.synchronize()
        assert (s, f) == (rs, rf)
        assert c == rc
        assert torch.equal(b1, r1) and torch.equal(b2, r2)
        assert s >= 3


def test_noalloc_refuses_bonds_beyond_its_scratch(hip):
    n12, a12g = problem(3, 2, 3)
    big = torch.zeros(32, 8, 4, dtype=torch.float64, device="cuda")  # D = 8, d = 4: nD * D = 256
    # operands of mismatched shapes never reach the library
    with pytest.raises(ValueError, match="shape"):
        hip.als_solve_noalloc(n12, a12g, big, big.clone(), 10, 1e-12, "cholesky", 1e-12)
    n12b = torch.zeros((32,) * 4, dtype=torch.float64, device="cuda")
    a12gb = torch.zeros(32, 32, 4, 4, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="at most 128"):
        hip.als_solve_noalloc(n12b, a12gb, big, big.clone(), 10, 1e-12, "cholesky", 1e-12)
