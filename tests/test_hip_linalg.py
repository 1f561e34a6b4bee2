"""acetn.linalg on the HIP engine (acetn_hip_rsvd_product): the reference's
own unit tests of svd_lowrank / fused_matmul_svd_lowrank /
fused_3matmul_svd_lowrank (tests/unit/test_linalg.py:6-368, same matrices,
seeds and bounds; its complex cases are refused here, the engine being real
FP64), and parity with the oracle's restatements of the three functions on
the same Philox sketch."""
import numpy as np
import pytest
import torch

from oracle import ctmrg as oc
from oracle.rng import gaussian_sketch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from acetn import _hip
    _hip.require_gpu()
    import acetn.linalg as linalg
    return linalg


def rel(X, Y):
    return float(torch.norm(X - Y) / torch.norm(X))


def recon(U, S, V):
    return U @ torch.diag(S) @ V.T


# ---- svd_lowrank (test_linalg.py:6-97)
@pytest.fixture
def random_matrix():
    torch.manual_seed(42)
    return torch.randn(100, 80, dtype=torch.float64)


@pytest.fixture
def low_rank_matrix():
    torch.manual_seed(42)
    U = torch.randn(100, 10, dtype=torch.float64)
    V = torch.randn(80, 10, dtype=torch.float64)
    return U @ V.T


def test_svd_lowrank_shapes_and_reconstruction(L, random_matrix):
    A = random_matrix
    U, S, V = L.svd_lowrank(A, q=20, niter=2)
    assert U.shape == (100, 20) and S.shape == (20,) and V.shape == (80, 20)
    U, S, V = L.svd_lowrank(A, q=40, niter=2)
    assert rel(A, recon(U, S, V)) < 0.5


def test_svd_lowrank_low_rank_exact_recovery(L, low_rank_matrix):
    A = low_rank_matrix
    U, S, V = L.svd_lowrank(A, q=15, niter=3)
    assert rel(A, recon(U, S, V)) < 1e-10


def test_svd_lowrank_ordering_and_orthogonality(L, random_matrix):
    U, S, V = L.svd_lowrank(random_matrix, q=20, niter=2)
    assert torch.all(S >= 0) and torch.all(S[:-1] >= S[1:])
    I = torch.eye(20, dtype=torch.float64)
    assert torch.allclose(U.T @ U, I, atol=1e-10)
    assert torch.allclose(V.T @ V, I, atol=1e-10)


def test_svd_lowrank_comparison_with_torch(L, random_matrix):
    A = random_matrix
    torch.manual_seed(456)
    U1, S1, V1 = L.svd_lowrank(A, q=20, niter=2)
    torch.manual_seed(456)
    U2, S2, V2 = torch.svd_lowrank(A, q=20, niter=2)
    assert abs(rel(A, recon(U1, S1, V1)) - rel(A, recon(U2, S2, V2))) < 0.1


# ---- fused_matmul_svd_lowrank (test_linalg.py:100-174)
@pytest.fixture
def matrix_pair():
    torch.manual_seed(42)
    return torch.randn(60, 50, dtype=torch.float64), torch.randn(50, 70, dtype=torch.float64)


def test_fused_matmul(L, matrix_pair):
    A, B = matrix_pair
    C = A @ B
    U, S, V = L.fused_matmul_svd_lowrank(A, B, q=20, niter=2)
    assert U.shape == (60, 20) and S.shape == (20,) and V.shape == (70, 20)
    torch.manual_seed(789)
    e1 = rel(C, recon(*L.svd_lowrank(C, q=25, niter=2)))
    torch.manual_seed(789)
    e2 = rel(C, recon(*L.fused_matmul_svd_lowrank(A, B, q=25, niter=2)))
    assert abs(e1 - e2) < 0.05
    assert rel(C, recon(*L.fused_matmul_svd_lowrank(A, B, q=30, niter=2))) < 0.5


def test_fused_matmul_low_rank_exact_recovery(L):
    torch.manual_seed(42)
    A = torch.randn(60, 8, dtype=torch.float64)
    B = torch.randn(8, 70, dtype=torch.float64)
    U, S, V = L.fused_matmul_svd_lowrank(A, B, q=12, niter=3)
    assert rel(A @ B, recon(U, S, V)) < 1e-10
    # the directions of the zero singular values still come out orthonormal
    assert torch.allclose(V.T @ V, torch.eye(12, dtype=torch.float64), atol=1e-10)


# ---- fused_3matmul_svd_lowrank (test_linalg.py:177-256)
@pytest.fixture
def matrix_quad():
    torch.manual_seed(42)
    return (torch.randn(40, 35, dtype=torch.float64), torch.randn(35, 30, dtype=torch.float64),
            torch.randn(30, 35, dtype=torch.float64), torch.randn(35, 45, dtype=torch.float64))


def test_fused_3matmul(L, matrix_quad):
    A, B, C, D = matrix_quad
    F = A @ B @ C @ D
    U, S, V = L.fused_3matmul_svd_lowrank(A, B, C, D, q=20, niter=2)
    assert U.shape == (40, 20) and S.shape == (20,) and V.shape == (45, 20)
    assert torch.all(S >= 0) and torch.all(S[:-1] >= S[1:])
    torch.manual_seed(101)
    e1 = rel(F, recon(*L.svd_lowrank(F, q=20, niter=2)))
    torch.manual_seed(101)
    e2 = rel(F, recon(*L.fused_3matmul_svd_lowrank(A, B, C, D, q=20, niter=2)))
    assert abs(e1 - e2) < 0.05
    assert rel(F, recon(*L.fused_3matmul_svd_lowrank(A, B, C, D, q=25, niter=2))) < 0.5


def test_fused_3matmul_low_rank_exact_recovery(L):
    torch.manual_seed(42)
    A = torch.randn(40, 6, dtype=torch.float64)
    B = torch.eye(6, dtype=torch.float64)
    C = torch.eye(6, dtype=torch.float64)
    D = torch.randn(6, 45, dtype=torch.float64)
    U, S, V = L.fused_3matmul_svd_lowrank(A, B, C, D, q=10, niter=3)
    assert rel(A @ B @ C @ D, recon(U, S, V)) < 1e-10


# ---- complex matrices (test_linalg.py:259-307): refused by the real engine
def test_complex_matrices_are_refused(L):
    torch.manual_seed(42)
    A = torch.randn(40, 35, dtype=torch.complex128)
    B = torch.randn(35, 45, dtype=torch.complex128)
    with pytest.raises(TypeError, match="complex"):
        L.fused_matmul_svd_lowrank(A, B, q=15, niter=2)
    with pytest.raises(TypeError, match="complex"):
        L.svd_lowrank(A, q=20, niter=2)


# ---- edge cases (test_linalg.py:310-368)
def test_edge_cases(L):
    torch.manual_seed(42)
    A = torch.randn(20, 15, dtype=torch.float64)
    U, S, V = L.svd_lowrank(A, q=50, niter=2)
    assert U.shape[1] <= 15 and S.shape[0] <= 15 and V.shape[1] <= 15
    # q = min(m, n): the full SVD
    assert rel(A, recon(U, S, V)) < 1e-12
    torch.manual_seed(42)
    A = torch.randn(30, 30, dtype=torch.float64)
    assert rel(A, recon(*L.svd_lowrank(A, q=15, niter=2))) < 0.6
    torch.manual_seed(42)
    U, S, V = L.svd_lowrank(torch.randn(100, 20, dtype=torch.float64), q=15, niter=2)
    assert U.shape == (100, 15) and V.shape == (20, 15)
    torch.manual_seed(42)
    U, S, V = L.svd_lowrank(torch.randn(20, 100, dtype=torch.float64), q=15, niter=2)
    assert U.shape == (20, 15) and V.shape == (100, 15)
    torch.manual_seed(42)
    A = torch.randn(50, 40, dtype=torch.float64)
    assert rel(A, recon(*L.svd_lowrank(A, q=20, niter=0))) < 1.0


# ---- parity with the oracle on the same sketch
@pytest.mark.parametrize("case", ["svd", "fused", "fused3"])
@pytest.mark.parametrize("niter", [0, 2])
def test_matches_oracle_on_the_same_sketch(L, case, niter):
    from acetn import _hip
    rng = np.random.default_rng(7)
    seed, stream, q = 12345, 3, 40
    if case == "svd":
        mats = [rng.standard_normal((300, 260))]
        ref = oc.svd_lowrank(mats[0], q, niter, lambda n, k: gaussian_sketch(n, k, seed, stream))
    elif case == "fused":
        mats = [rng.standard_normal((300, 200)), rng.standard_normal((200, 260))]
        ref = oc.fused_matmul_svd_lowrank(*mats, q, niter,
                                          lambda n, k: gaussian_sketch(n, k, seed, stream))
    else:
        mats = [rng.standard_normal(s) for s in ((150, 120), (120, 100), (100, 130), (130, 170))]
        ref = oc.fused_3matmul_svd_lowrank(*mats, q, niter,
                                           lambda n, k: gaussian_sketch(n, k, seed, stream))
    ops = [torch.as_tensor(a, device="cuda") for a in mats]
    U, S, V = _hip.rsvd_product(ops, q, niter, case == "fused3", seed, stream)
    Uo, So, Vo = ref
    S, U, V = S.cpu().numpy(), U.cpu().numpy(), V.cpu().numpy()
    assert S.shape == So.shape
    np.testing.assert_allclose(S, So, rtol=0, atol=1e-10 * So[0])
    R, Ro = (U * S) @ V.T, (Uo * So) @ Vo.T
    assert np.abs(R - Ro).max() <= 1e-10 * np.abs(Ro).max()
    np.testing.assert_allclose(U.T @ U, np.eye(q), atol=1e-12)
    np.testing.assert_allclose(V.T @ V, np.eye(q), atol=1e-12)


@pytest.mark.parametrize("niter", [0, 2])
def test_graded_spectrum_matches_oracle(L, niter):
    """svd_lowrank of a matrix whose 40 singular values fall from 1 to 1e-13
    (the grading of a CTMRG corner product at a product-state start): the
    power-iterated sketch is graded far beyond its Gram matrix's resolution,
    and every value and the low-rank product still agree with the oracle's
    Householder-QR path on the same sketch."""
    from acetn import _hip
    rng = np.random.default_rng(11)
    m, n, r, q = 300, 260, 40, 40
    s_true = 10.0 ** (-13.0 * np.arange(r) / (r - 1))
    Ua, _ = np.linalg.qr(rng.standard_normal((m, r)))
    Va, _ = np.linalg.qr(rng.standard_normal((n, r)))
    A = (Ua * s_true) @ Va.T
    seed, stream = 2024, 5
    Uo, So, Vo = oc.svd_lowrank(A, q, niter, lambda nn, k: gaussian_sketch(nn, k, seed, stream))
    U, S, V = _hip.rsvd_product([torch.as_tensor(A, device="cuda")], q, niter, False, seed, stream)
    S, U, V = S.cpu().numpy(), U.cpu().numpy(), V.cpu().numpy()
    live = So > 1e-12 * So[0]
    rel = np.abs(S[live] - So[live]) / So[live]
    print(f"niter={niter}: {int(live.sum())} values above 1e-12 s0, max rel diff {rel.max():.1e}")
    # 1e-9 relative, plus an absolute 1e-17 s0 (a tenth of eps s0, the
    # resolution both paths have for the smallest values: at 1e-12 s0 they
    # differ from the exact values by ~7e-7 relative themselves)
    assert np.all(rel <= 1e-9 + 1e-17 * So[0] / So[live])
    R, Ro = (U * S) @ V.T, (Uo * So) @ Vo.T
    assert np.abs(R - Ro).max() <= 1e-12 * np.abs(Ro).max()
