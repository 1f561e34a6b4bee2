"""Parity of the HIP building blocks (through the C-ABI) against numpy / the
oracle.  FP64 throughout; tolerances are stated per check."""
import numpy as np
import pytest
import torch

from oracle import ctmrg as oc
from oracle.rng import gaussian_sketch as np_sketch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from acetn import _hip
    _hip.require_gpu()
    print("device:", _hip.device_info(), flush=True)
    return _hip


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def test_sketch_matches_oracle_generator(hip):
    for n, q, seed, stream in [(7, 3, 0, 0), (1024, 66, 12345, 17), (5184, 146, 2 ** 40 + 3, 2 ** 33 + 5)]:
        z = hip.gaussian_sketch(n, q, seed, stream).cpu().numpy()
        ref = np_sketch(n, q, seed, stream)
        # same Philox bits; Box-Muller transcendental functions may differ in the last ulp
        np.testing.assert_allclose(z, ref, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("M,N,K,ta,tb", [(1, 1, 1, False, False), (37, 53, 29, False, False),
                                         (64, 64, 64, True, False), (130, 70, 300, False, True),
                                         (100, 146, 5000, True, False), (512, 66, 1024, False, False),
                                         # small-output long-K products (per-lane dot kernel)
                                         (4, 4, 100000, False, False), (3, 7, 5000, True, True),
                                         (8, 8, 4096, False, True), (1, 5, 70000, True, False)])
def test_gemm_matches_numpy(hip, M, N, K, ta, tb):
    rng = np.random.default_rng(M * 1000 + N)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    C0 = rng.standard_normal((M, N))
    ref = 0.5 * ((A.T if ta else A) @ (B.T if tb else B)) - 2.0 * C0
    out = hip.gemm(dev(A), dev(B), ta=ta, tb=tb, alpha=0.5, beta=-2.0, C=dev(C0)).cpu().numpy()
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12 * np.sqrt(K))


GEMM_SHAPES = [(1, 1, 1, False, False), (37, 53, 29, False, False), (64, 64, 64, True, False),
               (130, 70, 300, False, True), (100, 146, 5000, True, False),
               (1296, 146, 1296, False, False), (146, 146, 5184, True, False)]


@pytest.mark.parametrize("M,N,K,ta,tb", GEMM_SHAPES)
def test_gemm_split_error_model(hip, M, N, K, ta, tb):
    """Split-bf16 products: every element within the error model of the
    split (operand truncation c_p |A||B| plus FP32 accumulation), and the
    three-piece split an order of magnitude tighter than the two-piece one."""
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    C0 = rng.standard_normal((M, N))
    Ao, Bo = (A.T if ta else A), (B.T if tb else B)
    ref = 0.5 * (Ao @ Bo) - 2.0 * C0
    mag = 0.5 * (np.abs(Ao) @ np.abs(Bo))
    u32 = 2.0 ** -24
    errs = {}
    for pieces, c_p in ((2, 3 * 2.0 ** -17), (3, 4 * u32)):
        out = hip.gemm(dev(A), dev(B), ta=ta, tb=tb, alpha=0.5, beta=-2.0, C=dev(C0),
                       pieces=pieces).cpu().numpy()
        err = np.abs(out - ref)
        bound = (c_p + 8 * np.sqrt(K) * u32) * mag + 4 * u32 * np.abs(ref)
        assert np.all(err <= bound), (pieces, float((err / bound).max()))
        errs[pieces] = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    if K >= 29:
        assert errs[3] < errs[2] / 10, errs


def test_gemm_split_exact_on_small_integers(hip):
    """Operands exactly representable in one bf16 piece with sums below 2^24:
    both splits reproduce the product exactly."""
    rng = np.random.default_rng(11)
    A = rng.integers(-8, 9, size=(96, 64)).astype(np.float64)
    B = rng.integers(-8, 9, size=(64, 80)).astype(np.float64)
    for pieces in (2, 3):
        out = hip.gemm(dev(A), dev(B), pieces=pieces).cpu().numpy()
        np.testing.assert_array_equal(out, A @ B)


@pytest.mark.parametrize("q", [1, 2, 5, 16, 33, 66, 146, 202])
def test_eigh_matches_lapack(hip, q):
    rng = np.random.default_rng(q)
    X = rng.standard_normal((q, q))
    G = X @ X.T + np.diag(np.linspace(0, 1, q))
    w, V = hip.eigh(dev(G))
    w, V = w.cpu().numpy(), V.cpu().numpy()
    wr = np.linalg.eigvalsh(G)[::-1]
    np.testing.assert_allclose(w, wr, rtol=1e-12, atol=1e-12 * wr[0])
    np.testing.assert_allclose(V.T @ V, np.eye(q), atol=1e-12)
    np.testing.assert_allclose(G @ V, V * w, atol=1e-11 * wr[0])


def test_eigh_graded_and_degenerate(hip):
    # graded spectrum down to 1e-14 relative and an exact 3-fold degeneracy
    rng = np.random.default_rng(7)
    Qm, _ = np.linalg.qr(rng.standard_normal((40, 40)))
    lam = np.concatenate([[1, 1, 1], np.logspace(-1, -14, 37)])
    G = (Qm * lam) @ Qm.T
    w, V = hip.eigh(dev(G))
    w, V = w.cpu().numpy(), V.cpu().numpy()
    np.testing.assert_allclose(w, np.sort(lam)[::-1], atol=1e-15)
    np.testing.assert_allclose(V.T @ V, np.eye(40), atol=1e-12)


def _svd_checks(hip, M, rtol_vec=1e-12):
    """Singular values against LAPACK (absolute eps-class accuracy, as a
    backward-stable SVD has), orthonormal left vectors, and U^T M = S V^T
    (the rows of U^T M are orthogonal with norms s)."""
    rows, cols = M.shape
    s, U, sweeps = hip.svd(dev(M))
    s, U = s.cpu().numpy(), U.cpu().numpy()
    sr = np.linalg.svd(M, compute_uv=False)
    k = min(rows, cols)
    assert sweeps > 0, f"no convergence ({-sweeps} sweeps)"
    np.testing.assert_allclose(s, sr[:k], rtol=0, atol=4 * max(rows, cols) * 2.2e-16 * sr[0])
    # a singular vector is resolved to about eps s0 / s_k (as by LAPACK):
    # vector checks on the values above 1e-4 s0, scaled by s0 / s_min there
    live = s > 1e-4 * s[0]
    Ul = U[:, :k][:, live]
    amp = s[0] / s[live].min()
    np.testing.assert_allclose(Ul.T @ Ul, np.eye(Ul.shape[1]), atol=1e-14 * amp + 1e-13)
    W = Ul.T @ M   # = S V^T
    np.testing.assert_allclose(np.linalg.norm(W, axis=1), s[live], rtol=0,
                               atol=(1e-14 * amp + 1e-13) * sr[0])
    G = W @ W.T
    np.testing.assert_allclose(G - np.diag(np.diag(G)), 0, atol=(1e-14 * amp + 1e-13) * sr[0] ** 2)
    return s, U, sweeps


@pytest.mark.parametrize("rows,cols", [(1, 1), (5, 5), (22, 22), (66, 66), (146, 146), (150, 146),
                                       (146, 160), (192, 192), (202, 202), (272, 272), (300, 280)])
def test_jacobi_svd_matches_lapack(hip, rows, cols):
    rng = np.random.default_rng(rows * 1000 + cols)
    M = rng.standard_normal((rows, cols))
    s, U, sweeps = _svd_checks(hip, M)
    print(f"{rows}x{cols}: {sweeps} sweeps")
    # well-separated random spectrum: the vectors themselves match LAPACK's up to sign
    Ur = np.linalg.svd(M)[0]
    k = min(rows, cols)
    d = np.abs(np.sum(U[:, :k] * Ur[:, :k], axis=0))
    sr = np.linalg.svd(M, compute_uv=False)
    gap = np.minimum(np.abs(np.diff(np.r_[np.inf, sr])), np.abs(np.diff(np.r_[sr, -np.inf])))[:k]
    ok = gap > 1e-3 * sr[0]
    assert np.all(1 - d[ok] < 1e-14 + 1e-12 * (sr[0] / gap[ok]) ** 2)


def test_jacobi_svd_graded_degenerate_and_rank_deficient(hip):
    rng = np.random.default_rng(11)
    # graded spectrum down to 1e-14 relative with an exact 3-fold degeneracy
    q = 146
    U0, _ = np.linalg.qr(rng.standard_normal((q, q)))
    V0, _ = np.linalg.qr(rng.standard_normal((q, q)))
    sv = np.concatenate([[1, 1, 1], np.logspace(-1, -14, q - 3)])
    s, U, _ = _svd_checks(hip, (U0 * sv) @ V0.T)
    # values below q eps |M| are roundoff (the Jacobi sweeps leave columns
    # that small alone): absolute accuracy q eps |M|_F
    np.testing.assert_allclose(s, sv, rtol=0, atol=q * 1.1e-16 * np.sqrt(3.0))
    # rank 40 of 120 (zero singular values past the rank)
    M = rng.standard_normal((120, 40)) @ rng.standard_normal((40, 120))
    s, U, _ = _svd_checks(hip, M)
    assert np.all(s[40:] <= 1e-12 * s[0])
    # a zero matrix and a single nonzero column
    s, U, _ = hip.svd(dev(np.zeros((7, 7))))
    assert float(s.abs().max()) == 0.0
    M = np.zeros((9, 9)); M[:, 4] = rng.standard_normal(9)
    s, U, _ = _svd_checks(hip, M)
    np.testing.assert_allclose(s[0], np.linalg.norm(M[:, 4]), rtol=1e-15)


@pytest.mark.parametrize("m,q,rank", [(200, 20, 20), (1024, 66, 66), (300, 30, 1), (80, 22, 16),
                                      (16, 16, 12), (600, 100, 100), (2000, 146, 146),
                                      (1500, 146, 60), (2000, 170, 170), (900, 173, 150),
                                      (1200, 174, 174), (1200, 220, 200)])
def test_orthonormalize(hip, m, q, rank):
    # q orthonormal columns always (rank-deficient sketches are completed, as
    # Householder QR does), and their span contains range(Y)
    rng = np.random.default_rng(m + q)
    Y = rng.standard_normal((m, rank)) @ rng.standard_normal((rank, q))
    Qd = hip.orthonormalize(dev(Y)).cpu().numpy()
    np.testing.assert_allclose(Qd.T @ Qd, np.eye(q), atol=1e-13)
    np.testing.assert_allclose(Qd @ (Qd.T @ Y), Y, atol=1e-11 * np.abs(Y).max())


@pytest.mark.parametrize("D,chi,k", [(2, 4, 0), (2, 8, 1), (3, 9, 2), (4, 16, 3), (3, 16, 0)])
def test_quarter_tensor_matches_oracle(hip, D, chi, k):
    A = oc.random_site_tensors(1, 1, D, 2, seed=D * 10 + k)[(0, 0)]
    rng = np.random.default_rng(chi + k)
    C = [rng.standard_normal((chi, chi)) for _ in range(4)]
    E = [rng.standard_normal((chi, chi, D, D)) for _ in range(4)]
    site = {'A': A, 'C': C, 'E': E}
    Qo, _ = oc.make_quarter_tensor(site, k)
    Qh = hip.quarter_tensor((dev(A), [dev(c) for c in C], [dev(e) for e in E]), k).cpu().numpy()
    Qh = Qh / np.abs(Qh).max()
    np.testing.assert_allclose(Qh, Qo, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("D,chi,k", [(2, 8, 1), (3, 9, 2), (4, 16, 3), (6, 12, 0)])
def test_quarter_tensor_mixed_within_fp32_error(hip, D, chi, k):
    """Mixed precision: the T2 product on split-bf16 planes (three pieces)
    and the double-layer absorption in FP32 on the matrix cores; every
    element within FP32-class error of the oracle (relative to max |Q|)."""
    A = oc.random_site_tensors(1, 1, D, 2, seed=D * 10 + k + 5)[(0, 0)]
    rng = np.random.default_rng(chi + k + 5)
    C = [rng.standard_normal((chi, chi)) for _ in range(4)]
    E = [rng.standard_normal((chi, chi, D, D)) for _ in range(4)]
    site = {'A': A, 'C': C, 'E': E}
    Qo, _ = oc.make_quarter_tensor(site, k)
    Qh = hip.quarter_tensor((dev(A), [dev(c) for c in C], [dev(e) for e in E]), k,
                            precision="mixed").cpu().numpy()
    Qh = Qh / np.abs(Qh).max()
    err = np.abs(Qh - Qo).max()
    # FP32 accumulation over K = chi, then over D^4 d terms: a few 1e-7
    assert err < 5e-6, err
    assert err > 0.0  # the mixed path really ran (FP64 would agree to 1e-13)
