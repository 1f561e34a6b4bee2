"""Randomized SVDs of a matrix or of a product of matrices
(acetn/linalg/{svd_lowrank,fused_matmul_svd_lowrank,fused_3matmul_svd_lowrank}.py)
on the HIP engine (acetn_hip_rsvd_product): the product is never formed, the
sketch is a Philox Gaussian drawn under a seed taken from torch's default
generator (so torch.manual_seed makes calls reproducible, as it does for the
reference's torch.randn sketch), QR is CholeskyQR3 (the spans of Householder
QR) and the SVD of the projected matrix is the engine's two-pass Gram + Jacobi
eigensolver.  Same signatures and return convention as the reference:
(U, S, V) with A ~ U diag(S) V^T, V not transposed.

FP64 real matrices only; inputs on the CPU are moved to the GPU for the
computation and the results returned on the inputs' device."""
import torch

from .. import _hip

__all__ = ["svd_lowrank", "fused_matmul_svd_lowrank", "fused_3matmul_svd_lowrank"]


def _run(mats, q, niter, mid_qr):
    for t in mats:
        if t.is_complex():
            raise TypeError("acetn HIP engine: complex matrices are not supported (the kernels "
                            "are real FP64)")
    _hip.require_gpu()
    home = mats[0].device
    dev = home if home.type == "cuda" else torch.device("cuda", torch.cuda.current_device())
    ops = [t.detach().to(device=dev, dtype=torch.float64) for t in mats]
    seed = int(torch.randint(0, 2 ** 62, (1,), dtype=torch.int64))
    U, S, V = _hip.rsvd_product(ops, q, niter, mid_qr, seed, 0)
    dtype = mats[0].dtype if mats[0].dtype in (torch.float32, torch.float64) else torch.float64
    return tuple(x.to(device=home, dtype=dtype) for x in (U, S, V))


def svd_lowrank(A, q=6, niter=2):
    """linalg/svd_lowrank.py:4-44."""
    return _run([A], q, niter, False)


def fused_matmul_svd_lowrank(A, B, q=6, niter=2):
    """linalg/fused_matmul_svd_lowrank.py:4-52: the randomized SVD of A @ B."""
    return _run([A, B], q, niter, False)


def fused_3matmul_svd_lowrank(A, B, C, D, q=6, niter=2):
    """linalg/fused_3matmul_svd_lowrank.py:4-56: the randomized SVD of
    A @ B @ C @ D, with the extra orthonormalisation between the two halves
    of every power iteration."""
    return _run([A, B, C, D], q, niter, True)
