"""FP64 numpy restatement of the learned-start perceptron's forward and
backward passes (TEST INFRASTRUCTURE ONLY; checks ace-tn_amd/csrc/mlp.hip).

With ``emulate_bf16`` every matrix-product operand is rounded to bf16 (round
to nearest even) exactly where the device kernels round it; accumulation is
FP64 here and FP32 on the device.  Without it, the functions are plain FP64
calculus, pinned by finite differences in tests/test_oracle.py.
"""
import numpy as np


def bf16(x):
    x32 = np.asarray(x, dtype=np.float32)
    b = x32.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, fin, h, fout):
    p = np.asarray(params, dtype=np.float64)
    o = 0
    W1 = p[o:o + fin * h].reshape(fin, h); o += fin * h
    b1 = p[o:o + h]; o += h
    W2 = p[o:o + h * h].reshape(h, h); o += h * h
    b2 = p[o:o + h]; o += h
    W3 = p[o:o + h * fout].reshape(h, fout); o += h * fout
    b3 = p[o:o + fout]
    return W1, b1, W2, b2, W3, b3


def forward(params, fin, h, fout, X, emulate_bf16=True):
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(params, fin, h, fout)
    H1 = np.maximum(r(X) @ r(W1) + b1, 0.0)
    H2 = np.maximum(r(H1) @ r(W2) + b2, 0.0)
    return r(H2) @ r(W3) + b3, H1, H2


def loss_and_grad(params, fin, h, fout, X, T, emulate_bf16=True):
    """Mean squared error over all outputs and its gradient (flat, params layout)."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(params, fin, h, fout)
    Y, H1, H2 = forward(params, fin, h, fout, X, emulate_bf16)
    n = Y.size
    loss = float(np.sum((Y - T) ** 2) / n)
    dY = 2.0 * (Y - T) / n
    gW3 = r(H2).T @ r(dY)
    gb3 = dY.sum(axis=0)
    dH2 = (r(dY) @ r(W3).T) * (H2 > 0)
    gW2 = r(H1).T @ r(dH2)
    gb2 = dH2.sum(axis=0)
    dH1 = (r(dH2) @ r(W2).T) * (H1 > 0)
    gW1 = r(X).T @ r(dH1)
    gb1 = dH1.sum(axis=0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, grad
