"""Counter-based Gaussian sketch generator (oracle side).

The reference draws the range-finder sketch with ``torch.randn(n, q)`` inside
``acetn/linalg/fused_matmul_svd_lowrank.py:32`` (and ``svd_lowrank.py:29``,
``fused_3matmul_svd_lowrank.py:35``).  Any iid N(0,1) matrix is a valid
restatement; to make the device engine and this oracle see the *same* sketch
both evaluate the published Philox4x32-10 generator (Salmon, Moraes, Dror,
Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11) followed by the
Box-Muller transform (Box & Muller 1958).

Element (i, j) of an n x q sketch for stream ``stream`` and seed ``seed``:
    e = i*q + j;  call = e // 2;  which = e % 2
    (x0,x1,x2,x3) = philox4x32_10(ctr=(call lo, call hi, stream lo, stream hi),
                                  key=(seed lo, seed hi))
    ua = ((x0 | x1<<32) >> 11 + 0.5) * 2^-53,  ub likewise from (x2, x3)
    z  = sqrt(-2 ln ua) * (cos(2 pi ub) if which == 0 else sin(2 pi ub))
"""
import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = 0x9E3779B9
_W1 = 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32 with 10 rounds.  Inputs are uint32-valued arrays
    (held as uint64); returns four uint64 arrays with uint32 values."""
    c0 = np.asarray(c0, dtype=np.uint64)
    c1 = np.asarray(c1, dtype=np.uint64)
    c2 = np.asarray(c2, dtype=np.uint64)
    c3 = np.asarray(c3, dtype=np.uint64)
    k0 = int(k0) & 0xFFFFFFFF
    k1 = int(k1) & 0xFFFFFFFF
    for r in range(10):
        if r > 0:
            k0 = (k0 + _W0) & 0xFFFFFFFF
            k1 = (k1 + _W1) & 0xFFFFFFFF
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ np.uint64(k0), lo1,
                          hi0 ^ c3 ^ np.uint64(k1), lo0)
    return c0, c1, c2, c3


def gaussian_sketch(n, q, seed, stream):
    """The n x q N(0,1) sketch for (seed, stream), row-major element order."""
    e = np.arange(n * q, dtype=np.uint64)
    call = e >> np.uint64(1)
    which = e & np.uint64(1)
    s = int(stream)
    x0, x1, x2, x3 = philox4x32_10(call & _MASK, call >> np.uint64(32),
                                   np.full_like(call, s & 0xFFFFFFFF),
                                   np.full_like(call, (s >> 32) & 0xFFFFFFFF),
                                   int(seed) & 0xFFFFFFFF, (int(seed) >> 32) & 0xFFFFFFFF)
    a = x0 | (x1 << np.uint64(32))
    b = x2 | (x3 << np.uint64(32))
    ua = ((a >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    ub = ((b >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(ua))
    theta = (2.0 * np.pi) * ub
    z = np.where(which == 0, r * np.cos(theta), r * np.sin(theta))
    return z.reshape(n, q)


class SketchStream:
    """Hands out one sketch per rSVD call, numbering calls in the order the
    reference makes them (one per ``fused_matmul_svd_lowrank`` call)."""

    def __init__(self, seed, first_stream=0):
        self.seed = int(seed)
        self.next_stream = int(first_stream)

    def __call__(self, n, q):
        om = gaussian_sketch(n, q, self.seed, self.next_stream)
        self.next_stream += 1
        return om
