"""Dense factorisations of the oracle, in one place.

Test infrastructure only.  The oracle computes with numpy's LAPACK and fixes
the signs of projector columns and bond indices with its own gauge rules
(shared with the device engine).  The reference's trajectory instead depends
on the signs torch.linalg's CPU LAPACK returns (its rSVD sketches act on the
boundary tensors in whatever basis those signs leave), so a replay of a
recorded reference run switches to torch.linalg and drops the gauge rules
inside ``reference_conventions()``.
"""
import contextlib

import numpy as np


class _State:
    torch = False   # factorisations by torch.linalg (CPU) instead of numpy
    gauges = True   # the oracle's projector / bond sign gauges


state = _State()


def _t(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a))


def qr(a):
    if state.torch:
        import torch
        q, r = torch.linalg.qr(_t(a))
        return q.numpy(), r.numpy()
    return np.linalg.qr(a)


def svd(a, full_matrices=True):
    if state.torch:
        import torch
        u, s, vh = torch.linalg.svd(_t(a), full_matrices=full_matrices)
        return u.numpy(), s.numpy(), vh.numpy()
    return np.linalg.svd(a, full_matrices=full_matrices)


def eigh(a):
    if state.torch:
        import torch
        w, v = torch.linalg.eigh(_t(a))
        return w.numpy(), v.numpy()
    return np.linalg.eigh(a)


@contextlib.contextmanager
def reference_conventions():
    saved = (state.torch, state.gauges)
    state.torch, state.gauges = True, False
    try:
        yield
    finally:
        state.torch, state.gauges = saved
