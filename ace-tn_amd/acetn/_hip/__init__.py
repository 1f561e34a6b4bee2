"""ctypes boundary to ``libacetn_hip.so`` (the gfx950 engine, C-ABI declared in
include/acetn_hip.h).

Mirrors the role of acetn/evolution/_extensions/__init__.py:12-35 (extension
discovery + availability flag), with one deliberate difference: a caller
that asks for the 'hip' backend on a machine where the library or the GPU is
missing gets an exception, never a silent CPU path.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libacetn_hip.so")

c_i64 = ctypes.c_int64
c_vp = ctypes.c_void_p
c_dbl = ctypes.c_double
c_size = ctypes.c_size_t
c_u64 = ctypes.c_uint64


class AcetnSite(ctypes.Structure):
    """acetn_site of include/acetn_hip.h."""
    _fields_ = [("A", c_vp), ("C", c_vp * 4), ("E", c_vp * 4),
                ("C_dim", (c_i64 * 2) * 4), ("E_dim", (c_i64 * 2) * 4),
                ("D", c_i64), ("d", c_i64)]


_lib = None
_load_error = None


def _declare(lib):
    sig = {
        "acetn_hip_last_error": (ctypes.c_char_p, []),
        "acetn_hip_device_info": (ctypes.c_int, [ctypes.c_char_p, c_size]),
        "acetn_hip_autotune": (ctypes.c_int, [ctypes.c_int]),
        "acetn_hip_autotune_entries": (ctypes.c_int, [ctypes.POINTER(c_i64)]),
        "acetn_hip_autotune_clear": (ctypes.c_int, []),
        "acetn_hip_autotune_save": (ctypes.c_int, [ctypes.c_char_p]),
        "acetn_hip_autotune_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(c_i64)]),
        "acetn_hip_kernel_timing": (ctypes.c_int, [ctypes.c_int]),
        "acetn_hip_kernel_timing_bytes": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(c_dbl)]),
        "acetn_hip_kernel_timing_read": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(c_dbl),
                                                        ctypes.POINTER(c_dbl),
                                                        ctypes.POINTER(c_i64)]),
        "acetn_hip_quarter_tensor_workspace": (c_size, [ctypes.POINTER(AcetnSite), ctypes.c_int]),
        "acetn_hip_quarter_tensor": (ctypes.c_int, [ctypes.POINTER(AcetnSite), ctypes.c_int,
                                                    ctypes.c_int, c_vp, c_vp, c_size, c_vp]),
        "acetn_hip_half_projectors_workspace": (c_size, [ctypes.POINTER(AcetnSite),
                                                         ctypes.POINTER(AcetnSite), ctypes.c_int,
                                                         c_i64, c_i64]),
        "acetn_hip_half_projectors": (ctypes.c_int, [ctypes.POINTER(AcetnSite),
                                                     ctypes.POINTER(AcetnSite), ctypes.c_int, c_i64,
                                                     c_i64, ctypes.c_int, c_dbl, c_u64, c_u64,
                                                     ctypes.c_int, c_vp, c_vp, ctypes.POINTER(c_i64),
                                                     c_vp, c_vp, c_size, c_vp]),
        "acetn_hip_half_projectors_begin": (ctypes.c_int, [ctypes.POINTER(AcetnSite),
                                                           ctypes.POINTER(AcetnSite), ctypes.c_int,
                                                           c_i64, c_i64, ctypes.c_int, c_dbl, c_u64,
                                                           c_u64, ctypes.c_int, c_vp, c_vp, c_vp,
                                                           c_size, c_vp]),
        "acetn_hip_half_projectors_end": (ctypes.c_int, [ctypes.POINTER(AcetnSite),
                                                         ctypes.POINTER(AcetnSite), ctypes.c_int,
                                                         c_i64, c_i64, ctypes.c_int, ctypes.c_int,
                                                         c_i64, c_vp, c_vp, c_vp, c_size, c_vp]),
        "acetn_hip_projectors_workspace": (c_size, [ctypes.POINTER(AcetnSite) * 4, ctypes.c_int,
                                                    c_i64, c_i64, ctypes.c_int, ctypes.c_int]),
        "acetn_hip_projectors_begin": (ctypes.c_int, [ctypes.POINTER(AcetnSite) * 4, ctypes.c_int,
                                                      c_i64, c_i64, ctypes.c_int, c_dbl, c_u64,
                                                      c_u64, ctypes.c_int, ctypes.c_int,
                                                      ctypes.c_int, c_vp, c_vp, c_vp, c_size,
                                                      c_vp]),
        "acetn_hip_projectors_end": (ctypes.c_int, [ctypes.POINTER(AcetnSite) * 4, ctypes.c_int,
                                                    c_i64, c_i64, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_int, ctypes.c_int, c_i64, c_vp, c_vp,
                                                    c_vp, c_size, c_vp]),
        "acetn_hip_projectors_begin_mark": (ctypes.c_int, [ctypes.POINTER(AcetnSite) * 4,
                                                           ctypes.c_int, c_i64, c_i64,
                                                           ctypes.c_int, c_dbl, c_u64, c_u64,
                                                           ctypes.c_int, ctypes.c_int,
                                                           ctypes.c_int, c_vp, c_vp, c_vp,
                                                           c_size, c_vp, c_vp]),
        "acetn_hip_projectors_begin_graph": (ctypes.c_int, [ctypes.POINTER(AcetnSite) * 4,
                                                            ctypes.c_int, c_i64, c_i64,
                                                            ctypes.c_int, c_dbl, c_u64, c_u64,
                                                            c_vp, ctypes.c_int, ctypes.c_int,
                                                            ctypes.c_int, c_vp, c_vp, c_vp,
                                                            c_size, c_vp]),
        "acetn_hip_projectors_end_padded": (ctypes.c_int, [ctypes.POINTER(AcetnSite) * 4,
                                                           ctypes.c_int, c_i64, c_i64,
                                                           ctypes.c_int, ctypes.c_int,
                                                           ctypes.c_int, ctypes.c_int, c_i64, c_vp,
                                                           c_vp, c_vp, c_size, c_vp]),
        "acetn_hip_renormalize_workspace": (c_size, [ctypes.POINTER(AcetnSite), ctypes.c_int,
                                                     c_i64, c_i64]),
        "acetn_hip_renormalize_boundary": (ctypes.c_int, [ctypes.POINTER(AcetnSite), ctypes.c_int,
                                                          c_vp, c_vp, c_i64, c_vp, c_vp, c_i64,
                                                          ctypes.c_int, c_vp, c_vp, c_vp, c_vp,
                                                          c_size, c_vp]),
        "acetn_hip_renormalize_part_workspace": (c_size, [ctypes.POINTER(AcetnSite), ctypes.c_int,
                                                          ctypes.c_int, c_i64, c_i64]),
        "acetn_hip_renormalize_boundary_part": (ctypes.c_int, [ctypes.POINTER(AcetnSite), ctypes.c_int,
                                                               ctypes.c_int, ctypes.c_int, c_vp, c_vp,
                                                               c_i64, c_vp, c_vp, c_i64, c_i64,
                                                               ctypes.c_int, c_vp, c_vp, c_vp,
                                                               c_vp, c_size, c_vp]),
        "acetn_hip_site_rdm_workspace": (c_size, [ctypes.POINTER(AcetnSite)]),
        "acetn_hip_site_rdm": (ctypes.c_int, [ctypes.POINTER(AcetnSite), c_vp, c_vp, c_size, c_vp]),
        "acetn_hip_bond_rdm_workspace": (c_size, [ctypes.POINTER(AcetnSite), ctypes.POINTER(AcetnSite),
                                                  ctypes.c_int]),
        "acetn_hip_bond_rdm": (ctypes.c_int, [ctypes.POINTER(AcetnSite), ctypes.POINTER(AcetnSite),
                                              ctypes.c_int, c_vp, c_vp, c_size, c_vp]),
        "acetn_hip_gaussian_sketch": (ctypes.c_int, [c_vp, c_i64, c_i64, c_u64, c_u64, c_vp]),
        "acetn_hip_gemm_workspace": (c_size, [c_i64, c_i64, c_i64]),
        "acetn_hip_gemm": (ctypes.c_int, [c_i64, c_i64, c_i64, ctypes.c_int, c_vp, c_i64,
                                          ctypes.c_int, c_vp, c_i64, c_dbl, c_dbl, c_vp, c_i64,
                                          c_vp, c_size, c_vp]),
        "acetn_hip_gemm_split": (ctypes.c_int, [c_i64, c_i64, c_i64, ctypes.c_int, c_vp, c_i64,
                                                ctypes.c_int, c_vp, c_i64, c_dbl, c_dbl, c_vp,
                                                c_i64, ctypes.c_int, c_vp, c_size, c_vp]),
        "acetn_hip_bond_decompose_workspace": (c_size, [c_i64, c_i64]),
        "acetn_hip_bond_decompose": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int, c_i64, c_i64, c_vp,
                                                    c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
        "acetn_hip_norm_tensor_workspace": (c_size, [ctypes.POINTER(AcetnSite),
                                                     ctypes.POINTER(AcetnSite), ctypes.c_int]),
        "acetn_hip_norm_tensor": (ctypes.c_int, [ctypes.POINTER(AcetnSite),
                                                 ctypes.POINTER(AcetnSite), ctypes.c_int, c_vp,
                                                 c_vp, c_vp, c_vp, c_size, c_vp]),
        "acetn_hip_apply_gate_workspace": (c_size, [c_i64, c_i64, c_i64]),
        "acetn_hip_apply_gate": (ctypes.c_int, [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp,
                                                c_size, c_vp]),
        "acetn_hip_positive_approx_workspace": (c_size, [c_i64]),
        "acetn_hip_positive_approx": (ctypes.c_int, [c_vp, c_i64, c_dbl, c_vp,
                                                     ctypes.POINTER(ctypes.c_int32), c_vp, c_size,
                                                     c_vp]),
        "acetn_hip_gauge_fix_workspace": (c_size, [c_i64, c_i64]),
        "acetn_hip_gauge_fix": (ctypes.c_int, [c_vp, c_vp, c_i64, c_i64, c_dbl, c_vp, c_vp, c_vp,
                                               c_vp, c_vp, c_size, c_vp]),
        "acetn_hip_als_workspace": (c_size, [c_i64, c_i64, c_i64]),
        "acetn_hip_als_initialize": (ctypes.c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp,
                                                    c_size, c_vp]),
        "acetn_hip_als_batch_workspace": (c_size, [c_i64, c_i64, c_i64, c_i64]),
        "acetn_hip_als_solve_batch": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64,
                                                     c_i64, ctypes.c_int, c_dbl, ctypes.c_int,
                                                     c_dbl, ctypes.c_int, c_vp, c_vp, c_vp,
                                                     c_size, c_vp]),
        "acetn_hip_als_solve": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64,
                                               ctypes.c_int, c_dbl, ctypes.c_int, c_dbl,
                                               ctypes.POINTER(c_i64), ctypes.POINTER(c_dbl),
                                               ctypes.POINTER(ctypes.c_int32), c_vp, c_size,
                                               c_vp]),
        "acetn_hip_als_solve_noalloc": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64,
                                                       c_i64, ctypes.c_int, c_dbl, ctypes.c_int,
                                                       c_dbl,
                                                       ctypes.POINTER(c_i64), ctypes.POINTER(c_dbl),
                                                       ctypes.POINTER(ctypes.c_int32), c_vp]),
        "acetn_hip_finalize_reduced_workspace": (c_size, [c_i64, c_i64, c_i64]),
        "acetn_hip_finalize_reduced": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64,
                                                      c_vp, c_vp, c_vp, c_size, c_vp]),
        "acetn_hip_recompose_workspace": (c_size, [c_i64, c_i64]),
        "acetn_hip_recompose": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_int, ctypes.c_int,
                                               c_i64, c_i64, c_vp, c_vp, c_vp, c_size, c_vp]),
        "acetn_hip_finalize_bond_workspace": (c_size, [c_i64, c_i64]),
        "acetn_hip_finalize_bond": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_int,
                                                   c_i64, c_i64, c_vp, c_vp, c_vp, c_size, c_vp]),
        "acetn_hip_simple_update_workspace": (c_size, [c_i64, c_i64]),
        "acetn_hip_simple_update_bond": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int, c_vp * 4,
                                                        c_vp * 4, c_vp, c_i64, c_i64, c_dbl, c_vp,
                                                        c_vp, c_vp, c_vp, c_size, c_vp]),
        "acetn_hip_mlp_param_count": (c_i64, [c_i64, c_i64, c_i64]),
        "acetn_hip_mlp_workspace": (c_size, [c_i64, c_i64, c_i64, c_i64]),
        "acetn_hip_mlp_forward": (ctypes.c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_i64, c_vp,
                                                 c_vp, c_size, c_vp]),
        "acetn_hip_mlp_backward": (ctypes.c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_i64, c_vp,
                                                  c_vp, c_vp, ctypes.POINTER(ctypes.c_float),
                                                  c_vp, c_size, c_vp]),
        "acetn_hip_mlp_adam_step": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64,
                                                   ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                                   ctypes.c_float, ctypes.c_float, c_vp]),
        "acetn_hip_rsvd_product_workspace": (c_size, [ctypes.c_int, c_vp, c_i64, ctypes.c_int,
                                                      ctypes.c_int]),
        "acetn_hip_rsvd_product": (ctypes.c_int, [ctypes.c_int, c_vp, c_vp, c_i64, ctypes.c_int,
                                                  ctypes.c_int, c_u64, c_u64, c_vp, c_vp, c_vp,
                                                  c_vp, c_size, c_vp]),
        "acetn_hip_svd_workspace": (c_size, [c_i64, c_i64]),
        "acetn_hip_svd": (ctypes.c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp,
                                         ctypes.POINTER(ctypes.c_int32), c_vp, c_size, c_vp]),
        "acetn_hip_eigh_workspace": (c_size, [c_i64]),
        "acetn_hip_eigh": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_vp, c_size, c_vp]),
        "acetn_hip_orthonormalize_workspace": (c_size, [c_i64, c_i64]),
        "acetn_hip_orthonormalize": (ctypes.c_int, [c_vp, c_i64, c_i64, c_vp, c_size, c_vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def _try_load():
    global _lib, _load_error
    if _lib is not None or _load_error is not None:
        return
    if not os.path.exists(LIB_PATH):
        _load_error = f"{LIB_PATH} not built (run __graft_entry__.build())"
        return
    try:
        _lib = _declare(ctypes.CDLL(LIB_PATH))
    except OSError as e:  # pragma: no cover - depends on the runtime
        _load_error = f"failed to load {LIB_PATH}: {e}"


_try_load()
HIP_AVAILABLE = _lib is not None


def lib():
    """The loaded library; raises if it is not available."""
    if _lib is None:
        raise RuntimeError(f"acetn HIP engine unavailable: {_load_error}")
    return _lib


def require_gpu():
    """Raise unless the HIP library is loaded and a GPU is visible."""
    lib()
    if not torch.cuda.is_available():
        raise RuntimeError("acetn HIP engine needs a visible AMD GPU (torch.cuda.is_available() "
                           "is False)")


def to_engine(t, ipeps=None):
    """``t`` as a contiguous FP64 GPU tensor: a GPU tensor stays on its
    device, a host tensor goes to the network's device when ``ipeps`` is
    given, else to the current one.  For the public functions that, like the
    reference's, accept tensors on any device: the computation itself always
    runs on the GPU."""
    require_gpu()
    if t.is_complex():
        raise TypeError("acetn HIP engine: complex tensors are not supported (FP64 engine)")
    if t.device.type == "cuda":
        dev = t.device
    elif ipeps is not None:
        dev = ipeps.device
    else:
        dev = torch.device("cuda", torch.cuda.current_device())
    return t.to(device=dev, dtype=torch.float64).contiguous()


def check(status):
    if status != 0:
        msg = lib().acetn_hip_last_error().decode()
        raise RuntimeError(f"acetn HIP engine error {status}: {msg}")


def stream_handle(device=None):
    return c_vp(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    """Device address of a tensor for the C boundary.  Host tensors are
    refused: a kernel would take a host address for device memory."""
    if t is None:
        return c_vp(0)
    if t.device.type != "cuda":
        raise ValueError(f"acetn HIP engine: tensor on {t.device}, expected a GPU tensor")
    return c_vp(t.data_ptr())


class Workspace:
    """Grow-only device scratch owned by torch's caching allocator."""

    def __init__(self):
        self.buf = None

    def get(self, nbytes, device):
        nbytes = int(nbytes)
        if self.buf is None or self.buf.numel() < nbytes or self.buf.device != device:
            self.buf = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)
        return self.buf


_ws = Workspace()
_slot_ws = {}
_streams = {}


def workspace(nbytes, device):
    return _ws.get(nbytes, device)


def slot_workspace(slot, nbytes, device):
    """Per-slot scratch for work that runs concurrently on several streams."""
    w = _slot_ws.setdefault(slot, Workspace())
    return w.get(nbytes, device)


_serial = False


def serialize(flag):
    """When set, every slot runs on the caller's current stream (no overlap of
    independent projector pairs); used to time kernels without contention."""
    global _serial
    _serial = bool(flag)


N_SLOTS = 4  # concurrent projector streams (one per hardware queue)


def slot_stream(slot, device):
    """Default-priority streams, one per concurrent slot.  Slot 0 runs on the
    caller's current stream: the runtime spreads streams over a small fixed
    number of hardware queues (4 by default), and the current stream plus
    three side streams land on four distinct queues, where the current
    stream and four side streams put two slots on one queue and serialise
    them."""
    if _serial or slot == 0:
        return torch.cuda.current_stream(device)
    key = (slot, str(device))
    if key not in _streams:
        _streams[key] = torch.cuda.Stream(device=device)
    return _streams[key]


def make_site(A, C, E):
    """Build an AcetnSite view of contiguous FP64 device tensors; the returned
    tuple keeps the tensors alive.  Other dtypes are refused (the kernels read
    FP64)."""
    ts = [A, *C, *E]
    for t in ts:
        if t.dtype != torch.float64:
            raise TypeError(f"acetn HIP engine: site tensors must be float64, got {t.dtype}")
    for t in ts:
        if t.device.type != "cuda":
            raise ValueError(f"acetn HIP engine: site tensor on {t.device}, expected a GPU tensor")
    if A.dim() != 5 or len(set(A.shape[:4])) != 1:
        raise ValueError(f"acetn HIP engine: site tensor A has shape {tuple(A.shape)}, "
                         "expected (D, D, D, D, d)")
    D = A.shape[0]
    if len(C) != 4 or len(E) != 4 or any(c.dim() != 2 for c in C) or \
            any(e.dim() != 4 or tuple(e.shape[2:]) != (D, D) for e in E):
        raise ValueError("acetn HIP engine: a site needs four corners (chi, chi) and four "
                         "edges (chi, chi, D, D)")
    keep = [A.contiguous()] + [c.contiguous() for c in C] + [e.contiguous() for e in E]
    s = AcetnSite()
    s.A = keep[0].data_ptr()
    D = A.shape[0]
    s.D = D
    s.d = A.shape[4]
    for k in range(4):
        c, e = keep[1 + k], keep[5 + k]
        s.C[k] = c.data_ptr()
        s.E[k] = e.data_ptr()
        s.C_dim[k][0], s.C_dim[k][1] = c.shape[0], c.shape[1]
        s.E_dim[k][0], s.E_dim[k][1] = e.shape[0], e.shape[1]
    return s, keep


def device_info():
    buf = ctypes.create_string_buffer(512)
    check(lib().acetn_hip_device_info(buf, 512))
    return buf.value.decode()


AUTOTUNE_MODES = {"off": 0, "tune": 1, "table": 2}


def autotune(mode="tune"):
    """Tile-shape autotuner of the matrix-core products: 'off' (built-in
    thresholds), 'tune' (time the candidates of every new shape once, keep the
    fastest) or 'table' (use what was tuned or loaded, tune nothing new)."""
    check(lib().acetn_hip_autotune(AUTOTUNE_MODES[mode]))


def autotune_entries():
    n = c_i64(0)
    check(lib().acetn_hip_autotune_entries(ctypes.byref(n)))
    return n.value


def autotune_clear():
    check(lib().acetn_hip_autotune_clear())


def autotune_save(path):
    check(lib().acetn_hip_autotune_save(str(path).encode()))


def autotune_load(path):
    n = c_i64(0)
    check(lib().acetn_hip_autotune_load(str(path).encode(), ctypes.byref(n)))
    return n.value


def kernel_timing(enable):
    check(lib().acetn_hip_kernel_timing(int(enable)))


# kernel classes of acetn_hip_kernel_timing_read
K_GEMM_F64, K_GEMM_SPLIT, K_DOUBLE_LAYER, K_JACOBI, K_CHOLESKY = 0, 1, 2, 3, 4
# precision modes (ACETN_HIP_FP64 / ACETN_HIP_MIXED)
PRECISIONS = {"fp64": 0, "mixed": 1}
# OR-ed into a projector pair's precision code: the range finder refills
# dropped columns at random (cheaper; identical unless the spectrum is graded)
RF_FAST = 0x100


def precision_code(precision):
    try:
        return PRECISIONS[precision]
    except KeyError:
        raise ValueError(f"unknown precision {precision!r}; expected one of {sorted(PRECISIONS)}")


def kernel_timing_read(kernel=K_GEMM_F64):
    """(summed kernel ms, summed algorithmic flops, launches) of one kernel class."""
    ms, fl, n = c_dbl(0), c_dbl(0), c_i64(0)
    check(lib().acetn_hip_kernel_timing_read(kernel, ctypes.byref(ms), ctypes.byref(fl),
                                             ctypes.byref(n)))
    return ms.value, fl.value, n.value


def kernel_timing_bytes(kernel=K_GEMM_F64):
    """Summed algorithmic HBM bytes of one kernel class."""
    b = c_dbl(0)
    check(lib().acetn_hip_kernel_timing_bytes(kernel, ctypes.byref(b)))
    return b.value


# ---------------------------------------------------------------- wrappers
def gaussian_sketch(n, q, seed, stream_id, device="cuda"):
    out = torch.empty(n, q, dtype=torch.float64, device=device)
    check(lib().acetn_hip_gaussian_sketch(ptr(out), n, q, seed, stream_id, stream_handle()))
    return out


def gemm(A, B, ta=False, tb=False, alpha=1.0, beta=0.0, C=None, pieces=0):
    """Dense GEMM on the matrix cores: alpha*op(A) op(B) + beta*C, FP64
    (pieces=0) or split-bf16 with FP32 accumulation (pieces=2: bf16x3,
    pieces=3: bf16x6); FP64 storage either way."""
    if A.dim() != 2 or B.dim() != 2:
        raise ValueError("gemm: A and B must be matrices")
    A = A.contiguous()
    B = B.contiguous()
    M = A.shape[1] if ta else A.shape[0]
    K = A.shape[0] if ta else A.shape[1]
    N = B.shape[0] if tb else B.shape[1]
    _expect("B", B, (N, K) if tb else (K, N))
    _expect("A", A, (K, M) if ta else (M, K))
    if C is None:
        C = torch.zeros(M, N, dtype=torch.float64, device=A.device)
    else:
        _expect("C", C, (M, N))
        if not C.is_contiguous():
            raise ValueError("gemm: C must be contiguous")
    nb = lib().acetn_hip_gemm_workspace(M, N, K)
    ws = workspace(nb, A.device)
    if pieces:
        check(lib().acetn_hip_gemm_split(M, N, K, int(ta), ptr(A), A.shape[1], int(tb), ptr(B),
                                         B.shape[1], alpha, beta, ptr(C), C.shape[1], pieces,
                                         ptr(ws), ws.numel(), stream_handle()))
        return C
    check(lib().acetn_hip_gemm(M, N, K, int(ta), ptr(A), A.shape[1], int(tb), ptr(B), B.shape[1],
                               alpha, beta, ptr(C), C.shape[1], ptr(ws), ws.numel(), stream_handle()))
    return C


def eigh(G):
    """Descending eigen-decomposition of a symmetric FP64 matrix."""
    q = G.shape[0]
    G = G.clone().contiguous()
    w = torch.empty(q, dtype=torch.float64, device=G.device)
    V = torch.empty(q, q, dtype=torch.float64, device=G.device)
    nb = lib().acetn_hip_eigh_workspace(q)
    ws = workspace(nb, G.device)
    check(lib().acetn_hip_eigh(ptr(G), q, ptr(w), ptr(V), ptr(ws), ws.numel(), stream_handle()))
    return w, V


def rsvd_product(ops, q, niter, mid_qr, seed, stream_id):
    """Randomized SVD of ops[0] @ ops[1] @ ... (device FP64 matrices) without
    forming the product (acetn_hip_rsvd_product): (U, S, V) with q clamped to
    min(rows, cols)."""
    if not ops:
        raise ValueError("rsvd_product: no factor")
    for t in ops:
        if t.dtype != torch.float64:
            raise TypeError(f"rsvd_product: float64 matrices only, got {t.dtype}")
        if t.dim() != 2:
            raise ValueError("rsvd_product: factors must be matrices")
    for a, b in zip(ops[:-1], ops[1:]):
        if a.shape[1] != b.shape[0]:
            raise ValueError(f"rsvd_product: inner dimensions {tuple(a.shape)} x {tuple(b.shape)}")
    ops = [t.contiguous() for t in ops]
    for t in ops:
        if t.device.type != "cuda":
            raise ValueError(f"rsvd_product: tensor on {t.device}, expected a GPU tensor")
    dims = [ops[0].shape[0]] + [t.shape[1] for t in ops]
    m, n = dims[0], dims[-1]
    q = min(int(q), m, n)
    dev = ops[0].device
    darr = (c_i64 * len(dims))(*dims)
    parr = (c_vp * len(ops))(*[t.data_ptr() for t in ops])
    U = torch.empty(m, q, dtype=torch.float64, device=dev)
    S = torch.empty(q, dtype=torch.float64, device=dev)
    V = torch.empty(n, q, dtype=torch.float64, device=dev)
    nb = lib().acetn_hip_rsvd_product_workspace(len(ops), darr, q, int(niter), int(mid_qr))
    if nb == 0:
        check(1)
    ws = workspace(nb, dev)
    check(lib().acetn_hip_rsvd_product(len(ops), parr, darr, q, int(niter), int(mid_qr),
                                       int(seed), int(stream_id), ptr(U), ptr(S), ptr(V), ptr(ws),
                                       ws.numel(), stream_handle()))
    return U, S, V


def svd(M):
    """One-sided Jacobi SVD of a device FP64 matrix M (rows x cols):
    (s, U, sweeps) with the min(rows, cols) leading singular values and the
    left singular vectors as U's columns (rows x rows); sweeps < 0 means the
    sweep limit stopped it."""
    rows, cols = M.shape
    Mt = M.t().contiguous()  # column j of M contiguous
    s = torch.empty(rows, dtype=torch.float64, device=M.device)
    U = torch.empty(rows, rows, dtype=torch.float64, device=M.device)
    sweeps = ctypes.c_int32(0)
    nb = lib().acetn_hip_svd_workspace(rows, cols)
    if nb == 0:
        raise ValueError(f"svd: unsupported shape {tuple(M.shape)}")
    ws = workspace(nb, M.device)
    check(lib().acetn_hip_svd(ptr(Mt), rows, cols, rows, ptr(s), ptr(U), ctypes.byref(sweeps),
                              ptr(ws), ws.numel(), stream_handle()))
    k = min(rows, cols)
    return s[:k], U, sweeps.value


def orthonormalize(Y):
    Y = Y.clone().contiguous()
    m, q = Y.shape
    nb = lib().acetn_hip_orthonormalize_workspace(m, q)
    ws = workspace(nb, Y.device)
    check(lib().acetn_hip_orthonormalize(ptr(Y), m, q, ptr(ws), ws.numel(), stream_handle()))
    return Y


def quarter_tensor(site_tensors, k, precision="fp64"):
    """site_tensors: (A, [C0..C3], [E0..E3]) device tensors."""
    s, keep = make_site(*site_tensors)
    N1 = s.E_dim[k % 4][1] * s.D * s.D
    N2 = s.E_dim[(3 + k) % 4][0] * s.D * s.D
    Q = torch.empty(N1, N2, dtype=torch.float64, device=keep[0].device)
    nb = lib().acetn_hip_quarter_tensor_workspace(ctypes.byref(s), k)
    ws = workspace(nb, Q.device)
    check(lib().acetn_hip_quarter_tensor(ctypes.byref(s), k, precision_code(precision), ptr(Q),
                                         ptr(ws), ws.numel(), stream_handle()))
    return Q


def half_projectors(site1, site4, k, chi, oversampling, niter, cutoff, seed, stream_id,
                    precision="fp64"):
    """Returns (P1, P2, spectrum) with P1/P2 shaped (rows, cD_new)."""
    s1, keep1 = make_site(*site1)
    s4, keep4 = make_site(*site4)
    dev = keep1[0].device
    rows = s1.E_dim[(3 + k) % 4][0] * s1.D * s1.D
    P1 = torch.empty(rows * chi, dtype=torch.float64, device=dev)
    P2 = torch.empty(rows * chi, dtype=torch.float64, device=dev)
    spec = torch.zeros(chi + oversampling, dtype=torch.float64, device=dev)
    cD = c_i64(0)
    nb = lib().acetn_hip_half_projectors_workspace(ctypes.byref(s1), ctypes.byref(s4), k, chi,
                                                   oversampling)
    ws = workspace(nb, dev)
    check(lib().acetn_hip_half_projectors(ctypes.byref(s1), ctypes.byref(s4), k, chi, oversampling,
                                          niter, cutoff, seed, stream_id,
                                          precision_code(precision), ptr(P1), ptr(P2),
                                          ctypes.byref(cD), ptr(spec), ptr(ws), ws.numel(),
                                          stream_handle()))
    n = cD.value
    return P1[:rows * n].view(rows, n), P2[:rows * n].view(rows, n), spec


SYSTEMS = {"half-system": 0, "full-system": 1}
SVD_TYPES = {"rsvd": 0, "full-rank": 1}


class ProjectorJob:
    """One projector pair computed in two phases (acetn_hip_projectors_begin
    / _end): ``sites`` are the four plaquette sites (A, C, E) in s1..s4
    order; system 'half-system' | 'full-system', svd 'rsvd' | 'full-rank'."""

    def __init__(self, slot, sites, k, chi, oversampling, niter, cutoff, seed, stream_id,
                 precision="fp64", system="half-system", svd="rsvd", ws_key=None, graph=None,
                 mark=False, rf_fast=False):
        """With ``graph`` = (stream_base, counts, index) the pair is enqueued in
        the capturable form (acetn_hip_projectors_begin_graph / _end_padded):
        the sketch stream is stream_id + stream_base[0] (an int64 device
        tensor), the truncation count goes to counts[index] (an int32 device
        tensor), nothing synchronises, the workspace is private to the job,
        and ``finish`` returns min(chi, rows) columns with those past the
        count zero.  With ``mark`` the event ``self.mark`` is recorded on the
        pair's stream when its range-finder products are enqueued, ahead of
        the eigensolver (acetn_hip_projectors_begin_mark).  ``rf_fast`` sets
        RF_FAST (for pairs whose previous spectrum was not graded); the
        ungraded test reads the spectrum: ``self.spec_host`` receives its q
        computed values on the host by the time ``ready`` completes."""
        made = [make_site(*st) for st in sites]
        self.sv = [m[0] for m in made]
        self.keep = [m[1] for m in made]
        self.arr = (ctypes.POINTER(AcetnSite) * 4)(*[ctypes.pointer(x) for x in self.sv])
        self.k, self.chi, self.p, self.niter = k, chi, oversampling, niter
        self.prec = precision_code(precision) | (RF_FAST if rf_fast else 0)
        self.system, self.svd = SYSTEMS[system], SVD_TYPES[svd]
        self.spec_host = None
        s1 = self.sv[0]
        self.dev = self.keep[0][0].device
        self.rows = s1.E_dim[(3 + k) % 4][0] * s1.D * s1.D
        self.slot = slot
        self.stream = slot_stream(slot, self.dev)
        nb = lib().acetn_hip_projectors_workspace(self.arr, k, chi, oversampling, self.system,
                                                  self.svd)
        self.graph = graph
        nspec = max(chi + oversampling, 512) if self.svd else chi + oversampling
        if graph is not None:
            base, counts, index = graph
            if not (0 <= index < counts.numel()):
                raise ValueError("ProjectorJob: count index out of range")
            self.cols = min(chi, self.rows)
            with torch.cuda.stream(self.stream):
                self.ws = torch.empty(max(int(nb), 256), dtype=torch.uint8, device=self.dev)
                self.spec = torch.zeros(nspec, dtype=torch.float64, device=self.dev)
                check(lib().acetn_hip_projectors_begin_graph(
                    self.arr, k, chi, oversampling, niter, cutoff, seed, stream_id,
                    ptr(base), self.system, self.svd, self.prec,
                    c_vp(counts.data_ptr() + 4 * index), ptr(self.spec), ptr(self.ws),
                    self.ws.numel(), c_vp(self.stream.cuda_stream)))
            self.ready = None
            return
        # pairs queued behind each other on one slot keep their own scratch
        # until their second phase has run
        self.ws = slot_workspace(slot if ws_key is None else ("pair", ws_key), nb, self.dev)
        self.cd_host = torch.zeros(2, dtype=torch.int32, pin_memory=True)
        self.mark = None
        with torch.cuda.stream(self.stream):
            # zero-filled on the pair's stream, ahead of the kernels that write it
            self.spec = torch.zeros(nspec, dtype=torch.float64, device=self.dev)
            if mark:
                # the event object exists once recorded; the library records it
                # again at the mark, and waits issued later refer to that record
                self.mark = torch.cuda.Event()
                self.mark.record(self.stream)
                check(lib().acetn_hip_projectors_begin_mark(
                    self.arr, k, chi, oversampling, niter, cutoff, seed, stream_id, self.system,
                    self.svd, self.prec, c_vp(self.cd_host.data_ptr()), ptr(self.spec),
                    ptr(self.ws), self.ws.numel(), c_vp(self.stream.cuda_stream),
                    c_vp(self.mark.cuda_event)))
            else:
                check(lib().acetn_hip_projectors_begin(
                    self.arr, k, chi, oversampling, niter, cutoff, seed, stream_id, self.system,
                    self.svd, self.prec, c_vp(self.cd_host.data_ptr()), ptr(self.spec),
                    ptr(self.ws), self.ws.numel(), c_vp(self.stream.cuda_stream)))
            if self.svd == SVD_TYPES["rsvd"]:
                q = chi + oversampling
                self.spec_host = torch.empty(q, dtype=torch.float64, pin_memory=True)
                self.spec_host.copy_(self.spec[:q], non_blocking=True)
            # marks the truncation count's (and spectrum's) arrival in pinned
            # host memory
            self.ready = torch.cuda.Event()
            self.ready.record(self.stream)

    def finish(self):
        """Call once ``ready`` has completed (or the stream was synchronised);
        in the capturable form at once."""
        if self.graph is not None:
            with torch.cuda.stream(self.stream):
                P1 = torch.empty(self.rows, self.cols, dtype=torch.float64, device=self.dev)
                P2 = torch.empty(self.rows, self.cols, dtype=torch.float64, device=self.dev)
                check(lib().acetn_hip_projectors_end_padded(
                    self.arr, self.k, self.chi, self.p, self.niter, self.system, self.svd,
                    self.prec, self.cols, ptr(P1), ptr(P2), ptr(self.ws), self.ws.numel(),
                    c_vp(self.stream.cuda_stream)))
            return P1, P2
        n = int(self.cd_host[0])
        with torch.cuda.stream(self.stream):
            P1 = torch.empty(self.rows, n, dtype=torch.float64, device=self.dev)
            P2 = torch.empty(self.rows, n, dtype=torch.float64, device=self.dev)
            check(lib().acetn_hip_projectors_end(
                self.arr, self.k, self.chi, self.p, self.niter, self.system, self.svd, self.prec,
                n, ptr(P1), ptr(P2), ptr(self.ws), self.ws.numel(), c_vp(self.stream.cuda_stream)))
        return P1, P2


def renormalize_boundary(site1, k, P1i, P2i, P1j, P2j, slot=None, precision="fp64",
                         ws_key=None):
    """Returns (C_a, C_b, E_o): the new C[(3+k)%4], C[k], E[(3+k)%4] of s2.
    ``slot`` (or ``ws_key``, when several calls share a slot) selects a
    workspace of its own for concurrent calls."""
    s1, keep = make_site(*site1)
    dev = keep[0].device
    D = s1.D
    cDi, cDj = P1i.shape[1], P1j.shape[1]
    xa = s1.E_dim[(2 + k) % 4][0]
    xc = s1.E_dim[k % 4][1]
    Ca = torch.empty(xa, cDi, dtype=torch.float64, device=dev)
    Cb = torch.empty(cDj, xc, dtype=torch.float64, device=dev)
    Eo = torch.empty(cDi, cDj, D, D, dtype=torch.float64, device=dev)
    P1i, P2i, P1j, P2j = (t.contiguous() for t in (P1i, P2i, P1j, P2j))
    nb = lib().acetn_hip_renormalize_workspace(ctypes.byref(s1), k, cDi, cDj)
    # own per-slot scratch: a renormalisation may run while other slots'
    # projector pairs still use theirs
    key = slot if ws_key is None else ws_key
    ws = workspace(nb, dev) if key is None else slot_workspace(("renormalize", key), nb, dev)
    check(lib().acetn_hip_renormalize_boundary(ctypes.byref(s1), k, ptr(P1i), ptr(P2i), cDi,
                                               ptr(P1j), ptr(P2j), cDj, precision_code(precision),
                                               ptr(Ca), ptr(Cb), ptr(Eo),
                                               ptr(ws), ws.numel(), stream_handle()))
    return Ca, Cb, Eo


class RenormalizeParts:
    """renormalize_boundary in two calls on the current stream: ``first(P1, P2)``
    with the projectors of the pair named by ``first_pair`` ('j' or 'i'),
    then ``second(P1, P2)`` with the other pair's; ``second`` returns
    (C_a, C_b, E_o).  ``cd_bound`` >= the second pair's count (chi) fixes the
    shared workspace layout.  With a ``slot`` or ``ws_key`` the workspace is
    that key's, which no other call may use between the two parts."""

    def __init__(self, site1, k, cd_bound, first_pair="j", slot=None, precision="fp64",
                 ws_key=None):
        if first_pair not in ("i", "j"):
            raise ValueError("first_pair must be 'i' or 'j'")
        self.s1, self.keep = make_site(*site1)
        self.k, self.bound = k, int(cd_bound)
        self.first_code = 0 if first_pair == "j" else 1
        self.prec = precision_code(precision)
        self.key = slot if ws_key is None else ws_key
        self.dev = self.keep[0].device

    def _call(self, part, P1i, P2i, cDi, P1j, P2j, cDj, Ca, Cb, Eo):
        check(lib().acetn_hip_renormalize_boundary_part(
            ctypes.byref(self.s1), self.k, part, self.first_code, ptr(P1i), ptr(P2i), cDi,
            ptr(P1j), ptr(P2j), cDj, self.bound, self.prec, ptr(Ca), ptr(Cb), ptr(Eo),
            ptr(self.ws), self.ws.numel(), stream_handle()))

    def first(self, P1, P2):
        s1, k = self.s1, self.k
        self.P1f, self.P2f = P1.contiguous(), P2.contiguous()
        cf = self.P1f.shape[1]
        cDi, cDj = (self.bound, cf) if self.first_code == 0 else (cf, self.bound)
        nb = lib().acetn_hip_renormalize_part_workspace(ctypes.byref(s1), k, self.first_code,
                                                        cDi, cDj)
        # the two calls share what part 1 leaves in the workspace: without a
        # slot or key it is private to this object (the shared scratch could
        # be reused by other calls in between)
        self.ws = (torch.empty(max(int(nb), 256), dtype=torch.uint8, device=self.dev)
                   if self.key is None
                   else slot_workspace(("renormalize", self.key), nb, self.dev))
        if self.first_code == 0:
            xc = s1.E_dim[k % 4][1]
            self.Cf = torch.empty(cf, xc, dtype=torch.float64, device=self.dev)
            self._call(1, None, None, 0, self.P1f, self.P2f, cf, None, self.Cf, None)
        else:
            xa = s1.E_dim[(2 + k) % 4][0]
            self.Cf = torch.empty(xa, cf, dtype=torch.float64, device=self.dev)
            self._call(1, self.P1f, self.P2f, cf, None, None, 0, self.Cf, None, None)
        return self.Cf

    def second(self, P1, P2):
        s1, k, D = self.s1, self.k, self.s1.D
        P1, P2 = P1.contiguous(), P2.contiguous()
        cs = P1.shape[1]
        if cs > self.bound:
            raise ValueError("renormalize: the second pair's count exceeds the layout bound")
        if self.first_code == 0:
            P1i, P2i, P1j, P2j, cDi, cDj = P1, P2, self.P1f, self.P2f, cs, self.P1f.shape[1]
            Cb = self.Cf
            Ca = torch.empty(s1.E_dim[(2 + k) % 4][0], cDi, dtype=torch.float64, device=self.dev)
        else:
            P1i, P2i, P1j, P2j, cDi, cDj = self.P1f, self.P2f, P1, P2, self.P1f.shape[1], cs
            Ca = self.Cf
            Cb = torch.empty(cDj, s1.E_dim[k % 4][1], dtype=torch.float64, device=self.dev)
        Eo = torch.empty(cDi, cDj, D, D, dtype=torch.float64, device=self.dev)
        self._call(2, P1i, P2i, cDi, P1j, P2j, cDj, Ca, Cb, Eo)
        return Ca, Cb, Eo


def site_rdm(site):
    """rho[P][p] of one site; site = (A, [C0..C3], [E0..E3])."""
    s, keep = make_site(*site)
    dev = keep[0].device
    rho = torch.empty(s.d, s.d, dtype=torch.float64, device=dev)
    nb = lib().acetn_hip_site_rdm_workspace(ctypes.byref(s))
    ws = workspace(nb, dev)
    check(lib().acetn_hip_site_rdm(ctypes.byref(s), ptr(rho), ptr(ws), ws.numel(), stream_handle()))
    return rho


def bond_rdm(site1, site2, k):
    """rho[P][Q][p][q] of the bond (s1, s2, k)."""
    s1, keep1 = make_site(*site1)
    s2, keep2 = make_site(*site2)
    dev = keep1[0].device
    d = s1.d
    rho = torch.empty(d, d, d, d, dtype=torch.float64, device=dev)
    nb = lib().acetn_hip_bond_rdm_workspace(ctypes.byref(s1), ctypes.byref(s2), k)
    ws = workspace(nb, dev)
    check(lib().acetn_hip_bond_rdm(ctypes.byref(s1), ctypes.byref(s2), k, ptr(rho), ptr(ws),
                                   ws.numel(), stream_handle()))
    return rho


# ------------------------------------------------------------ full update
ALS_METHODS = {"cholesky": 0, "pinv": 1}


def _f64(*shape, device):
    return torch.empty(*shape, dtype=torch.float64, device=device)


def bond_decompose(A1, A2, k):
    """(a1q, a1r, a2q, a2r) of the bond frame k (tensor_update.py:52-68)."""
    if A1.dim() != 5:
        raise ValueError("acetn HIP engine: site tensors are (D, D, D, D, d)")
    D, d = A1.shape[0], A1.shape[4]
    _expect("A1", A1, (D, D, D, D, d))
    _expect("A2", A2, (D, D, D, D, d))
    nD = D * d
    dev = A1.device
    A1, A2 = A1.contiguous(), A2.contiguous()
    a1q, a2q = _f64(D, D, D, nD, device=dev), _f64(D, D, D, nD, device=dev)
    a1r, a2r = _f64(nD, D, d, device=dev), _f64(nD, D, d, device=dev)
    ws = workspace(lib().acetn_hip_bond_decompose_workspace(D, d), dev)
    check(lib().acetn_hip_bond_decompose(ptr(A1), ptr(A2), k, D, d, ptr(a1q), ptr(a1r), ptr(a2q),
                                         ptr(a2r), ptr(ws), ws.numel(), stream_handle()))
    return a1q, a1r, a2q, a2r


def norm_tensor(site1, site2, k, a1q, a2q):
    """n12[y][x][Y][X] (full_update.py:163-227); site = (A, C, E)."""
    s1, keep1 = make_site(*site1)
    s2, keep2 = make_site(*site2)
    D, d = s1.D, s1.d
    if (s2.D, s2.d) != (D, d):
        raise ValueError("norm_tensor: the two sites' (D, d) differ")
    nD = D * d
    _expect("a1q", a1q, (D, D, D, nD))
    _expect("a2q", a2q, (D, D, D, nD))
    dev = a1q.device
    n12 = _f64(nD, nD, nD, nD, device=dev)
    ws = workspace(lib().acetn_hip_norm_tensor_workspace(ctypes.byref(s1), ctypes.byref(s2), k), dev)
    a1q, a2q = a1q.contiguous(), a2q.contiguous()  # held until the call has enqueued
    check(lib().acetn_hip_norm_tensor(ctypes.byref(s1), ctypes.byref(s2), k, ptr(a1q),
                                      ptr(a2q), ptr(n12), ptr(ws), ws.numel(),
                                      stream_handle()))
    return n12


def _expect(name, t, shape):
    """Shape and dtype check of a bond-solver operand before its address
    reaches a kernel (which reads exactly ``shape`` FP64 elements)."""
    if t.dtype != torch.float64:
        raise TypeError(f"acetn HIP engine: {name} must be float64, got {t.dtype}")
    if tuple(t.shape) != tuple(int(x) for x in shape):
        raise ValueError(f"acetn HIP engine: {name} has shape {tuple(t.shape)}, "
                         f"expected {tuple(int(x) for x in shape)}")


def apply_gate(a1r, a2r, gate):
    """a12g (nD, nD, d, d) of full_update.py:77-86 from a1r, a2r (nD, D, d)."""
    nD, D, d = a1r.shape
    _expect("a2r", a2r, (nD, D, d))
    _expect("gate", gate, (d, d, d, d))
    a12g = _f64(nD, nD, d, d, device=a1r.device)
    ws = workspace(lib().acetn_hip_apply_gate_workspace(nD, D, d), a1r.device)
    a1r, a2r, gate = a1r.contiguous(), a2r.contiguous(), gate.contiguous()
    check(lib().acetn_hip_apply_gate(ptr(a1r), ptr(a2r), ptr(gate), nD, D, d, ptr(a12g), ptr(ws),
                                     ws.numel(), stream_handle()))
    return a12g


def positive_approx(n12, cutoff):
    """nz (nD, nD, nD^2) of full_update.py:262-297 and the number of diagonal
    shifts applied."""
    nD = n12.shape[0]
    _expect("n12", n12, (nD, nD, nD, nD))
    nz = _f64(nD, nD, nD * nD, device=n12.device)
    shifts = ctypes.c_int32(0)
    ws = workspace(lib().acetn_hip_positive_approx_workspace(nD), n12.device)
    n12 = n12.contiguous()
    check(lib().acetn_hip_positive_approx(ptr(n12), nD, float(cutoff), ptr(nz),
                                          ctypes.byref(shifts), ptr(ws), ws.numel(),
                                          stream_handle()))
    return nz, shifts.value


def gauge_fix(nz, a12g, atol):
    """(n12, a12g, nzxr_inv, nzyr_inv) of full_update.py:300-343."""
    nD = nz.shape[0]
    _expect("nz", nz, (nD, nD, nD * nD))
    d = a12g.shape[2] if a12g.dim() == 4 else 0
    _expect("a12g", a12g, (nD, nD, d, d))
    dev = nz.device
    n12 = _f64(nD, nD, nD, nD, device=dev)
    a12g_out = _f64(nD, nD, d, d, device=dev)
    nzxr_inv, nzyr_inv = _f64(nD, nD, device=dev), _f64(nD, nD, device=dev)
    ws = workspace(lib().acetn_hip_gauge_fix_workspace(nD, d), dev)
    nz, a12g = nz.contiguous(), a12g.contiguous()
    check(lib().acetn_hip_gauge_fix(ptr(nz), ptr(a12g), nD, d, float(atol),
                                    ptr(n12), ptr(a12g_out), ptr(nzxr_inv), ptr(nzyr_inv), ptr(ws),
                                    ws.numel(), stream_handle()))
    return n12, a12g_out, nzxr_inv, nzyr_inv


def als_initialize(a12g, D):
    """The plain ALS start (als_solver.py:117-146): (a1r, a2r), each (nD, D, d)."""
    nD = a12g.shape[0]
    d = a12g.shape[2] if a12g.dim() == 4 else 0
    _expect("a12g", a12g, (nD, nD, d, d))
    dev = a12g.device
    a1r, a2r = _f64(nD, D, d, device=dev), _f64(nD, D, d, device=dev)
    ws = workspace(lib().acetn_hip_als_workspace(nD, D, d), dev)
    a12g = a12g.contiguous()
    check(lib().acetn_hip_als_initialize(ptr(a12g), nD, D, d, ptr(a1r), ptr(a2r), ptr(ws),
                                         ws.numel(), stream_handle()))
    return a1r, a2r


def _als_operands(n12, a12g, a1r, a2r):
    nD, D, d = a1r.shape
    _expect("a1r", a1r, (nD, D, d))
    _expect("a2r", a2r, (nD, D, d))
    _expect("n12", n12, (nD, nD, nD, nD))
    _expect("a12g", a12g, (nD, nD, d, d))
    return nD, D, d


def als_solve(n12, a12g, a1r, a2r, niter, tol, method, epsilon):
    """Device ALS from the start (a1r, a2r); returns (a1r, a2r, sweeps, cost, flags)."""
    nD, D, d = _als_operands(n12, a12g, a1r, a2r)
    dev = a1r.device
    if method not in ALS_METHODS:
        raise ValueError(f"Unknown ALS method: {method}")
    b1, b2 = a1r.clone().contiguous(), a2r.clone().contiguous()
    sweeps, cost, info = c_i64(0), c_dbl(0.0), ctypes.c_int32(0)
    ws = workspace(lib().acetn_hip_als_workspace(nD, D, d), dev)
    n12, a12g = n12.contiguous(), a12g.contiguous()
    check(lib().acetn_hip_als_solve(ptr(n12), ptr(a12g), ptr(b1), ptr(b2), nD, D, d, int(niter),
                                    float(tol), ALS_METHODS[method], float(epsilon),
                                    ctypes.byref(sweeps), ctypes.byref(cost), ctypes.byref(info),
                                    ptr(ws), ws.numel(), stream_handle()))
    return b1, b2, sweeps.value, cost.value, info.value


def als_solve_batch(n12, a12g, D, niter, tol, method, epsilon, starts=None):
    """``count`` ALS problems in one launch (acetn_hip_als_solve_batch):
    n12 (count, nD, nD, nD, nD), a12g (count, nD, nD, d, d) on the device;
    ``starts`` = (a1r, a2r), each (count, nD, D, d), or None for the plain
    start.  Returns (a1r, a2r, sweeps, cost, flags) as device tensors; no
    synchronisation."""
    if method not in ALS_METHODS:
        raise ValueError(f"Unknown ALS method: {method}")
    if a12g.dim() != 5:
        raise ValueError("als_solve_batch: a12g must be (count, nD, nD, d, d)")
    count, nD, d = int(a12g.shape[0]), int(a12g.shape[1]), int(a12g.shape[3])
    _expect("a12g", a12g, (count, nD, nD, d, d))
    _expect("n12", n12, (count, nD, nD, nD, nD))
    dev = a12g.device
    n12, a12g = n12.contiguous(), a12g.contiguous()
    if starts is None:
        a1 = torch.empty(count, nD, D, d, dtype=torch.float64, device=dev)
        a2 = torch.empty_like(a1)
    else:
        a1, a2 = (t.clone().contiguous() for t in starts)
        _expect("a1r", a1, (count, nD, D, d))
        _expect("a2r", a2, (count, nD, D, d))
    cost = torch.empty(count, dtype=torch.float64, device=dev)
    info = torch.empty(count, 2, dtype=torch.int32, device=dev)
    ws = workspace(lib().acetn_hip_als_batch_workspace(count, nD, D, d), dev)
    check(lib().acetn_hip_als_solve_batch(ptr(n12), ptr(a12g), ptr(a1), ptr(a2), count, nD, D, d,
                                          int(niter), float(tol), ALS_METHODS[method],
                                          float(epsilon), 1 if starts is None else 0, ptr(cost),
                                          ptr(info), ptr(ws), ws.numel(), stream_handle()))
    return a1, a2, info[:, 0], cost, info[:, 1]


def als_solve_noalloc(n12, a12g, a1r, a2r, niter, tol, method, epsilon):
    """als_solve through the entry point that allocates nothing (its scratch
    is a fixed device array of the library, nD * D <= 128); a1r/a2r are
    updated in place.  Returns (sweeps, cost, flags)."""
    nD, D, d = _als_operands(n12, a12g, a1r, a2r)
    if method not in ALS_METHODS:
        raise ValueError(f"Unknown ALS method: {method}")
    for t in (n12, a12g, a1r, a2r):
        if not t.is_contiguous():
            raise ValueError("als_solve_noalloc: tensors must be contiguous")
    sweeps, cost, info = c_i64(0), c_dbl(0.0), ctypes.c_int32(0)
    check(lib().acetn_hip_als_solve_noalloc(ptr(n12), ptr(a12g), ptr(a1r), ptr(a2r), nD, D, d,
                                            int(niter), float(tol), ALS_METHODS[method],
                                            float(epsilon), ctypes.byref(sweeps),
                                            ctypes.byref(cost), ctypes.byref(info),
                                            stream_handle()))
    return sweeps.value, cost.value, info.value


def finalize_bond(a1r, a2r, nzyr_inv, nzxr_inv, a1q, a2q, k):
    """Updated site tensors (site frame, unit norm) of full_update.py:123-161
    + tensor_update.py:26-38,70-75."""
    nD, D, d = a1r.shape
    if nD != D * d:
        raise ValueError(f"finalize_bond: reduced dimension {nD} != D d = {D * d}")
    _expect("a2r", a2r, (nD, D, d))
    _expect("a1q", a1q, (D, D, D, nD))
    _expect("a2q", a2q, (D, D, D, nD))
    for name, g in (("nzyr_inv", nzyr_inv), ("nzxr_inv", nzxr_inv)):
        if g is not None:
            _expect(name, g, (nD, nD))
    dev = a1r.device
    A1, A2 = _f64(D, D, D, D, d, device=dev), _f64(D, D, D, D, d, device=dev)
    ws = workspace(lib().acetn_hip_finalize_bond_workspace(D, d), dev)
    nzyr_inv = nzyr_inv.contiguous() if nzyr_inv is not None else None
    nzxr_inv = nzxr_inv.contiguous() if nzxr_inv is not None else None
    a1r, a2r, a1q, a2q = (t.contiguous() for t in (a1r, a2r, a1q, a2q))
    gy = ptr(nzyr_inv) if nzyr_inv is not None else None
    gx = ptr(nzxr_inv) if nzxr_inv is not None else None
    check(lib().acetn_hip_finalize_bond(ptr(a1r), ptr(a2r), gy, gx,
                                        ptr(a1q), ptr(a2q), k, D, d,
                                        ptr(A1), ptr(A2), ptr(ws), ws.numel(), stream_handle()))
    return A1, A2


def finalize_reduced(a1r, a2r, nzyr_inv, nzxr_inv):
    """finalize_reduced_tensors (full_update.py:123-161): the rebalanced
    (a1r, a2r), each (nD, D, d); the gauge inverses (nD, nD) or None."""
    nD, D, d = a1r.shape
    _expect("a2r", a2r, (nD, D, d))
    if (nzyr_inv is None) != (nzxr_inv is None):
        raise ValueError("finalize_reduced: give both gauge inverses or neither")
    for name, g in (("nzyr_inv", nzyr_inv), ("nzxr_inv", nzxr_inv)):
        if g is not None:
            _expect(name, g, (nD, nD))
    dev = a1r.device
    b1, b2 = _f64(nD, D, d, device=dev), _f64(nD, D, d, device=dev)
    ws = workspace(lib().acetn_hip_finalize_reduced_workspace(nD, D, d), dev)
    a1r, a2r = a1r.contiguous(), a2r.contiguous()
    gy = nzyr_inv.contiguous() if nzyr_inv is not None else None
    gx = nzxr_inv.contiguous() if nzxr_inv is not None else None
    check(lib().acetn_hip_finalize_reduced(ptr(a1r), ptr(a2r), ptr(gy), ptr(gx), nD, D, d,
                                           ptr(b1), ptr(b2), ptr(ws), ws.numel(),
                                           stream_handle()))
    return b1, b2


def recompose(a1r, a2r, a1q, a2q, k=0, normalize=False):
    """recompose_site_tensors (tensor_update.py:70-75): (a1, a2), each
    (D, D, D, D, d), rotated from the bond frame by k (k = 0: the bond frame),
    each divided by its norm when ``normalize``."""
    nD, D, d = a1r.shape
    if nD != D * d:
        raise ValueError(f"recompose: reduced dimension {nD} != D d = {D * d}")
    _expect("a2r", a2r, (nD, D, d))
    _expect("a1q", a1q, (D, D, D, nD))
    _expect("a2q", a2q, (D, D, D, nD))
    dev = a1r.device
    A1, A2 = _f64(D, D, D, D, d, device=dev), _f64(D, D, D, D, d, device=dev)
    ws = workspace(lib().acetn_hip_recompose_workspace(D, d), dev)
    a1r, a2r, a1q, a2q = (t.contiguous() for t in (a1r, a2r, a1q, a2q))
    check(lib().acetn_hip_recompose(ptr(a1r), ptr(a2r), ptr(a1q), ptr(a2q), int(k) & 3,
                                    1 if normalize else 0, D, d, ptr(A1), ptr(A2), ptr(ws),
                                    ws.numel(), stream_handle()))
    return A1, A2


def simple_update_bond(A1, A2, k, lam1, lam2, gate, cutoff=1e-12):
    """(A1', A2', lambda') of simple_update.py:28-76; lam1/lam2 are four
    device D-vectors (None for the bond's own leg)."""
    if A1.dim() != 5:
        raise ValueError("acetn HIP engine: site tensors are (D, D, D, D, d)")
    D, d = A1.shape[0], A1.shape[4]
    _expect("A1", A1, (D, D, D, D, d))
    _expect("A2", A2, (D, D, D, D, d))
    _expect("gate", gate, (d, d, d, d))
    for lam in (*lam1, *lam2):
        if lam is not None:
            _expect("lambda", lam, (D,))
    dev = A1.device
    A1o, A2o = _f64(D, D, D, D, d, device=dev), _f64(D, D, D, D, d, device=dev)
    lam = _f64(D, device=dev)
    lam1 = [t.contiguous() if t is not None else None for t in lam1]
    lam2 = [t.contiguous() if t is not None else None for t in lam2]
    A1, A2, gate = A1.contiguous(), A2.contiguous(), gate.contiguous()
    l1 = (c_vp * 4)(*[ptr(t) if t is not None else None for t in lam1])
    l2 = (c_vp * 4)(*[ptr(t) if t is not None else None for t in lam2])
    ws = workspace(lib().acetn_hip_simple_update_workspace(D, d), dev)
    check(lib().acetn_hip_simple_update_bond(ptr(A1), ptr(A2), k, l1, l2,
                                             ptr(gate), D, d, cutoff, ptr(A1o),
                                             ptr(A2o), ptr(lam), ptr(ws), ws.numel(),
                                             stream_handle()))
    return A1o, A2o, lam


# ------------------------------------------------------------ learned ALS start
def mlp_param_count(fin, hidden, fout):
    return int(lib().acetn_hip_mlp_param_count(fin, hidden, fout))


def mlp_forward(params, fin, hidden, fout, X):
    """Y = MLP(X) for X (batch x fin) FP32 on the device."""
    X = X.contiguous()
    B = X.shape[0]
    Y = torch.empty(B, fout, dtype=torch.float32, device=X.device)
    ws = workspace(lib().acetn_hip_mlp_workspace(fin, hidden, fout, B), X.device)
    check(lib().acetn_hip_mlp_forward(ptr(params), fin, hidden, fout, ptr(X), B, ptr(Y), ptr(ws),
                                      ws.numel(), stream_handle()))
    return Y


def mlp_backward(params, fin, hidden, fout, X, T, grad):
    """Forward + MSE loss + backward; fills ``grad``; returns (loss, dL/dY)."""
    X, T = X.contiguous(), T.contiguous()
    B = X.shape[0]
    Y = torch.empty(B, fout, dtype=torch.float32, device=X.device)
    loss = ctypes.c_float(0.0)
    ws = workspace(lib().acetn_hip_mlp_workspace(fin, hidden, fout, B), X.device)
    check(lib().acetn_hip_mlp_backward(ptr(params), fin, hidden, fout, ptr(X), B, ptr(T), ptr(Y),
                                       ptr(grad), ctypes.byref(loss), ptr(ws), ws.numel(),
                                       stream_handle()))
    return loss.value, Y


def mlp_adam_step(params, grad, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8,
                  weight_decay=0.0):
    check(lib().acetn_hip_mlp_adam_step(ptr(params), ptr(grad), ptr(m), ptr(v), params.numel(),
                                        int(step), lr, beta1, beta2, eps, weight_decay,
                                        stream_handle()))
