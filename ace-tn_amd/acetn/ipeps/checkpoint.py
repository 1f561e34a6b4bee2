"""Checkpoint reading and writing (acetn/ipeps/tensor_network.py:94-122).

The reference saves ``{"tensor_network": TensorNetwork, "tensors": {site:
SiteTensor}}`` with torch.save, i.e. pickled acetn objects, and reads it back
with ``torch.load(weights_only=False)``, which executes the pickle.  Here such
a file is read without executing anything and without the reference package:

  * the tensor payloads are the raw little-endian storages ``<name>/data/<key>``
    of the zip archive;
  * each storage's shape and stride come from walking ``data.pkl`` with
    ``pickletools.genops``, which yields opcodes and their literal arguments
    only (nothing is imported, resolved or called);
  * the storages appear in the order the reference's site dictionary
    serialises them: per site, in the x-major site order of
    tensor_network.py:179-193, A then C[0..3] then E[0..3].

The engine's own checkpoints hold plain tensors only and load with
``torch.load(weights_only=True)``.
"""
import pickletools
import zipfile

import numpy as np

TENSORS_PER_SITE = 9  # A, C[0..3], E[0..3]


def _storage_layouts(pkl_bytes):
    """[(storage_key, shape, stride)] in serialisation order, from the literal
    opcode arguments around each persistent-id reference (torch's
    _rebuild_tensor_v2(storage, offset, shape, stride, ...) arguments)."""
    ops = list(pickletools.genops(pkl_bytes))
    out = []
    for i, (op, arg, _) in enumerate(ops):
        if op.name != "BINPERSID":
            continue
        keys = [a for (o, a, _) in ops[max(0, i - 8):i]
                if o.name in ("BINUNICODE", "SHORT_BINUNICODE") and str(a).isdigit()]
        if not keys:
            raise ValueError("checkpoint: a storage reference without a storage key")
        groups, cur, j = [], [], i + 1
        while j < len(ops) and ops[j][0].name != "REDUCE":
            o, a, _ = ops[j]
            if o.name.startswith("BININT") or o.name == "INT":
                cur.append(int(a))
            elif o.name in ("TUPLE", "TUPLE1", "TUPLE2", "TUPLE3", "EMPTY_TUPLE"):
                groups.append(cur)
                cur = []
            j += 1
        if len(groups) < 2 or not groups[0]:
            raise ValueError("checkpoint: unexpected tensor record")
        offset, shape, stride = groups[0][0], tuple(groups[0][1:]), tuple(groups[1])
        kinds = [str(a) for (o, a, _) in ops[max(0, i - 8):i] if o.name == "GLOBAL"]
        kind = kinds[-1].split()[-1] if kinds else "DoubleStorage"
        if kind not in _STORAGE_DTYPES:
            raise ValueError(f"checkpoint: unsupported storage type {kind} (real float64 or "
                             "float32 tensors only)")
        out.append((keys[0], offset, shape, stride, _STORAGE_DTYPES[kind]))
    return out


_STORAGE_DTYPES = {"DoubleStorage": "<f8", "FloatStorage": "<f4"}


def is_reference_checkpoint(path):
    """Whether ``path`` is a torch zip archive whose pickle names acetn
    classes (the reference's save format)."""
    try:
        with zipfile.ZipFile(path) as z:
            pkl = [n for n in z.namelist() if n.endswith("/data.pkl")]
            if not pkl:
                return False
            return any(op.name in ("GLOBAL", "STACK_GLOBAL") and str(arg).startswith("acetn.")
                       for op, arg, _ in pickletools.genops(z.read(pkl[0])))
    except (zipfile.BadZipFile, ValueError):
        return False


def read_reference_checkpoint(path, nx, ny):
    """Site tensors of a reference checkpoint as {site: {'A', 'C', 'E'}} numpy
    float64 arrays, sites in x-major order over an nx x ny cell."""
    sites = [(x, y) for x in range(nx) for y in range(ny)]
    with zipfile.ZipFile(path) as z:
        names = z.namelist()
        pkl = [n for n in names if n.endswith("/data.pkl")]
        if len(pkl) != 1:
            raise ValueError(f"{path}: not a torch zip checkpoint")
        root = pkl[0][:-len("data.pkl")]
        bo = root + "byteorder"
        if bo in names and z.read(bo).decode().strip() != "little":
            raise ValueError(f"{path}: only little-endian checkpoints are read")
        layouts = _storage_layouts(z.read(pkl[0]))
        if len(layouts) != TENSORS_PER_SITE * len(sites):
            raise ValueError(f"{path}: {len(layouts)} tensors, expected "
                             f"{TENSORS_PER_SITE * len(sites)} for a {nx}x{ny} cell")
        out = {}
        for n, (key, offset, shape, stride, dt) in enumerate(layouts):
            raw = np.frombuffer(z.read(f"{root}data/{key}"), dtype=dt)
            need = offset + sum((s - 1) * t for s, t in zip(shape, stride)) + 1
            if raw.size < need:
                raise ValueError(f"{path}: storage {key} is shorter than its tensor")
            t = np.lib.stride_tricks.as_strided(raw[offset:], shape=shape,
                                                strides=tuple(raw.itemsize * s for s in stride))
            t = t.astype(np.float64)
            site, slot = sites[n // TENSORS_PER_SITE], n % TENSORS_PER_SITE
            st = out.setdefault(site, {'A': None, 'C': [None] * 4, 'E': [None] * 4})
            if slot == 0:
                st['A'] = t
            elif slot <= 4:
                st['C'][slot - 1] = t
            else:
                st['E'][slot - 5] = t
    for site, st in out.items():
        A = st['A']
        if A.ndim != 5 or any(c.ndim != 2 for c in st['C']) or any(e.ndim != 4 for e in st['E']):
            raise ValueError(f"{path}: site {site} does not hold A (5-d), C (2-d), E (4-d)")
    return out
