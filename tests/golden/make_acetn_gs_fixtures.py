"""Convert ace-tn's converged ground-state fixtures into plain .npz arrays.

Source: ace-tn tests/integration/ipeps_gs/{ising_dims_2_20_dtau_001_hx_295,
heisenberg_dims_3_16_dtau_001}.pt and energies.csv (read from the reference
checkout in the build container; the output .npz files are committed).

``torch.load(weights_only=True)`` refuses these files (they hold pickled
acetn classes), and no loader that would construct those objects is used.
Instead:
  * the tensor payloads are the raw little-endian float64 storages
    ``<case>/data/<key>`` inside the zip, read with ``numpy.frombuffer``;
  * each storage's shape and stride come from tokenizing ``data.pkl`` with
    ``pickletools.genops`` -- a disassembler that yields opcodes and their
    literal arguments only; nothing is imported, resolved or called;
  * storages appear in serialization order: per site (x-major site list)
    A, C0..C3, E0..E3.
The resulting tensors are validated by the golden energies (rel 1e-10) in
tests/test_oracle.py::test_measure_matches_golden_energy.

Usage:  python tests/golden/make_acetn_gs_fixtures.py /root/reference
"""
import csv
import os
import pickletools
import sys
import zipfile

import numpy as np

CASES = {
    "ising_dims_2_20_dtau_001_hx_295": dict(model="ising", D=2, chi=20, d=2,
                                           params=dict(jz=1.0, hx=2.95)),
    "heisenberg_dims_3_16_dtau_001": dict(model="heisenberg", D=3, chi=16, d=2,
                                         params=dict(J=1.0)),
}
SITES = [(0, 0), (0, 1), (1, 0), (1, 1)]


def storage_layouts(pkl_bytes):
    """Return [(storage_key, shape, stride)] in serialization order, from the
    literal opcode arguments that follow each persistent-id reference."""
    ops = list(pickletools.genops(pkl_bytes))
    out = []
    for i, (op, arg, _) in enumerate(ops):
        if op.name != "BINPERSID":
            continue
        key = [a for (o, a, _) in ops[max(0, i - 8):i]
               if o.name in ("BINUNICODE", "SHORT_BINUNICODE") and str(a).isdigit()][0]
        # literal ints up to the REDUCE: storage offset, then the shape tuple,
        # then the stride tuple
        groups, cur, j = [], [], i + 1
        while ops[j][0].name != "REDUCE":
            o, a, _ = ops[j]
            if o.name.startswith("BININT") or o.name == "INT":
                cur.append(int(a))
            elif o.name in ("TUPLE", "TUPLE1", "TUPLE2", "TUPLE3", "EMPTY_TUPLE"):
                groups.append(cur)
                cur = []
            j += 1
        offset, shape = groups[0][0], tuple(groups[0][1:])
        stride = tuple(groups[1])
        assert offset == 0
        out.append((key, shape, stride))
    return out


def main(ref_root):
    src = os.path.join(ref_root, "tests", "integration", "ipeps_gs")
    energies = {}
    with open(os.path.join(src, "energies.csv"), newline="") as f:
        for row in csv.reader(f):
            energies[row[0]] = float(row[1])
    here = os.path.dirname(os.path.abspath(__file__))
    for case, meta in CASES.items():
        z = zipfile.ZipFile(os.path.join(src, case + ".pt"))
        layouts = storage_layouts(z.read(f"{case}/data.pkl"))
        assert z.read(f"{case}/byteorder").decode() == "little"
        assert len(layouts) == 36
        arrays = {}
        for n, (key, shape, stride) in enumerate(layouts):
            raw = np.frombuffer(z.read(f"{case}/data/{key}"), dtype="<f8")
            t = np.lib.stride_tricks.as_strided(
                raw, shape=shape, strides=tuple(8 * s for s in stride)).copy()
            site = SITES[n // 9]
            slot = n % 9
            name = "A" if slot == 0 else (f"C{slot - 1}" if slot <= 4 else f"E{slot - 5}")
            arrays[f"{site[0]}{site[1]}_{name}"] = t
        np.savez_compressed(os.path.join(here, case + ".npz"),
                            energy=np.array(energies[case]),
                            D=meta["D"], chi=meta["chi"], d=meta["d"], **arrays)
        print(case, energies[case], {k: v.shape for k, v in list(arrays.items())[:9]})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
