"""Convert ace-tn's converged ground-state fixtures into plain .npz arrays.

Source: ace-tn tests/integration/ipeps_gs/{ising_dims_2_20_dtau_001_hx_295,
heisenberg_dims_3_16_dtau_001}.pt and energies.csv (read from the reference
checkout in the build container; the output .npz files are committed).

``torch.load(weights_only=True)`` refuses these files (they hold pickled
acetn classes), and no loader that would construct those objects is used:
they are read by acetn.ipeps.checkpoint.read_reference_checkpoint, which
takes the raw storages out of the zip and their shapes from the pickle's
opcodes (pickletools.genops: nothing is imported, resolved or called).
The resulting tensors are validated by the golden energies (rel 1e-10) in
tests/test_oracle.py::test_measure_matches_golden_energy.

Usage:  python tests/golden/make_acetn_gs_fixtures.py /root/reference
"""
import csv
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "ace-tn_amd"))
from acetn.ipeps.checkpoint import read_reference_checkpoint  # noqa: E402

CASES = {
    "ising_dims_2_20_dtau_001_hx_295": dict(model="ising", D=2, chi=20, d=2,
                                           params=dict(jz=1.0, hx=2.95)),
    "heisenberg_dims_3_16_dtau_001": dict(model="heisenberg", D=3, chi=16, d=2,
                                         params=dict(J=1.0)),
}
SITES = [(0, 0), (0, 1), (1, 0), (1, 1)]


def main(ref_root):
    src = os.path.join(ref_root, "tests", "integration", "ipeps_gs")
    energies = {}
    with open(os.path.join(src, "energies.csv"), newline="") as f:
        for row in csv.reader(f):
            energies[row[0]] = float(row[1])
    here = os.path.dirname(os.path.abspath(__file__))
    for case, meta in CASES.items():
        st = read_reference_checkpoint(os.path.join(src, case + ".pt"), 2, 2)
        arrays = {}
        for site in SITES:
            arrays[f"{site[0]}{site[1]}_A"] = st[site]['A']
            for k in range(4):
                arrays[f"{site[0]}{site[1]}_C{k}"] = st[site]['C'][k]
                arrays[f"{site[0]}{site[1]}_E{k}"] = st[site]['E'][k]
        np.savez_compressed(os.path.join(here, case + ".npz"),
                            energy=np.array(energies[case]),
                            D=meta["D"], chi=meta["chi"], d=meta["d"], **arrays)
        print(case, energies[case], {k: v.shape for k, v in list(arrays.items())[:9]})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
