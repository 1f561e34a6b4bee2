"""Headline benchmark: CTMRG iterations/s of the MI355X engine.

Workload (one GPU): 2x2 unit cell, d=2, D=6, chi=144 (enlarged-corner matrix
dim chi*D^2 = 5184), site tensors seeded iid N(0,1) plus 0.5 on the spin-up
[0,0,0,0,0] slot, each normalised; environment initialised from A (as
acetn/ipeps/site_tensor.py:152-202 does); ``--warmup`` CTMRG iterations (the
first saturates chi), then ``--steps`` timed iterations, each the full
reference order of 8 directional moves (16 projector pairs + 16 boundary
renormalisations).  Inputs are resident in HBM before the timed region.

Prints ONE JSON line (rank 0).  Diagnostics go to stderr.
  roofline     -- the FP64 matrix-core GEMM kernel (the dominant kernel):
                  summed algorithmic flops (2*M*N*K per launch) over the timed
                  region / summed event-timed launch durations on its stream.
  cpu_baseline -- the FP64 numpy oracle (a restatement of the reference
                  torch path) on the host cores: one left move (2 of the 16
                  projector+renormalisation units of an iteration) on a
                  chi-saturated random environment, scaled to iterations/s.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "ace-tn_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (matrix), AMD published


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def random_state(D, d, seed, bias):
    """Seeded N(0,1) site tensors (+bias on [0,0,0,0,0]), unit norm, in
    x-major site order; environment from A."""
    import torch
    g = torch.Generator().manual_seed(seed)
    sites = {}
    for s in [(0, 0), (0, 1), (1, 0), (1, 1)]:
        a = torch.randn(D, D, D, D, d, generator=g, dtype=torch.float64)
        a[0, 0, 0, 0, 0] += bias
        sites[s] = a / a.norm()
    return sites


class DeviceNetwork:
    """The unit cell on the GPU, in the container shape the movers expect."""

    def __init__(self, A, D, chi, seed):
        import torch
        from acetn.ipeps.site_tensor import SiteTensor
        from acetn.renormalization.projectors import SketchCounter
        self.nx = self.ny = 2
        self.dims = {'bond': D, 'chi': chi, 'phys': A[(0, 0)].shape[-1]}
        self.site_list = list(A)
        self.sites = {}
        for s, a in A.items():
            st = SiteTensor.__new__(SiteTensor)
            st.dims, st.dtype, st.device = self.dims, torch.float64, torch.device("cuda", 0)
            st['A'] = a
            st.initialize_corner_tensors()
            st.initialize_edge_tensors()
            self.sites[s] = st
        self.sketch = SketchCounter(seed)

    def __getitem__(self, s):
        return self.sites[tuple(s)]


def cpu_baseline(D, chi, d=2):
    """Oracle timing on the host (bounded sample, see module docstring)."""
    import numpy as np
    from oracle import ctmrg as oc
    from oracle.rng import SketchStream
    threads = len(os.sched_getaffinity(0))
    rng = np.random.default_rng(5)
    A = oc.random_site_tensors(2, 2, D, d, seed=3, bias=0.5)
    st = {}
    for s, a in A.items():
        C = [rng.standard_normal((chi, chi)) for _ in range(4)]
        E = [rng.standard_normal((chi, chi, D, D)) for _ in range(4)]
        st[s] = {'A': a, 'C': [c / np.linalg.norm(c) for c in C],
                 'E': [e / np.linalg.norm(e) for e in E]}
    mover = oc.DirectionalMover(oc.ProjectorCalculator(omega_fn=SketchStream(1)), 2, 2, chi)
    t0 = time.perf_counter()
    mover.left_move(st, 0)
    dt = time.perf_counter() - t0
    return {"value": 1.0 / (8.0 * dt), "unit": "iters/s", "cores": threads, "kind": "port",
            "sample": f"1 left move (2 of 16 projector+renormalisation units) at D={D} chi={chi} "
                      f"on a chi-saturated random environment, FP64 numpy oracle, {dt:.2f} s, x8"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--D", type=int, default=6)
    ap.add_argument("--chi", type=int, default=144)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note, not part of this build)", file=sys.stderr)
        sys.exit(2)

    import torch
    from acetn import _hip
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover
    _hip.require_gpu()
    log("device:", _hip.device_info())

    class Cfg:
        projectors, svd_type, svd_cutoff = "half-system", "rsvd", 1e-12
        rsvd_niter, rsvd_oversampling = 2, 2

    D, chi = args.D, args.chi
    net = DeviceNetwork(random_state(D, 2, args.seed, 0.5), D, chi, seed=args.seed)
    mover = DirectionalMover(Cfg(), net.sketch)
    for i in range(args.warmup):
        ctmrg_iteration(mover, net)
        log(f"warmup {i}: C0 {tuple(net[(0, 0)]['C'][0].shape)}")
    torch.cuda.synchronize()
    _hip.kernel_timing(True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctmrg_iteration(mover, net)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kms, kflops, klaunch = _hip.kernel_timing_read()
    _hip.kernel_timing(False)
    its = args.steps / dt
    achieved = kflops / (kms * 1e-3) / 1e12 if kms > 0 else 0.0
    log(f"{its:.3f} it/s; gemm: {klaunch} launches, {kms:.1f} ms, {kflops/1e12:.2f} TFLOP -> "
        f"{achieved:.2f} TFLOP/s")
    traffic = None
    tf_path = os.path.join(ROOT, "profiles", "gemm_f64_traffic.json")
    if os.path.exists(tf_path):
        with open(tf_path) as f:
            traffic = json.load(f).get("bytes_per_launch")
    out = {
        "metric": "CTMRG iters/s at D=6 chi=144 (2x2 cell, all four directional moves)",
        "value": its,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded N(0,1) site tensors + 0.5 spin-up bias)",
        "config": {"workload": f"CTMRG iteration, 2x2 unit cell, d=2, D={D}, chi={chi}, "
                               f"half-system rSVD projectors (q=chi+2, 2 power iterations)",
                   "D": D, "chi": chi, "unit_cell": "2x2", "precision": "fp64"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": traffic,
                     "kernel": "gemm_f64_kernel",
                     "flops_per_launch": kflops / max(klaunch, 1),
                     "avg_launch_ms": kms / max(klaunch, 1)},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(D, chi)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
