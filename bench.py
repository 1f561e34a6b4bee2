"""Headline benchmark: CTMRG iterations/s of the MI355X engine.

Workload (one GPU): 2x2 unit cell, d=2, D=6, chi=144 (enlarged-corner matrix
dim chi*D^2 = 5184), site tensors seeded iid N(0,1) plus 0.5 on the spin-up
[0,0,0,0,0] slot, each normalised; environment initialised from A (as
acetn/ipeps/site_tensor.py:152-202 does); ``--warmup`` CTMRG iterations (the
first saturates chi), then ``--steps`` timed iterations, each the full
reference order of 8 directional moves (16 projector pairs + 16 boundary
renormalisations).  Inputs are resident in HBM before the timed region.

The same workload runs twice from the same initial state: with the plain
FP64 engine and with the mixed-precision engine (split-bf16 matrix-core
contractions, FP32 accumulation; FP64 QR / SVD / normalisations).  ``value``
is the mixed run's iterations/s; the FP64 rate, the speedup and the energy
per site of both final environments (Heisenberg J=1 bond operator from the
HIP bond reduced density matrices) are reported beside it.

Prints ONE JSON line (rank 0).  Diagnostics go to stderr.
  roofline     -- the matrix-core kernel class with the most event-timed
                  time in the mixed run, from one extra iteration run after
                  the timed region with the projector pairs on one stream and
                  every launch bracketed by HIP events on that stream: for the
                  FP64 classes summed algorithmic flops (2*M*N*K per GEMM
                  launch; 4*d*D^6 per double-layer block) / summed launch
                  durations against the FP64 matrix peak; for the split-bf16
                  class (HBM-bound by construction) summed algorithmic bytes
                  (operand planes read once, C written once) / summed launch
                  durations against HBM bandwidth, its FP64-equivalent flop
                  rate reported beside.
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

FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (matrix), AMD published
HBM_PEAK_GBS = 8000.0     # MI355X HBM3E, AMD published
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 (matrix), AMD published


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


class _Cfg:
    projectors, svd_type, svd_cutoff = "half-system", "rsvd", 1e-12
    rsvd_niter, rsvd_oversampling = 2, 2

    def __init__(self, precision):
        self.precision = precision


def heisenberg_energy(net):
    """Energy per site of the J=1 Heisenberg bond operator from the HIP bond
    reduced density matrices (the 8 bonds of the 2x2 cell)."""
    import torch
    from acetn import _hip
    from acetn.model.pauli_matrix import pauli_matrices
    X, Y, Z, I = pauli_matrices(torch.float64)
    h = (0.25 * (X * X + Y * Y + Z * Z)).mat
    h = torch.as_tensor(h, dtype=torch.float64).cpu().reshape(2, 2, 2, 2)
    bonds = [((x, y), ((x + 1) % 2, y), 2) for x in range(2) for y in range(2)] + \
            [((x, y), (x, (y + 1) % 2), 1) for y in range(2) for x in range(2)]
    e = 0.0
    for s1, s2, k in bonds:
        a, b = net[s1], net[s2]
        rho = _hip.bond_rdm((a['A'], a['C'], a['E']), (b['A'], b['C'], b['E']), k).cpu()
        e += float(torch.einsum("PQpq,pqPQ->", rho, h) / torch.einsum("pqpq->", rho))
    return e / 4.0


def run_ctmrg(precision, args):
    """Warmup + timed CTMRG iterations at one precision; returns timings,
    per-class matrix-core kernel statistics and the final energy/site."""
    import torch
    from acetn import _hip
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover
    D, chi = args.D, args.chi
    net = DeviceNetwork(random_state(D, 2, args.seed, 0.5), D, chi, seed=args.seed)
    mover = DirectionalMover(_Cfg(precision), net.sketch)
    for i in range(args.warmup):
        ctmrg_iteration(mover, net)
        log(f"[{precision}] warmup {i}: C0 {tuple(net[(0, 0)]['C'][0].shape)}")
    torch.cuda.synchronize()
    _hip.serialize(args.serial)
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctmrg_iteration(mover, net)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # roofline pass: one more iteration with the projector pairs serialised on
    # one stream, every matrix-core launch bracketed by HIP events on that
    # stream (with overlapping streams an event interval would also count time
    # spent waiting for compute units held by other streams' kernels)
    _hip.serialize(True)
    _hip.kernel_timing(True)
    ctmrg_iteration(mover, net)
    torch.cuda.synchronize()
    _hip.serialize(args.serial)
    stats = {}
    for name, cls in (("gemm_f64_kernel", _hip.K_GEMM_F64),
                      ("gemm_split_kernel", _hip.K_GEMM_SPLIT),
                      ("dl_mfma_kernel", _hip.K_DOUBLE_LAYER)):
        ms, fl, n = _hip.kernel_timing_read(cls)
        if n:
            stats[name] = {"ms": ms, "flops": fl, "launches": n,
                           "bytes": _hip.kernel_timing_bytes(cls)}
    _hip.kernel_timing(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    energy = heisenberg_energy(net)
    t_measure = time.perf_counter() - t0
    spec = mover.projector_calculator.last_spectrum
    if spec is not None:
        sv = spec.cpu()
        sv = sv[sv > 0]
        log(f"[{precision}] last projector spectrum: s_min/s_0 = {float(sv[-1] / sv[0]):.3e} "
            f"over {len(sv)} values")
    its = args.steps / dt
    log(f"[{precision}] {its:.3f} it/s, E/site {energy:.12f}; " +
        "; ".join(f"{k}: {v['launches']} launches {v['ms']:.1f} ms "
                  f"{v['flops'] / v['ms'] / 1e9:.2f} TFLOP/s" for k, v in stats.items()))
    log(f"[{precision}] energy from the 8 bond RDMs in {1e3 * t_measure:.1f} ms")
    return {"its": its, "dt": dt, "stats": stats, "energy": energy, "measure_s": t_measure}


def run_als(args):
    """The full-update ALS workload: Heisenberg J=1 on the 2x2 cell, d=2,
    D=6, chi=72, gate exp(-0.01 H).  Seeded N(0,1) site tensors, CTMRG
    environment, then fast-full-update sweeps that record every bond's ALS
    problem (gauge-fixed n12, a12g) and converged solution.  The perceptron is
    trained on the first 3/4 of the problems (device kernels) and both starts
    are run on the rest to relative cost change < 1e-10."""
    import torch
    from acetn.evolution.learned_start import (MLP, MLPProposer, dataset, evaluate,
                                               feature_sizes, generate_problems, solve_from, train)
    from acetn.ipeps import Ipeps
    from acetn.model.model_factory import model_factory
    from acetn import _hip
    D, chi = 6, 72
    config = {'dtype': 'float64', 'device': 'cuda',
              'TN': {'nx': 2, 'ny': 2, 'dims': {'phys': 2, 'bond': D, 'chi': chi}},
              'model': {'name': 'heisenberg', 'params': {'J': 1.0}},
              'ctmrg': {'steps': 3, 'sketch_seed': args.seed, 'disable_progressbar': True},
              'evolution': {'disable_progressbar': True}}
    ipeps = Ipeps(config)
    g = torch.Generator().manual_seed(args.seed)
    for s in ipeps.site_list:
        a = torch.randn(D, D, D, D, 2, generator=g, dtype=torch.float64)
        ipeps[s]['A'] = a / a.norm()
        ipeps[s].initialize_corner_tensors()
        ipeps[s].initialize_edge_tensors()
    ipeps.site_states_initialized = True
    ipeps.renormalize()
    model = model_factory.create(ipeps.config.model)
    t0 = time.perf_counter()
    probs = generate_problems(ipeps, model, 0.01, args.als_steps, ipeps.config.evolution)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    ntr = (3 * len(probs)) // 4
    train_p, test_p = probs[:ntr], probs[ntr:]
    fin, fout = feature_sizes(2 * D, 2)
    X, T = dataset(train_p)
    mlp = MLP(fin, fout, hidden=512, seed=0)
    t0 = time.perf_counter()
    losses = train(mlp, X, T, epochs=args.als_epochs, batch=32, lr=1e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    cfg = ipeps.config.evolution
    res = evaluate(MLPProposer(mlp, cfg), test_p, cfg, tol=1e-10)
    # the reference's default stopping tolerance (als_tol = 1e-15)
    ref_tol = evaluate(MLPProposer(mlp, cfg), test_p, cfg, tol=cfg.als_tol)
    res["at_default_tol"] = {"tol": cfg.als_tol, **ref_tol,
                             "sweep_reduction": 1.0 - ref_tol["mean_sweeps_learned"] /
                             ref_tol["mean_sweeps_plain"]}
    # plain ALS solves/s (plain start included), to rel. cost change < 1e-10
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for pr in test_p:
        a1, a2 = _hip.als_initialize(pr['a12g'], D)
        solve_from(pr['n12'], pr['a12g'], a1, a2, cfg, 1e-10)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    res.update({"als_solves_per_s": len(test_p) / dt, "train_problems": ntr,
                "train_loss_first_last": [losses[0], losses[-1]],
                "sweep_reduction": 1.0 - res["mean_sweeps_learned"] / res["mean_sweeps_plain"],
                "generate_s": t_gen, "train_s": t_train,
                "workload": f"Heisenberg 2x2, D={D}, chi={chi}, dtau=0.01, {args.als_steps} "
                            f"fast-full-update sweeps from seeded N(0,1) tensors; MLP 3 layers, "
                            f"width 512, bf16 MFMA, {args.als_epochs} epochs"})
    log(f"[als] {res}")
    return res


def run_sweep(args):
    """Parameter sweep of independent instances: 12 transverse-field Ising
    iPEPS (2x2 cell, d=2, J=1, h uniform in [2.5, 3.5]), D drawn from
    {3, 4, 5, 6}, chi = 4 D^2, site tensors seeded N(0,1) + 0.5 on the
    spin-up component, mixed-precision contractions and FP64 projectors.
    CTMRG only (the field h enters the energy, not the boundary), so h is
    reported with each instance.  The instances run together
    (ctmrg_iteration_many) and, for comparison, one after another with the
    single-network driver; both must leave identical tensors."""
    import torch
    from acetn.renormalization.ctmrg import ctmrg_iteration, ctmrg_iteration_many
    from acetn.renormalization.directional_mover import DirectionalMover
    g = torch.Generator().manual_seed(args.seed)
    Ds = [int(x) for x in torch.randint(3, 7, (12,), generator=g)]
    hs = [float(x) for x in 2.5 + torch.rand(12, generator=g, dtype=torch.float64)]
    insts = [(D, 4 * D * D, args.seed * 100 + i) for i, D in enumerate(Ds)]

    def nets():
        out = []
        for D, chi, seed in insts:
            net = DeviceNetwork(random_state(D, 2, seed, 0.5), D, chi, seed=seed)
            out.append((DirectionalMover(_Cfg("mixed"), net.sketch), net))
        return out

    steps, warm = args.sweep_steps, 2
    # together
    pairs = nets()
    for _ in range(warm):
        ctmrg_iteration_many(pairs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctmrg_iteration_many(pairs)
    torch.cuda.synchronize()
    dt_many = time.perf_counter() - t0
    # one after another
    alone = nets()
    for m, net in alone:
        for _ in range(warm):
            ctmrg_iteration(m, net)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for m, net in alone:
        for _ in range(steps):
            ctmrg_iteration(m, net)
    torch.cuda.synchronize()
    dt_seq = time.perf_counter() - t0
    same = all(torch.equal(a[s][key][k], b[s][key][k])
               for (_, a), (_, b) in zip(pairs, alone) for s in a.site_list
               for key in ('C', 'E') for k in range(4))
    flops = sum(ctmrg_flops(D, chi) for D, chi, _ in insts)
    res = {"workload": "12 transverse-field Ising iPEPS, 2x2 cell, d=2, J=1, D drawn from "
                       "{3,4,5,6}, chi=4D^2, mixed-precision contractions, FP64 projectors",
           "instances": [{"D": D, "chi": chi, "h": h} for (D, chi, _), h in zip(insts, hs)],
           "steps": steps, "warmup": warm,
           "instance_iters_per_s_together": 12 * steps / dt_many,
           "instance_iters_per_s_one_by_one": 12 * steps / dt_seq,
           "speedup_together": dt_seq / dt_many,
           "tflops_together": flops * steps / dt_many / 1e12,
           "identical_tensors": same}
    log(f"[sweep] {res}")
    return res


def ctmrg_flops(D, chi, d=2):
    """Algorithmic FLOPs of one CTMRG iteration (2x2 cell, half-system rSVD,
    q = chi + 2, 2 power iterations): 16 projector pairs of two quarter
    tensors, 12 skinny range-finder products, projector products and the
    renormalisation of C and E (the per-row counts of the hot-path table)."""
    N, q = chi * D * D, chi + 2
    quarter = 2 * chi ** 3 * D ** 2 + 2 * chi ** 3 * D ** 4 + 4 * chi ** 2 * D ** 6 * d
    rsvd = 24 * N * N * q
    proj = 4 * N * N * chi
    renorm = 4 * chi ** 3 * D ** 4 + 4 * chi ** 2 * D ** 6 * d
    return 16 * (2 * quarter + rsvd + proj) + 16 * renorm


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--D", type=int, default=6)
    ap.add_argument("--chi", type=int, default=144)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--als-steps", type=int, default=64,
                    help="fast-full-update sweeps recording ALS problems (8 per sweep)")
    ap.add_argument("--als-epochs", type=int, default=60)
    ap.add_argument("--no-als", action="store_true")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the 12-instance parameter sweep")
    ap.add_argument("--sweep-steps", type=int, default=3)
    ap.add_argument("--serial", action="store_true",
                    help="run the timed iterations with the projector pairs serialised")
    ap.add_argument("--only", choices=["fp64", "mixed"], default=None,
                    help="run one precision only (profiling runs); prints no JSON line")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB once with a plain device copy (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {args.copy_calibration} MiB")
        return
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note, not part of this build)", file=sys.stderr)
        sys.exit(2)

    from acetn import _hip
    _hip.require_gpu()
    log("device:", _hip.device_info())
    D, chi = args.D, args.chi
    if args.only:
        run_ctmrg(args.only, args)
        return
    ref = run_ctmrg("fp64", args)
    mix = run_ctmrg("mixed", args)
    dE = abs(mix["energy"] - ref["energy"])
    # roofline of the dominant matrix-core kernel of the headline (mixed) run
    # (the split-bf16 class is HBM-bound by construction: its MFMA work is 3 or
    # 6 bf16 piece products per FP64-equivalent flop, about 250 flop per byte
    # of planes moved, below the bf16 ridge of 312; it is measured against
    # HBM bandwidth with its algorithmic bytes, the FP64 classes against the
    # FP64 matrix peak)
    name, st = max(mix["stats"].items(), key=lambda kv: kv[1]["ms"])
    tflops = st["flops"] / (st["ms"] * 1e-3) / 1e12
    if name == "gemm_split_kernel":
        bound, unit, peak = "hbm", "GB/s", HBM_PEAK_GBS
        achieved = st["bytes"] / (st["ms"] * 1e-3) / 1e9
    else:
        bound, unit = "mfma", "TFLOP/s"
        peak = {"gemm_f64_kernel": FP64_PEAK_TFLOPS, "dl_mfma_kernel": FP64_PEAK_TFLOPS}[name]
        achieved = tflops
    traffic = None
    tf_path = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if os.path.exists(tf_path):
        with open(tf_path) as f:
            traffic = json.load(f).get("bytes_per_launch")
    out = {
        "metric": "CTMRG iters/s at D=6 chi=144, mixed vs FP64 at |dE/site|<=1e-6; ALS sweeps",
        "value": mix["its"],
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * mix["dt"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3/bf16x6 split MFMA (FP32 accumulate) + f64",
        "data": "synthetic (seeded N(0,1) site tensors + 0.5 spin-up bias)",
        "config": {"workload": f"CTMRG iteration, 2x2 unit cell, d=2, D={D}, chi={chi}, "
                               f"half-system rSVD projectors (q=chi+2, 2 power iterations), "
                               f"all four directional moves",
                   "D": D, "chi": chi, "unit_cell": "2x2", "precision": "mixed"},
        "fp64_iters_per_s": ref["its"],
        "speedup_vs_fp64": mix["its"] / ref["its"],
        "bond_rdm_energy_ms": 1e3 * ref["measure_s"],
        "energy_per_site": {"mixed": mix["energy"], "fp64": ref["energy"], "abs_diff": dE,
                            "note": "Heisenberg J=1 bond operator on the final environments; "
                                    "the FP64 HIP path is the parity-tested restatement"},
        "roofline": {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
                     "frac": achieved / peak, "traffic": traffic, "kernel": name,
                     "algorithmic_bytes_per_launch": st["bytes"] / st["launches"],
                     "flops_per_launch": st["flops"] / st["launches"],
                     "fp64_equivalent_tflops": tflops,
                     "avg_launch_ms": st["ms"] / st["launches"]},
        "kernels": {"mixed": mix["stats"], "fp64": ref["stats"]},
    }
    if not args.no_sweep:
        out["sweep"] = run_sweep(args)
    if not args.no_als:
        out["als"] = run_als(args)
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(D, chi)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
