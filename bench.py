"""Headline benchmark: CTMRG iterations/s of the MI355X engine.

Workload (one GPU): Heisenberg J=1 on a 2x2 unit cell, d=2, D=6, chi=144
(enlarged-corner matrix dim chi*D^2 = 5184).  The site tensors are a physical
state: 50 device simple-update steps (dtau = 0.05) from seeded N(0,1) tensors,
each normalised; the environment is initialised from A (as
acetn/ipeps/site_tensor.py:152-202 does); ``--warmup`` CTMRG iterations (the
first saturates chi), then ``--steps`` timed iterations, each the full
reference order of 8 directional moves (16 projector pairs + 16 boundary
renormalisations).  Inputs are resident in HBM before the timed region.

The same workload runs twice from the same state: with the plain FP64 engine
and with the mixed-precision engine (split-bf16 matrix-core contractions,
FP32 accumulation; FP64 QR / SVD / normalisations).  ``value`` is the mixed
run's iterations/s; beside it: the FP64 rate and the speedup, and for both
runs the energy per site and the nearest-neighbour correlators <XX>, <YY>,
<ZZ> (bond reduced density matrices) after ``--energy-iters`` further
untimed iterations and as means over samples every 5 of them, the projector
truncation errors of the last timed iteration, and the mixed-minus-FP64
differences, next to the same differences between two FP64 runs that differ
only in their tile choices (the environment's own spread).  The FP64 engine is the path the tests hold to the FP64 oracle
at this size (one left move, projectors and boundary tensors within 1e-9).

Prints ONE JSON line (rank 0).  Diagnostics go to stderr.
  roofline     -- from one extra iteration run after the timed region with
                  the projector pairs on one stream and every launch of the
                  timed kernel classes bracketed by HIP events on that stream:
                  the matrix-core class with the most event time (summed
                  algorithmic flops / summed launch time against the FP64
                  matrix peak; the split-bf16 class, HBM-bound by
                  construction, as algorithmic bytes / time against HBM
                  bandwidth).  ``kernel_time_shares`` lists every timed class
                  with its share, including the one-workgroup eigensolver and
                  Cholesky kernels (latency-bound, no roofline).
  cpu_baseline -- the FP64 numpy oracle (a restatement of the reference
                  torch path) on the host cores: one full iteration (16
                  projector pairs + 16 renormalisations) from the chi-saturated
                  state the GPU run reached after its warmup.
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


def physical_state(D, seed, steps=50, dtau=0.05):
    """Site tensors of a physical state: ``steps`` device simple-update steps
    of the J=1 Heisenberg model (acetn/evolution/simple_update.py on the HIP
    engine) from seeded N(0,1) tensors, each normalised; host copies."""
    import torch
    from acetn.evolution.evolve import evolve
    from acetn.ipeps import Ipeps
    from acetn.model.model_factory import model_factory
    config = {'dtype': 'float64', 'device': 'cuda',
              'TN': {'nx': 2, 'ny': 2, 'dims': {'phys': 2, 'bond': D, 'chi': D * D}},
              'model': {'name': 'heisenberg', 'params': {'J': 1.0}},
              'ctmrg': {'steps': 1, 'sketch_seed': seed, 'disable_progressbar': True},
              'evolution': {'update_type': 'simple', 'disable_progressbar': True}}
    ipeps = Ipeps(config)
    A = random_state(D, 2, seed, 0.0)
    for s in ipeps.site_list:
        ipeps[s]['A'] = A[tuple(s)]
    ipeps.site_states_initialized = True
    # evolve reports its progress on stdout (as the reference's does); the
    # bench's stdout carries only its JSON line
    import contextlib
    with contextlib.redirect_stdout(sys.stderr):
        evolve(ipeps, dtau, steps, model_factory.create(ipeps.config.model), ipeps.config.evolution)
    out = {}
    for s in ipeps.site_list:
        a = ipeps[s]['A'].double().cpu()
        out[tuple(s)] = a / a.norm()
    return out


def bond_observables(net):
    """Energy per site of the J=1 Heisenberg bond operator and the
    nearest-neighbour correlators <XX>, <YY>, <ZZ> averaged over the 8 bonds of
    the 2x2 cell, from the HIP bond reduced density matrices."""
    import torch
    from acetn import _hip
    from acetn.model.pauli_matrix import pauli_matrices
    X, Y, Z, I = pauli_matrices(torch.float64)
    ops = {k: torch.as_tensor((a * a).mat, dtype=torch.float64).cpu().reshape(2, 2, 2, 2)
           for k, a in (("XX", X), ("YY", Y), ("ZZ", Z))}
    h = 0.25 * (ops["XX"] + ops["YY"] + ops["ZZ"])
    bonds = [((x, y), ((x + 1) % 2, y), 2) for x in range(2) for y in range(2)] + \
            [((x, y), (x, (y + 1) % 2), 1) for y in range(2) for x in range(2)]
    e, corr = 0.0, {k: 0.0 for k in ops}
    for s1, s2, k in bonds:
        a, b = net[s1], net[s2]
        rho = _hip.bond_rdm((a['A'], a['C'], a['E']), (b['A'], b['C'], b['E']), k).cpu()
        nrm = float(torch.einsum("pqpq->", rho))
        e += float(torch.einsum("PQpq,pqPQ->", rho, h)) / nrm
        for key, op in ops.items():
            corr[key] += float(torch.einsum("PQpq,pqPQ->", rho, op)) / nrm / len(bonds)
    return e / 4.0, corr


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


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def cpu_baseline(state, D, chi, seed):
    """Oracle timing on the host: one full CTMRG iteration (8 moves: 16
    projector pairs, 16 renormalisations) from ``state``, the chi-saturated
    environment of the GPU run after its warmup (numpy arrays)."""
    import copy

    import torch
    from oracle import ctmrg as oc
    from oracle.rng import SketchStream
    threads = len(os.sched_getaffinity(0))
    st = copy.deepcopy(state)
    mover = oc.DirectionalMover(oc.ProjectorCalculator(omega_fn=SketchStream(seed)), 2, 2, chi)
    t0 = time.perf_counter()
    oc.ctmrg_iteration(mover, st)
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "iters/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_info(), "blas_threads": os.environ.get("OMP_NUM_THREADS", "default"),
            "torch_interop_threads": torch.get_num_interop_threads(),
            "sample": f"1 full CTMRG iteration (8 moves, 16 projector pairs) at D={D} chi={chi} "
                      f"from the chi-saturated state of the GPU run after its warmup, FP64 numpy "
                      f"oracle on {threads} host threads, {dt:.1f} s"}


class _Cfg:
    projectors, svd_type, svd_cutoff = "half-system", "rsvd", 1e-12
    rsvd_niter, rsvd_oversampling = 2, 2

    def __init__(self, precision):
        self.precision = precision


KERNEL_CLASSES = (("gemm_f64_kernel", 0), ("gemm_split_kernel", 1), ("dl_mfma_kernel", 2),
                  ("jacobi_eigh", 3), ("chol_rinv", 4))


def run_ctmrg(precision, args, A, keep_state=False):
    """Warmup + timed CTMRG iterations at one precision from site tensors A;
    returns timings, per-class kernel statistics, the final energy/site, the
    nearest-neighbour correlators and the truncation errors of the last
    iteration's 16 projector pairs."""
    import torch
    from acetn import _hip
    from acetn.renormalization.ctmrg import ctmrg_iteration
    from acetn.renormalization.directional_mover import DirectionalMover
    D, chi = args.D, args.chi
    net = DeviceNetwork(A, D, chi, seed=args.seed)
    mover = DirectionalMover(_Cfg(precision), net.sketch)
    for i in range(args.warmup):
        ctmrg_iteration(mover, net)
        log(f"[{precision}] warmup {i}: C0 {tuple(net[(0, 0)]['C'][0].shape)}")
    torch.cuda.synchronize()
    mode = args.autotune if precision == "fp64" else args.autotune_mixed
    if args.tile_table:
        # a saved tile table (e.g. from an earlier run's --save-tile-table):
        # the same kernel choices without a tuning iteration, as profiling
        # runs need
        _hip.autotune_clear()
        n = _hip.autotune_load(args.tile_table)
        _hip.autotune("table")
        log(f"[{precision}] loaded {n} tile choices from {args.tile_table}")
    elif mode == "off":
        # the built-in thresholds only (no table left from an earlier run)
        _hip.autotune("off")
        _hip.autotune_clear()
    else:
        # one more iteration tunes the tile shape of each product shape (with
        # the projector pairs serialised, every product timed alone, or
        # concurrent as in the timed loop); the timed iterations then use the
        # table
        _hip.autotune_clear()
        _hip.autotune("tune")
        _hip.serialize(mode == "serial")
        t0 = time.perf_counter()
        ctmrg_iteration(mover, net)
        torch.cuda.synchronize()
        _hip.serialize(False)
        _hip.autotune("table")
        log(f"[{precision}] autotuned {_hip.autotune_entries()} product shapes in "
            f"{time.perf_counter() - t0:.1f} s")
        if args.save_tile_table:
            _hip.autotune_save(f"{args.save_tile_table}.{precision}")
    warm = None
    if keep_state:
        warm = {s: {'A': net[s]['A'].cpu().numpy(), 'C': [c.cpu().numpy() for c in net[s]['C']],
                    'E': [e.cpu().numpy() for e in net[s]['E']]} for s in net.site_list}
        warm_stream = net.sketch.next
    _hip.serialize(args.serial)
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctmrg_iteration(mover, net)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    trunc = mover.projector_calculator.truncation_errors(16)
    # roofline pass: one more iteration with the projector pairs serialised on
    # one stream, every timed launch bracketed by HIP events on that stream
    # (with overlapping streams an event interval would also count time
    # spent waiting for compute units held by other streams' kernels)
    _hip.serialize(True)
    _hip.kernel_timing(True)
    ctmrg_iteration(mover, net)
    torch.cuda.synchronize()
    _hip.serialize(args.serial)
    stats = {}
    for name, cls in KERNEL_CLASSES:
        ms, fl, n = _hip.kernel_timing_read(cls)
        if n:
            stats[name] = {"ms": ms, "flops": fl, "launches": n,
                           "bytes": _hip.kernel_timing_bytes(cls)}
    _hip.kernel_timing(False)
    # energy and correlators sampled every 5 of the untimed iterations: at
    # this size the randomized projectors keep the environment fluctuating
    # around its fixed point (energy spread ~1e-6 from one iteration to the
    # next, and between FP64 runs that differ only in summation order), so
    # the precisions are compared on the window means as well as on the last
    # sample
    samples = []
    for i in range(args.energy_iters):
        ctmrg_iteration(mover, net)
        if (i + 1) % 5 == 0 and i + 1 != args.energy_iters:
            samples.append(bond_observables(net))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    energy, corr = bond_observables(net)
    t_measure = time.perf_counter() - t0
    samples.append((energy, corr))
    e_s = [e for e, _ in samples]
    e_mean = sum(e_s) / len(e_s)
    e_sd = (sum((e - e_mean) ** 2 for e in e_s) / len(e_s)) ** 0.5
    corr_mean = {k: sum(c[k] for _, c in samples) / len(samples) for k in corr}
    its = args.steps / dt
    log(f"[{precision}] {its:.3f} it/s, E/site {energy:.12f}, correlators {corr}, "
        f"truncation error max {max(trunc):.3e}; " +
        "; ".join(f"{k}: {v['launches']} launches {v['ms']:.1f} ms" for k, v in stats.items()))
    log(f"[{precision}] energy and correlators from the 8 bond RDMs in {1e3 * t_measure:.1f} ms")
    out = {"its": its, "dt": dt, "stats": stats, "energy": energy, "correlators": corr,
           "energy_mean": e_mean, "energy_sd": e_sd, "energy_samples": len(e_s),
           "correlators_mean": corr_mean,
           "truncation_error": {"max": max(trunc), "mean": sum(trunc) / len(trunc),
                                "pairs": len(trunc)},
           "measure_s": t_measure}
    if keep_state:
        out["warm_state"], out["warm_stream"] = warm, warm_stream
    return out


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
    # plain ALS solves/s (plain start included), to rel. cost change < 1e-10:
    # one bond at a time (the full update's order, a host readback each) and
    # all problems in one batched launch (one workgroup per problem)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for pr in test_p:
        a1, a2 = _hip.als_initialize(pr['a12g'], D)
        solve_from(pr['n12'], pr['a12g'], a1, a2, cfg, 1e-10)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    n12s = torch.stack([pr['n12'] for pr in probs])
    a12s = torch.stack([pr['a12g'] for pr in probs])
    _hip.als_solve_batch(n12s, a12s, D, cfg.als_niter, 1e-10, cfg.als_method, cfg.als_epsilon)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, _, sw_b, _, _ = _hip.als_solve_batch(n12s, a12s, D, cfg.als_niter, 1e-10, cfg.als_method,
                                            cfg.als_epsilon)
    torch.cuda.synchronize()
    dt_batch = time.perf_counter() - t0
    vals, cnts = torch.unique(sw_b.cpu(), return_counts=True)
    hist = {int(v): int(c) for v, c in zip(vals, cnts)}
    # the FP64 oracle's ALS on the host, same problems and start, first 16
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import full_update as ofu
        sample = test_p[:16]
        t0 = time.perf_counter()
        for pr in sample:
            a12g = pr['a12g'].cpu().numpy()
            shape = (a12g.shape[0], D, a12g.shape[2])
            ofu.als_solve(pr['n12'].cpu().numpy(), a12g, shape, cfg.als_niter, 1e-10,
                          start=ofu.initialize_reduced_tensors(a12g, shape))
        cpu = {"value": len(sample) / (time.perf_counter() - t0), "unit": "ALS solves/s",
               "cores": len(os.sched_getaffinity(0)), "kind": "port",
               "sample": f"{len(sample)} of the test problems, plain start, rel. cost change "
                         f"< 1e-10, FP64 numpy oracle"}
    res.update({"als_solves_per_s": len(test_p) / dt,
                "als_solves_per_s_batched": len(probs) / dt_batch,
                "batched_problems": len(probs),
                "plain_sweep_histogram_all_problems": hist,
                "train_problems": ntr,
                "cpu_baseline": cpu,
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
    # CPU baseline of the sweep: the instances one after another on the FP64
    # oracle, one iteration per distinct D from the instance's own initial
    # state (D=6 taken from the headline baseline, the same size)
    cpu = None
    if not args.no_cpu_baseline:
        import numpy as np
        from oracle import ctmrg as oc
        from oracle.rng import SketchStream
        per_d = {}
        for D, chi, seed in insts:
            if D in per_d or D == 6:
                continue
            A = {s: a.numpy() for s, a in random_state(D, 2, seed, 0.5).items()}
            st = oc.make_state(A)
            mover = oc.DirectionalMover(oc.ProjectorCalculator(omega_fn=SketchStream(seed)), 2, 2,
                                        chi)
            oc.ctmrg_iteration(mover, st)   # grows the boundary to chi
            t0 = time.perf_counter()
            oc.ctmrg_iteration(mover, st)
            per_d[D] = time.perf_counter() - t0
        cpu = {"seconds_per_iteration_by_D": per_d}
    res = {"workload": "12 transverse-field Ising iPEPS, 2x2 cell, d=2, J=1, D drawn from "
                       "{3,4,5,6}, chi=4D^2, mixed-precision contractions, FP64 projectors",
           "instances": [{"D": D, "chi": chi, "h": h} for (D, chi, _), h in zip(insts, hs)],
           "steps": steps, "warmup": warm,
           "instance_iters_per_s_together": 12 * steps / dt_many,
           "instance_iters_per_s_one_by_one": 12 * steps / dt_seq,
           "speedup_together": dt_seq / dt_many,
           "tflops_together": flops * steps / dt_many / 1e12,
           "identical_tensors": same}
    if cpu is not None:
        res["cpu_sequential"] = cpu
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
    ap.add_argument("--autotune", choices=["off", "serial", "concurrent"], default="serial",
                    help="tile autotuning of the FP64 run: off (built-in thresholds), or one "
                         "tuning iteration with the projector pairs serialised / concurrent")
    ap.add_argument("--autotune-mixed", choices=["off", "serial", "concurrent"], default="serial",
                    help="the same for the mixed-precision run")
    ap.add_argument("--energy-iters", type=int, default=100,
                    help="untimed iterations after the timed ones, over which the energy "
                         "and correlators are sampled every 5 iterations (window means and the "
                         "last sample are reported)")
    ap.add_argument("--serial", action="store_true",
                    help="run the timed iterations with the projector pairs serialised")
    ap.add_argument("--only", choices=["fp64", "mixed"], default=None,
                    help="run one precision only (profiling runs); prints no JSON line")
    ap.add_argument("--tile-table", default=None, metavar="PATH",
                    help="load this tile table instead of tuning (both precisions)")
    ap.add_argument("--save-tile-table", default=None, metavar="PREFIX",
                    help="write each precision's tuned table to PREFIX.<precision>")
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
    t0 = time.perf_counter()
    A = physical_state(D, args.seed)
    log(f"physical state: 50 simple-update steps at D={D} in {time.perf_counter() - t0:.1f} s")
    if args.only:
        run_ctmrg(args.only, args, A)
        return
    ref = run_ctmrg("fp64", args, A, keep_state=not args.no_cpu_baseline)
    mix = run_ctmrg("mixed", args, A)
    dE = abs(mix["energy"] - ref["energy"])
    # the FP64 run's own rounding spread: the same FP64 run with the other
    # tile choice (built-in thresholds, or tuned), i.e. the same arithmetic
    # summed in another order
    import copy as _copy
    alt = _copy.copy(args)
    alt.autotune = "off" if args.autotune != "off" else "serial"
    alt.tile_table = None
    ref2 = run_ctmrg("fp64", alt, A)
    spread = abs(ref2["energy"] - ref["energy"])
    # Roofline of the matrix-core class with the most event time in the
    # headline (mixed) run.  The split-bf16 class is HBM-bound by
    # construction (its MFMA work is 3 or 6 bf16 piece products per
    # FP64-equivalent flop, about 250 flop per byte of planes moved, below the
    # bf16 ridge of 312): it is measured against HBM bandwidth with its
    # algorithmic bytes, the FP64 classes against the FP64 matrix peak.  The
    # one-workgroup eigensolver and Cholesky classes are latency-bound and
    # listed with their time shares only.
    mfma_classes = ("gemm_f64_kernel", "gemm_split_kernel", "dl_mfma_kernel")
    timed_ms = sum(v["ms"] for v in mix["stats"].values())
    shares = {k: {"ms": v["ms"], "share": v["ms"] / timed_ms, "launches": v["launches"],
                  "bound": "mfma/hbm" if k in mfma_classes else "latency (one workgroup)"}
              for k, v in sorted(mix["stats"].items(), key=lambda kv: -kv[1]["ms"])}
    dominant = next(iter(shares))
    name, st = max(((k, v) for k, v in mix["stats"].items() if k in mfma_classes),
                   key=lambda kv: kv[1]["ms"])
    tflops = st["flops"] / (st["ms"] * 1e-3) / 1e12
    if name == "gemm_split_kernel":
        bound, unit, peak = "hbm", "GB/s", HBM_PEAK_GBS
        achieved = st["bytes"] / (st["ms"] * 1e-3) / 1e9
    else:
        bound, unit, peak = "mfma", "TFLOP/s", FP64_PEAK_TFLOPS
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
        "data": "synthetic: Heisenberg state from 50 device simple-update steps (dtau 0.05) on "
                "seeded N(0,1) site tensors",
        "config": {"workload": f"CTMRG iteration, Heisenberg J=1, 2x2 unit cell, d=2, D={D}, "
                               f"chi={chi}, half-system rSVD projectors (q=chi+2, 2 power "
                               f"iterations), all four directional moves",
                   "D": D, "chi": chi, "unit_cell": "2x2", "precision": "mixed"},
        "fp64_iters_per_s": ref["its"],
        "speedup_vs_fp64": mix["its"] / ref["its"],
        "bond_rdm_measure_ms": 1e3 * ref["measure_s"],
        "energy_per_site": {"mixed": mix["energy"], "fp64": ref["energy"], "abs_diff": dE,
                            "rel_diff": dE / abs(ref["energy"]),
                            "fp64_rounding_spread": spread,
                            "fp64_other_tiles": ref2["energy"],
                            "iterations_before_measurement": args.warmup + 1 + args.steps + 1 +
                            args.energy_iters,
                            "window_mean": {"mixed": mix["energy_mean"], "fp64": ref["energy_mean"],
                                            "fp64_other_tiles": ref2["energy_mean"],
                                            "abs_diff": abs(mix["energy_mean"] - ref["energy_mean"]),
                                            "fp64_rounding_spread": abs(ref2["energy_mean"] -
                                                                        ref["energy_mean"]),
                                            "sd_per_sample": {"mixed": mix["energy_sd"],
                                                              "fp64": ref["energy_sd"]},
                                            "samples": mix["energy_samples"]},
                            "note": "Heisenberg J=1 bond operator: last sample on the final "
                                    f"environments after {args.energy_iters} untimed iterations "
                                    "beyond the timed ones, and window means over samples every "
                                    "5 of those iterations (the randomized projectors keep the "
                                    "environment fluctuating at the 1e-6 level in energy, as the "
                                    "FP64 run with other tiles shows); the FP64 HIP path is held "
                                    "to the FP64 oracle by the tests"},
        "nn_correlators": {"mixed": mix["correlators"], "fp64": ref["correlators"],
                           "max_abs_diff": max(abs(mix["correlators"][k] - ref["correlators"][k])
                                               for k in ref["correlators"]),
                           "window_mean_max_abs_diff": max(abs(mix["correlators_mean"][k] -
                                                               ref["correlators_mean"][k])
                                                           for k in ref["correlators"])},
        "truncation_error": {"mixed": mix["truncation_error"], "fp64": ref["truncation_error"],
                             "definition": "sum_{i>=cD} s_i^2 / sum_i s_i^2 over the q computed "
                                           "singular values of each projector pair, last "
                                           "iteration"},
        "roofline": {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
                     "frac": achieved / peak, "traffic": traffic, "kernel": name,
                     "share_of_timed_kernel_time": st["ms"] / timed_ms,
                     "algorithmic_bytes_per_launch": st["bytes"] / st["launches"],
                     "flops_per_launch": st["flops"] / st["launches"],
                     "fp64_equivalent_tflops": tflops,
                     "avg_launch_ms": st["ms"] / st["launches"]},
        "dominant_kernel_by_time": dominant,
        "kernel_time_shares": shares,
        "autotune": {"fp64": args.autotune, "mixed": args.autotune_mixed},
        "kernels": {"mixed": mix["stats"], "fp64": ref["stats"]},
    }
    if not args.no_sweep:
        out["sweep"] = run_sweep(args)
    if not args.no_als:
        out["als"] = run_als(args)
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(ref["warm_state"], D, chi, args.seed)
        log(f"[cpu] {out['cpu_baseline']}")
        sw = out.get("sweep", {}).get("cpu_sequential")
        if sw is not None and D == 6 and chi == 144:
            per_d = dict(sw["seconds_per_iteration_by_D"])
            per_d[6] = 1.0 / out["cpu_baseline"]["value"]
            total = sum(per_d[inst["D"]] for inst in out["sweep"]["instances"])
            sw["instance_iters_per_s"] = len(out["sweep"]["instances"]) / total
            sw["note"] = ("instances run one after another on the FP64 oracle: one timed "
                          "iteration per distinct D (after one that grows the boundary to chi), "
                          "D=6 from the headline baseline")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
